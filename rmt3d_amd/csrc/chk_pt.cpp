// Host side of the checkpoint and slice formats: the files of
// fluid_3d::write_chk_pt / initialize_from_chk_point (fluid_3d_chk_pt.cc:
// 3-163), the recover.cfg of sim_params::write_params (Sim/sim_params.cc:
// 489-606) and the three slice formats of fluid_3d::write_slice
// (fluid_3d_io.cc:3-90).  The field values themselves are gathered and
// scattered on the device (k_io.hip); this file only moves bytes.
#include <stdio.h>
#include <string.h>
#include <sys/stat.h>
#include <string>
#include <vector>
#include "launch.h"
#include "../../include/rmt3d.h"

namespace rmt {

// write_params::default_names[0,1,2,7,9,10,11,12] (IO/write_params.cc:3)
const char *const CHK_NAMES[8] = {"vx", "vy", "vz", "p", "X", "Y", "Z", "OID"};

int chk_write_files(const char *dir, const double *img, int o, size_t L, std::string &err) {
    for (int f = 0; f < 8; f++) {
        std::string path = std::string(dir) + "/" + CHK_NAMES[f];
        FILE *fh = fopen(path.c_str(), "wb");
        if (!fh) {
            err = "cannot open " + path + " for writing";
            return RMT3D_ERR_ARG;
        }
        size_t n = (size_t)o * L;
        size_t w = fwrite(img + (size_t)f * n, sizeof(double), n, fh);
        int c = fclose(fh);
        if (w != n || c != 0) {
            err = "short write to " + path;
            return RMT3D_ERR_ARG;
        }
    }
    return 0;
}

static bool exists(const std::string &p) {
    struct stat st;
    return stat(p.c_str(), &st) == 0;
}

int chk_read_files(const char *dir, double *img, int o, size_t L, std::string &err) {
    // one slice buffer for every field and slice, zeroed once: a short read
    // leaves the previous slice's values in place, as in the reference
    std::vector<double> buf(L, 0.0);
    int status = 0;
    for (int f = 0; f < 8; f++) {
        std::string path = std::string(dir) + "/" + CHK_NAMES[f];
        if (!exists(path) && exists(path + ".xz")) {
            err = path + ".xz is compressed: decompress it (xz -d) first";
            return RMT3D_ERR_ARG;
        }
        FILE *fh = fopen(path.c_str(), "rb");
        if (!fh) {
            printf("fluid_3d::initialize_from_chk_point: File %s cannot be read!\n", path.c_str());
            return 1;
        }
        for (int k = 0; k < o; k++) {
            size_t bs = fread(buf.data(), sizeof(double), L, fh);
            if (bs == 0) {
                printf("fluid_3d::initialize_from_chk_point: Checkpoint file is empty!\n");
                status = 2;
            }
            memcpy(img + ((size_t)f * o + k) * L, buf.data(), sizeof(double) * L);
        }
        fclose(fh);
    }
    return status;
}

int slice_write_file(const char *filename, int format, const double *v, int d1, int d2, const double *a1,
                     const double *a2, int m) {
    FILE *fh = fopen(filename, "w");
    if (!fh) return RMT3D_ERR_ARG;
    if (format == 0) {
        for (int j = 0; j < d2; j++) {
            for (int i = 0; i < d1; i++) fprintf(fh, "%17.14g ", v[i + (size_t)j * d1]);
            fputc('\n', fh);
        }
    } else if (format == 1) {
        for (int j = 0; j < d2; j++) {
            for (int i = 0; i < d1; i++) fprintf(fh, "%17.14f %17.14f %17.14f\n", a1[i], a2[j], v[i + (size_t)j * d1]);
            fputc('\n', fh);
        }
    } else {
        // rows of d1+1 floats: the header row holds m and the first
        // coordinates, each following row its second coordinate and values
        std::vector<float> row(d1 + 1);
        row[0] = (float)m;
        for (int i = 0; i < d1; i++) row[i + 1] = (float)a1[i];
        fwrite(row.data(), sizeof(float), d1 + 1, fh);
        for (int j = 0; j < d2; j++) {
            row[0] = (float)a2[j];
            for (int i = 0; i < d1; i++) row[i + 1] = (float)v[i + (size_t)j * d1];
            fwrite(row.data(), sizeof(float), d1 + 1, fh);
        }
    }
    return fclose(fh) == 0 ? 0 : RMT3D_ERR_ARG;
}

}  // namespace rmt

// sim_type keyword of an RMT3D_SIM_* value (Execs/run_sim.cc:13-32)
static const char *sim_type_name(int t) {
    switch (t) {
        case RMT3D_SIM_STEST: return "stest";
        case RMT3D_SIM_OBJECTS: return "objects";
        case RMT3D_SIM_MILLS: return "object_mills";
        default: return "ftest";
    }
}

extern "C" int rmt3d_write_params(const rmt3d_config *c, const rmt3d_object *objs, const char *dir) {
    if (!c || !dir || (c->n_obj > 0 && !objs)) return RMT3D_ERR_ARG;
    std::string path = std::string(dir) + "/recover.cfg";
    FILE *fh = fopen(path.c_str(), "w");
    if (!fh) return RMT3D_ERR_ARG;
    // The keyword table of the reference's recover.cfg, value for value:
    // it is read back by the same .cfg parser, so the layout is the format.
    fprintf(fh, "#Keyword                  Input(default)           Type\n");
    fprintf(fh, "#System parameters\n");
    fprintf(fh, "suppress                  %d#(0)                   bool\n", c->suppress != 0);
    fprintf(fh, "display                   %d#(0)                   bool\n", c->display != 0);
    fprintf(fh, "implicit                  %d#(0)                   bool\n", c->impl != 0);
    fprintf(fh, "godunov                   %d#(0)                   bool\n", c->godunov != 0);
    fprintf(fh, "x_prd                     %d#(0)                   bool\n", c->prd[0] != 0);
    fprintf(fh, "y_prd                     %d#(0)                   bool\n", c->prd[1] != 0);
    fprintf(fh, "z_prd                     %d#(0)                   bool\n", c->prd[2] != 0);
    fprintf(fh, "has_exact                 %d#(0)                   bool\n", c->has_exact != 0);
    fprintf(fh, "recover_dirname           %s#                      string\n", c->dirname);
    fprintf(fh, "chk_num                   %d#(0)                   int\n", c->chk_num);
    fprintf(fh, "chk_step                  %d#(0)                   int\n", c->chk_step);
    fprintf(fh, "sys_size                  %d %d %d#({96,96,96})    int\n", c->sys_size[0], c->sys_size[1],
            c->sys_size[2]);
    fprintf(fh, "frames                    %d#(96)                  int\n", c->frames);
    fprintf(fh, "stop_short                %d#(96)                  int\n", c->stop_short);
    fprintf(fh, "chkpt_freq                %d#(1)                   int\n", c->chkpt_freq);
    fprintf(fh, "omp_num_thr               %d#(1)                   int\n", c->omp_num_thr);
    fprintf(fh, "debug_flag                %d#(0)                   int\n", c->debug_flag);
    fprintf(fh, "debug_obj_id              %d#(-1)                   int\n", c->debug_obj_id);
    fprintf(fh, "obj_body                  %d#(-1)                  int\n", c->obj_body);
    fprintf(fh, "data_fmt                  %d#(0)                   int\n", c->data_fmt);
    fprintf(fh, "out_flag                  %d#(135)                 int\n", c->out_flag);
    fprintf(fh, "output_dim                %d#0/1/2(2)              int\n", c->output_dim);
    fprintf(fh, "output_ind                %d#(48)                  int\n", c->output_ind);
    fprintf(fh, "dump_code                 %u#(3)                   unsigned int\n", (unsigned)c->dump_code);
    fprintf(fh, "ntracers                  %d#(1000)                int\n", c->ntracers);
    fprintf(fh, "num_iters                 %d#(5)                   int\n", c->num_iters);
    fprintf(fh, "dt                        %g#(1.)                  double\n", c->dt);
    fprintf(fh, "T                         %g#(1.)                  double\n", c->T);
    fprintf(fh, "current_time              %g#(0.)                  double\n", c->current_time);
    fprintf(fh, "xyz_bounds                %g %g %g %g %g %g\n", c->xyz_bounds[0], c->xyz_bounds[1],
            c->xyz_bounds[2], c->xyz_bounds[3], c->xyz_bounds[4], c->xyz_bounds[5]);
    fprintf(fh, "                          #(0 1 0 1 0 1)         doubles\n");
    fprintf(fh, "# SIMULATION TYPE CONFIGURATIONS\n");
    fprintf(fh, "sim_type                  %s#(\"ftest\")           string\n", sim_type_name(c->sim_type));
    fprintf(fh, "# FLUID PROPERTIES\n");
    fprintf(fh, "fmu                       %g#(0.002)               double\n", c->fmu);
    fprintf(fh, "fdt_pad                   %g#(0.5)                 double\n", c->fdt_pad);
    fprintf(fh, "# SOLID PROPERTIES\n");
    fprintf(fh, "n_obj                     %d#(0)                   int\n", c->n_obj);
    fprintf(fh, "nlayers                   %d#(8)                   int\n", c->nlayers);
    fprintf(fh, "wt_n                      %g#(2.5)                 double\n", c->wt_n);
    fprintf(fh, "ex_visc_mult              %g#(1.)                  double\n", c->ex_visc_mult);
    fprintf(fh, "ev_trans_mult             %g#(1.)                  double\n", c->ev_trans_mult);
    fprintf(fh, "sdt_pad                   %g#(0.25)                double\n", c->sdt_pad);
    fprintf(fh, "dt_ex_pad                 %g#(0.8)                 double\n", c->dt_ex_pad);
    fprintf(fh, "gravity                   %g#(10.)                 double\n", c->gravity);
    fprintf(fh, "# SOLIDS CONFIGURATIONS\n");
    int n = c->n_obj;
    for (int i = 0; i < n; i++) fprintf(fh, "object_list               %d %d\n", i, objs[i].type);
    for (int i = 0; i < n; i++) {
        fprintf(fh, "basic_specs               %d ", i);
        for (int j = 0; j < 4; j++) fprintf(fh, "%g ", objs[i].basic[j]);
        fprintf(fh, "\n");
    }
    for (int i = 0; i < n; i++) {
        fprintf(fh, "extra_specs               %d ", i);
        for (int j = 0; j < RMT3D_N_EXTRA_SPECS; j++) fprintf(fh, "%g ", objs[i].extra[j]);
        fprintf(fh, "\n");
    }
    for (int i = 0; i < n; i++) fprintf(fh, "srho                      %d %g\n", i, objs[i].srho);
    for (int i = 0; i < n; i++) fprintf(fh, "shear_mod                 %d %g\n", i, objs[i].shear_mod);
    fprintf(fh, "vel_prof_num              %d#(1)                   int\n", c->vel_prof_num);
    for (int i = 0; i < c->vel_prof_num && i < RMT3D_MAX_VEL_PROF; i++)
        fprintf(fh, "vel_profile               %d %d %g\n", i, (int)c->vel_profile[2 * i], c->vel_profile[2 * i + 1]);
    fprintf(fh, "# DOMAIN BOUNDARY CONDITIONS\n");
    static const char *face[6] = {"bc_xl", "bc_xh", "bc_yl", "bc_yh", "bc_zl", "bc_zh"};
    for (int f = 0; f < 6; f++) {
        fprintf(fh, "%s                     %d %d %d %g %g %g\n", face[f], c->bct[0][f], c->bct[1][f], c->bct[2][f],
                c->bcv[0][f], c->bcv[1][f], c->bcv[2][f]);
    }
    fprintf(fh, "# WALL BOUNDARIES\n");
    fprintf(fh, "walls                     %d %d %d %d %d %d\n", c->walls[0] != 0, c->walls[1] != 0, c->walls[2] != 0,
            c->walls[3] != 0, c->walls[4] != 0, c->walls[5] != 0);
    fprintf(fh, "wall_pos                  %g %g %g %g %g %g\n", c->wall_pos[0], c->wall_pos[1], c->wall_pos[2],
            c->wall_pos[3], c->wall_pos[4], c->wall_pos[5]);
    fprintf(fh, "wall_dist                 %g\n", c->wall_dist);
    return fclose(fh) == 0 ? 0 : RMT3D_ERR_ARG;
}
