// .cfg keyword parser: sim_params::sim_params(const char*) (Sim/sim_params.cc:10-410)
// restated over the POD rmt3d_config.  Same keywords, same argument
// order, same defaults, same post-parse consolidation; an unknown keyword or
// a missing argument is an error (the reference calls exit(1)).
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <string>
#include "../../include/rmt3d.h"

extern "C" void rmt3d_default_config(rmt3d_config *c) {
    memset(c, 0, sizeof *c);
    for (int i = 0; i < 3; i++) { c->sys_size[i] = 96; c->prd[i] = 1; }
    const double bd[6] = {0, 1, 0, 1, 0, 1};
    for (int i = 0; i < 6; i++) { c->xyz_bounds[i] = bd[i]; c->wall_pos[i] = bd[i]; }
    c->wall_dist = 0;
    c->fmu = 0.002; c->frho = 1; c->fdt_pad = 0.5;
    c->dt = 1.; c->T = 1.;
    c->frames = 100; c->stop_short = -1; c->num_iters = 5;
    c->nlayers = 8; c->max_extrap_rs = 5;
    c->wt_n = 2.5; c->ex_visc_mult = 1.; c->ev_trans_mult = 1.;
    c->sdt_pad = 0.25; c->dt_ex_pad = 0.8; c->weight_fac = 0.25; c->gravity = 0;
    c->godunov = 1; c->impl = 0; c->has_exact = 0;
    c->sim_type = RMT3D_SIM_FTEST;
    c->dump_code = 1 | 2; c->ntracers = 1000; c->debug_flag = 0;
    c->n_obj = 0;
    c->chkpt_freq = 1; c->omp_num_thr = 1; c->debug_obj_id = -1; c->obj_body = -1;
    c->data_fmt = 0; c->out_flag = 1 | 2 | 4 | 128; c->output_dim = 2; c->output_ind = -1;
    c->current_time = 0;
}

namespace {
struct Tok {
    int ln;
    bool ok = true;
    const char *next() {
        char *t = strtok(nullptr, " \t\n");
        if (!t) { fprintf(stderr, "Not enough arguments at input line %d\n", ln); ok = false; return "0"; }
        return t;
    }
    int i() { return atoi(next()); }
    double d() { return atof(next()); }
};
}  // namespace

extern "C" int rmt3d_parse_cfg(const char *path, rmt3d_config *c, rmt3d_object *objs, int max_obj) {
    rmt3d_default_config(c);
    size_t len = strlen(path);
    if (len < 4 || strcmp(path + len - 4, ".cfg") != 0) {
        fprintf(stderr, "sim_params::sim_params(const char*): Filename must end in '.cfg'\n");
        return RMT3D_ERR_CFG;
    }
    // the .cfg root names the output directory, which is also where
    // checkpoints are sought unless recover_dirname says otherwise
    // (sim_params.cc:53-61)
    if (len - 4 + 4 >= sizeof c->dirname) {
        fprintf(stderr, "sim_params: configuration path too long\n");
        return RMT3D_ERR_CFG;
    }
    memcpy(c->dirname, path, len - 3);
    memcpy(c->dirname + len - 3, "out", 4);
    memcpy(c->recover_dirname, c->dirname, len + 1);
    FILE *f = fopen(path, "r");
    if (!f) return RMT3D_ERR_CFG;
    char buf[1024];
    int ln = 1;
    std::string sim_type = "ftest";
    int n_obj = 0;
    bool have_obj = false;
    const int NB = 4, NE = RMT3D_N_EXTRA_SPECS;
    int rc = 0;
    while (fgets(buf, sizeof buf, f)) {
        for (char *bp = buf; *bp; bp++)
            if (*bp == '#') { *bp = 0; break; }
        char *kw = strtok(buf, " \t\n");
        if (!kw) { ln++; continue; }
        Tok T;
        T.ln = ln;
        std::string k(kw);
        if (k == "suppress") c->suppress = T.i();
        else if (k == "display") c->display = T.i();
        else if (k == "omp_num_thr") c->omp_num_thr = T.i();
        else if (k == "debug_obj_id") c->debug_obj_id = T.i();
        else if (k == "out_flag") c->out_flag = T.i();
        else if (k == "obj_body") c->obj_body = T.i();
        else if (k == "data_fmt") c->data_fmt = T.i();
        else if (k == "chkpt_freq") c->chkpt_freq = T.i();
        else if (k == "implicit") c->impl = T.i();
        else if (k == "godunov") c->godunov = T.i();
        else if (k == "x_prd") c->prd[0] = T.i();
        else if (k == "y_prd") c->prd[1] = T.i();
        else if (k == "z_prd") c->prd[2] = T.i();
        else if (k == "has_exact") c->has_exact = T.i();
        else if (k == "walls") { for (int i = 0; i < 6; i++) c->walls[i] = T.i(); }
        else if (k == "wall_pos") { for (int i = 0; i < 6; i++) c->wall_pos[i] = T.d(); }
        else if (k == "wall_dist") c->wall_dist = T.d();
        else if (k == "recover_dirname") snprintf(c->recover_dirname, sizeof c->recover_dirname, "%s", T.next());
        else if (k == "chk_num") c->chk_num = T.i();
        else if (k == "chk_step") c->chk_step = T.i();
        else if (k == "sys_size") {
            for (int i = 0; i < 3; i++) { int t = T.i(); if (t > 0) c->sys_size[i] = t; }
        } else if (k == "frames") c->frames = T.i();
        else if (k == "stop_short") c->stop_short = T.i();
        else if (k == "debug_flag") c->debug_flag = T.i();
        else if (k == "output_dim") {
            c->output_dim = T.i();
            if (c->output_dim < 0 || c->output_dim > 2) c->output_dim = -1;
        } else if (k == "output_ind") {
            // compared with sys_size as parsed so far (sim_params.cc:145-147)
            c->output_ind = T.i();
            if (c->output_ind < 0 || c->output_dim < 0 || c->output_ind >= c->sys_size[c->output_dim])
                c->output_ind = -1;
        }
        else if (k == "dump_code") c->dump_code = (int)strtoul(T.next(), nullptr, 10);
        else if (k == "ntracers") c->ntracers = T.i();
        else if (k == "num_iters") c->num_iters = T.i();
        else if (k == "dt") c->dt = T.d();
        else if (k == "T") { double t = T.d(); if (t > 0) c->T = t; }
        else if (k == "current_time") { double t = T.d(); if (t > 0) c->current_time = t; }
        else if (k == "xyz_bounds") { for (int i = 0; i < 6; i++) c->xyz_bounds[i] = T.d(); }
        else if (k == "fmu") c->fmu = T.d();
        else if (k == "fdt_pad") c->fdt_pad = T.d();
        else if (k == "vel_prof_num") {
            c->vel_prof_num = T.i();
            if (c->vel_prof_num > RMT3D_MAX_VEL_PROF) { rc = RMT3D_ERR_CFG; break; }
        } else if (k == "n_obj") {
            n_obj = T.i();
            if (n_obj > 0 && !have_obj) {
                if (n_obj > max_obj) { fprintf(stderr, "too many objects for the caller's array\n"); rc = RMT3D_ERR_CFG; break; }
                have_obj = true;
                for (int i = 0; i < n_obj; i++) {
                    memset(&objs[i], 0, sizeof objs[i]);
                    objs[i].type = 0;
                    objs[i].srho = 1.; objs[i].shear_mod = 1.;
                    objs[i].extra[0] = 1.;
                }
            }
        } else if (k == "nlayers") c->nlayers = T.i();
        else if (k == "max_extrap_rs") { c->max_extrap_rs = T.i(); if (c->max_extrap_rs < 2) c->max_extrap_rs = 2; }
        else if (k == "wt_n") c->wt_n = T.d();
        else if (k == "ex_visc_mult") c->ex_visc_mult = T.d();
        else if (k == "ev_trans_mult") c->ev_trans_mult = T.d();
        else if (k == "sdt_pad") c->sdt_pad = T.d();
        else if (k == "dt_ex_pad") c->dt_ex_pad = T.d();
        else if (k == "weight_fac") c->weight_fac = T.d();
        else if (k == "gravity") c->gravity = T.d();
        else if (k.size() == 5 && k.compare(0, 3, "bc_") == 0 && (k[3] == 'x' || k[3] == 'y' || k[3] == 'z') &&
                 (k[4] == 'l' || k[4] == 'h')) {
            int face = 2 * (k[3] - 'x') + (k[4] == 'h');
            int ty[3];
            double va[3];
            for (int i = 0; i < 3; i++) ty[i] = T.i();
            for (int i = 0; i < 3; i++) va[i] = T.d();
            for (int i = 0; i < 3; i++) { c->bct[i][face] = ty[i]; c->bcv[i][face] = va[i]; }
        } else if (k == "basic_specs") {
            int idx = T.i();
            if (idx < n_obj) for (int i = 0; i < NB; i++) objs[idx].basic[i] = T.d();
        } else if (k == "vel_profile") {
            int idx = T.i();
            if (idx < c->vel_prof_num) {
                int vt = T.i();
                c->vel_profile[2 * idx] = (double)vt;
                if (vt != 0) c->vel_profile[2 * idx + 1] = T.d();
            }
        } else if (k == "srho") {
            int idx = T.i();
            if (idx < n_obj) objs[idx].srho = T.d();
        } else if (k == "shear_mod") {
            int idx = T.i();
            if (idx < n_obj) objs[idx].shear_mod = T.d();
        } else if (k == "object_list") {
            int idx = T.i();
            if (idx < n_obj) objs[idx].type = T.i();
        } else if (k == "extra_specs") {
            int idx = T.i();
            if (idx < n_obj) for (int i = 0; i < NE; i++) objs[idx].extra[i] = T.d();
        } else if (k == "sim_type") {
            sim_type = T.next();
        } else {
            fprintf(stderr, "sim_params::sim_params(const char*): Unrecognized keyword '%s' at line %d of file %s\n",
                    kw, ln, path);
            rc = RMT3D_ERR_CFG;
            break;
        }
        if (!T.ok) { rc = RMT3D_ERR_CFG; break; }
        ln++;
    }
    fclose(f);
    if (rc) return rc;
    // consolidation (sim_params.cc:326-372)
    if (c->output_ind == -1 && c->output_dim >= 0) c->output_ind = c->sys_size[c->output_dim] / 2;
    for (int d = 0; d < 3; d++) {
        int s = c->bct[0][2 * d] + c->bct[1][2 * d] + c->bct[2][2 * d];
        if (s >= 0 && s < 3) {
            for (int comp = 0; comp < 3; comp++) { c->bct[comp][2 * d] = 0; c->bct[comp][2 * d + 1] = 0; }
            c->prd[d] = 1;
        } else {
            c->prd[d] = 0;
        }
    }
    for (int d = 0; d < 3; d++)
        if (!c->prd[d]) c->walls[2 * d] = c->walls[2 * d + 1] = 1;
    double dxs[3];
    for (int d = 0; d < 3; d++) dxs[d] = (c->xyz_bounds[2 * d + 1] - c->xyz_bounds[2 * d]) / c->sys_size[d];
    double mdh = dxs[0] < dxs[1] ? dxs[0] : dxs[1];
    mdh = mdh < dxs[2] ? mdh : dxs[2];
    c->wall_dist = mdh > c->wall_dist ? mdh : c->wall_dist;
    // a checkpoint is used only when its recover.cfg can be opened
    // (sim_params.cc:374-398)
    c->chk_dirname[0] = 0;
    if (c->chk_num > 0) {
        char probe[600];
        snprintf(c->chk_dirname, sizeof c->chk_dirname, "%s/chk.%05d", c->recover_dirname, c->chk_num);
        snprintf(probe, sizeof probe, "%s/recover.cfg", c->chk_dirname);
        FILE *fh = fopen(probe, "rb");
        if (!fh) {
            printf("sim_params::Cannot find checkpt directory %s. Reset to none.\n", c->chk_dirname);
            c->chk_num = 0;
            c->chk_step = 0;
            c->chk_dirname[0] = 0;
        } else {
            fclose(fh);
        }
    }
    if (sim_type == "ftest") { c->sim_type = RMT3D_SIM_FTEST; n_obj = 0; }
    else if (sim_type == "stest") { c->sim_type = RMT3D_SIM_STEST; n_obj = 0; }
    else if (sim_type == "objects") c->sim_type = RMT3D_SIM_OBJECTS;
    else if (sim_type == "object_mills") c->sim_type = RMT3D_SIM_MILLS;
    else {
        fprintf(stderr, "Simulation types can be:\n(1) ftest\n(2) stest\n(3) objects\nProvided: |%s|\n", sim_type.c_str());
        return RMT3D_ERR_CFG;
    }
    c->n_obj = n_obj;
    return 0;
}
