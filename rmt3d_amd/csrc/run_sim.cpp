// Command-line driver over the C ABI, following the reference's run_sim
// sequence (Execs/run_sim.cc:39-165) and the stepper's frame loop
// (Sim/stepper.hh:211-376, 580-610): read a .cfg, create the simulation,
// initialize (from a checkpoint when chk_num names one), set up the stepper
// (dt clamp, frame step count, init_iter warm-up), then step frame by frame.
// After each frame (and once before the first) it writes what dump_code
// asks for: tracers (bit 1, tr.<frame> binary and tr_st.<frame> text),
// contours (bit 2, ctr.<frame>), slices (bit 4, one per out_flag bit, write_params naming
// "<dirname>/<field>.<frame>"), checkpoints (bit 8, every chkpt_freq
// frames, "<dirname>/chk.<frame>") and the macro.dat / centroid.dat records
// (bit 32).  Usage:
//   run_sim <file.cfg> [--mixed] [--steps N] [--quiet]
// --steps N stops after N steps (as stop_short does).
#include <sys/stat.h>
#include <cfloat>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "rmt3d.h"

// write_params::default_names (IO/write_params.cc:3)
static const char *const FIELD_NAMES[22] = {"vx", "vy", "vz", "wx",  "wy",     "wz",   "w",    "p",
                                            "phi", "X", "Y",  "Z",   "OID",    "LID",  "pr_phi", "pr_X",
                                            "pr_Y", "pr_Z", "div_u", "div_eu", "rho", "J"};

// stepper::write_files (stepper.hh:580-610)
static int write_files(rmt3d_ctx *ctx, const rmt3d_config &cfg, int snum, int fnum) {
    char buf[1024];
    if ((cfg.dump_code & 1) && cfg.ntracers > 0) {
        snprintf(buf, sizeof buf, "%s/tr.%05d", cfg.dirname, fnum);
        int rc = rmt3d_write_tracers(ctx, buf, 1);
        if (rc) return rc;
        snprintf(buf, sizeof buf, "%s/tr_st.%05d", cfg.dirname, fnum);
        if ((rc = rmt3d_write_tracers(ctx, buf, 0))) return rc;
    }
    if (cfg.dump_code & 2) {
        snprintf(buf, sizeof buf, "%s/ctr.%05d", cfg.dirname, fnum);
        int rc = rmt3d_write_contours(ctx, buf);
        if (rc) return rc;
    }
    if (cfg.dump_code & 4) {
        for (int t = 0; t < 22; t++) {
            if (!((1 << t) & cfg.out_flag)) continue;
            snprintf(buf, sizeof buf, "%s/%s.%05d", cfg.dirname, FIELD_NAMES[t], fnum);
            int rc = rmt3d_write_slice(ctx, cfg.output_dim, cfg.output_ind, t, cfg.obj_body, cfg.data_fmt, buf);
            if (rc) return rc;
        }
    }
    if ((cfg.dump_code & 8) && cfg.chkpt_freq > 0 && fnum % cfg.chkpt_freq == 0) {
        snprintf(buf, sizeof buf, "%s/chk.%05d", cfg.dirname, fnum);
        mkdir(buf, S_IRWXU | S_IRWXG | S_IROTH | S_IXOTH);
        int rc = rmt3d_write_checkpoint(ctx, snum, fnum, buf);
        if (rc) return rc;
    }
    return 0;
}

static bool file_exists(const char *path) {
    FILE *f = fopen(path, "r");
    if (!f) return false;
    fclose(f);
    return true;
}

// stepper::collect_macro_data (stepper.hh:378-475): one macro.dat record
// per call (the reference ends it without a newline), and with objects one
// centroid.dat line; the headers are rewritten at the first call of a run
static int collect_macro_data(rmt3d_ctx *ctx, const rmt3d_config &cfg, int frame, long long tot_step,
                              long long init_step) {
    if (!(cfg.dump_code & 32)) return 0;
    char fn[1024];
    snprintf(fn, sizeof fn, "%s/macro.dat", cfg.dirname);
    if (!file_exists(fn) || tot_step == init_step) {
        FILE *fh = fopen(fn, "w");
        if (!fh) return RMT3D_ERR_ARG;
        fprintf(fh,
                "%10s %6s %6s %10s %6s %10s %6s %10s %10s %10s %10s %10s %10s %10s %10s %10s %10s %10s %16s %16s "
                "%10s\n",
                "#Time", "frame", "steps", "div_u_max", "rank", "div_u_min", "rank", "divu_L2", "tot_mmt_x",
                "tot_mmt_y", "tot_mmt_z", "tot_energy", "kinetic", "elastic", "potential", "power", "max_dev_detF",
                "dev_detF_L2", "sing. obj. cent.", "volume", "change in u");
        fclose(fh);
    }
    rmt3d_macro m;
    int rc = rmt3d_macro_data(ctx, &m);
    if (rc) return rc;
    const int digs = DBL_DIG;
    double t = rmt3d_get_time(ctx);
    FILE *fh = fopen(fn, "a");
    if (!fh) return RMT3D_ERR_ARG;
    fprintf(fh, "%.*e %06d %06lld %.*e %06d %.*e %06d ", digs, t, frame, tot_step, digs, m.divu_max, 0, digs,
            m.divu_min, 0);
    fprintf(fh, "%.*e ", digs, m.divu_l2);
    fprintf(fh, "%.*e %.*e %.*e %.*e ", digs, m.momentum[0], digs, m.momentum[1], digs, m.momentum[2], digs,
            m.energy);
    fprintf(fh, "%.*e %.*e %.*e %.*e ", digs, m.kinetic, digs, m.elastic, digs, m.potential, digs, m.power);
    fprintf(fh, "%.*e %.*e ", digs, m.max_dev_detF, digs, m.dev_detF_l2);
    fclose(fh);
    if (cfg.n_obj > 0 && cfg.sim_type != RMT3D_SIM_FTEST) {
        std::vector<double> c(6 * (size_t)cfg.n_obj);
        if ((rc = rmt3d_centroids(ctx, c.data()))) return rc;
        snprintf(fn, sizeof fn, "%s/centroid.dat", cfg.dirname);
        if (!file_exists(fn) || tot_step == init_step) {
            fh = fopen(fn, "w");
            if (!fh) return RMT3D_ERR_ARG;
            fprintf(fh, "%10s %6s %6s %6s %6s %6s %6s %6s %6s->\n", "#Time", "frame", "steps", "x", "y", "z", "vx",
                    "vy", "vz");
            fclose(fh);
        }
        fh = fopen(fn, "a");
        if (!fh) return RMT3D_ERR_ARG;
        fprintf(fh, "%.*e %d %lld ", digs, t, frame, tot_step);
        for (int i = 0; i < cfg.n_obj; i++) {
            fprintf(fh, "%.*e %.*e %.*e ", digs, c[6 * i], digs, c[6 * i + 1], digs, c[6 * i + 2]);
            fprintf(fh, "%.*e %.*e %.*e ", digs, c[6 * i + 3], digs, c[6 * i + 4], digs, c[6 * i + 5]);
        }
        fprintf(fh, "\n");
        fclose(fh);
    }
    return 0;
}

int main(int argc, char **argv) {
    if (argc < 2) {
        fprintf(stderr, "usage: %s <file.cfg> [--mixed] [--steps N] [--quiet]\n", argv[0]);
        return 1;
    }
    int precision = RMT3D_PREC_FP64, steps_override = -1, quiet = 0;
    for (int a = 2; a < argc; a++) {
        if (!strcmp(argv[a], "--mixed")) precision = RMT3D_PREC_MIXED;
        else if (!strcmp(argv[a], "--quiet")) quiet = 1;
        else if (!strcmp(argv[a], "--steps") && a + 1 < argc) steps_override = atoi(argv[++a]);
        else {
            fprintf(stderr, "unknown argument %s\n", argv[a]);
            return 1;
        }
    }
    rmt3d_config cfg;
    std::vector<rmt3d_object> objs(1024);
    int rc = rmt3d_parse_cfg(argv[1], &cfg, objs.data(), (int)objs.size());
    if (rc) {
        fprintf(stderr, "run_sim: cannot read %s (%d)\n", argv[1], rc);
        return 1;
    }
    rmt3d_ctx *ctx = nullptr;
    rc = rmt3d_create(&cfg, objs.data(), precision, &ctx);
    if (rc) {
        fprintf(stderr, "run_sim: create failed (%d): %s\n", rc, rmt3d_last_error(nullptr));
        return 1;
    }
    auto t0 = std::chrono::steady_clock::now();
    int init_err = rmt3d_initialize(ctx);
    if (init_err < 0) {
        fprintf(stderr, "run_sim: initialize failed: %s\n", rmt3d_last_error(ctx));
        rmt3d_destroy(ctx);
        return 1;
    }
    double dt = 0;
    int total = 0;
    rc = rmt3d_prepare_run(ctx, init_err, &dt, &total);
    if (rc) {
        fprintf(stderr, "run_sim: prepare_run failed (%d): %s\n", rc, rmt3d_last_error(ctx));
        rmt3d_destroy(ctx);
        return 1;
    }
    // the stepper's frame bookkeeping (stepper.hh:211-235): steps per frame,
    // the first frame and step of a resumed run, and the stop_short bound
    // counted from that step
    int init_frame = init_err > 0 ? 0 : cfg.chk_num, init_step = cfg.chk_step;
    double df = cfg.T / (cfg.frames - 1);
    int steps = (int)llround(df / dt);
    if (steps < 1) steps = 1;
    long long stop = (cfg.stop_short <= 0) ? (long long)(cfg.frames - 1) * steps + 1 : cfg.stop_short + init_step;
    if (steps_override >= 0) stop = (long long)init_step + steps_override;
    long long cells = (long long)cfg.sys_size[0] * cfg.sys_size[1] * cfg.sys_size[2];
    printf("# %s: %dx%dx%d, %d object(s), %s, dt %.6g, %d steps per frame, %d step(s) to run\n", argv[1],
           cfg.sys_size[0], cfg.sys_size[1], cfg.sys_size[2], cfg.n_obj,
           precision == RMT3D_PREC_MIXED ? "mixed" : "fp64", dt, steps, total);
    if (cfg.dump_code & (1 | 2 | 4 | 8 | 32)) mkdir(cfg.dirname, S_IRWXU | S_IRWXG | S_IROTH | S_IXOTH);
    // the first macro record is taken with the frame counter still at 0
    // (stepper.hh:216, 341-344)
    if ((rc = write_files(ctx, cfg, init_step, init_frame)) ||
        (rc = collect_macro_data(ctx, cfg, 0, init_step, init_step))) {
        fprintf(stderr, "run_sim: output failed: %s\n", rmt3d_last_error(ctx));
        rmt3d_destroy(ctx);
        return 1;
    }
    rmt3d_device_synchronize(ctx);
    auto t1 = std::chrono::steady_clock::now();
    int status = 0, done = 0;
    long long tot_step = init_step;
    for (int frame = init_frame + 1; frame <= cfg.frames - 1 && tot_step < stop && status == 0; frame++) {
        for (int s = 0; s < steps && tot_step < stop; s++) {
            tot_step++;
            status = rmt3d_step(ctx, 0);
            if (status) break;
            done++;
        }
        if (status) break;
        if (!quiet) {
            double divu = 0, nmac = 0, nfem = 0;
            rmt3d_diag(ctx, 0, &divu);
            rmt3d_diag(ctx, 1, &nmac);
            rmt3d_diag(ctx, 2, &nfem);
            printf("frame %5d step %6lld  time %.6f  div_u %.3e  V-cycles MAC %2.0f FEM %2.0f\n", frame, tot_step,
                   rmt3d_get_time(ctx), divu, nmac, nfem);
            fflush(stdout);
        }
        if ((rc = write_files(ctx, cfg, (int)tot_step, frame)) ||
            (rc = collect_macro_data(ctx, cfg, frame, tot_step, init_step))) {
            fprintf(stderr, "run_sim: output failed: %s\n", rmt3d_last_error(ctx));
            status = rc;
            break;
        }
    }
    rmt3d_device_synchronize(ctx);
    auto t2 = std::chrono::steady_clock::now();
    double init_s = std::chrono::duration<double>(t1 - t0).count();
    double run_s = std::chrono::duration<double>(t2 - t1).count();
    if (status) printf("# stopped at step %lld with code %d (%s)\n", tot_step, status, rmt3d_last_error(ctx));
    printf("# initialization %.3f s, %d steps %.3f s, %.3f ms/step, %.4g cell-steps/s\n", init_s, done, run_s,
           done ? 1e3 * run_s / done : 0.0, done ? cells * done / run_s : 0.0);
    rmt3d_destroy(ctx);
    return status < 0 ? 1 : 0;
}
