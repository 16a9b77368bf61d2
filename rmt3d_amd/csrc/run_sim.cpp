// Command-line driver over the C ABI, following the reference's run_sim
// sequence (Execs/run_sim.cc:39-165): read a .cfg, create the simulation,
// initialize, set up the stepper (dt clamp, step count, init_iter warm-up),
// then step frame by frame, printing per-frame diagnostics and the total
// wall time.  Usage:
//   run_sim <file.cfg> [--mixed] [--steps N] [--quiet]
// --steps N overrides the step count of the stepper (stop_short / T / dt).
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "rmt3d.h"

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
    if (steps_override >= 0) total = steps_override;
    int frames = cfg.frames > 0 ? cfg.frames : 1;
    int per_frame = total / frames > 0 ? total / frames : 1;
    long long cells = (long long)cfg.sys_size[0] * cfg.sys_size[1] * cfg.sys_size[2];
    printf("# %s: %dx%dx%d, %d object(s), %s, dt %.6g, %d steps\n", argv[1], cfg.sys_size[0], cfg.sys_size[1],
           cfg.sys_size[2], cfg.n_obj, precision == RMT3D_PREC_MIXED ? "mixed" : "fp64", dt, total);
    rmt3d_device_synchronize(ctx);
    auto t1 = std::chrono::steady_clock::now();
    int status = 0, done = 0;
    for (int s = 0; s < total; s++) {
        status = rmt3d_step(ctx, 0);
        if (status) break;
        done++;
        if (!quiet && (done % per_frame == 0 || done == total)) {
            double divu = 0, nmac = 0, nfem = 0;
            rmt3d_diag(ctx, 0, &divu);
            rmt3d_diag(ctx, 1, &nmac);
            rmt3d_diag(ctx, 2, &nfem);
            printf("step %6d  time %.6f  div_u %.3e  V-cycles MAC %2.0f FEM %2.0f\n", done, rmt3d_get_time(ctx), divu,
                   nmac, nfem);
            fflush(stdout);
        }
    }
    rmt3d_device_synchronize(ctx);
    auto t2 = std::chrono::steady_clock::now();
    double init_s = std::chrono::duration<double>(t1 - t0).count();
    double run_s = std::chrono::duration<double>(t2 - t1).count();
    if (status) printf("# stopped at step %d with code %d (%s)\n", done + 1, status, rmt3d_last_error(ctx));
    printf("# initialization %.3f s, %d steps %.3f s, %.3f ms/step, %.4g cell-steps/s\n", init_s, done, run_s,
           done ? 1e3 * run_s / done : 0.0, done ? cells * done / run_s : 0.0);
    rmt3d_destroy(ctx);
    return status < 0 ? 1 : 0;
}
