// fluid_3d_gpu.hh -- the p_class adaptor a maintainer of RMT3D adds to run
// its stepper (Sim/stepper.hh) and run_sim (Execs/run_sim.cc) on librmt3d.
//
// Include it after the reference's own headers (geometry.hh, sim_params.hh,
// sim_manager.hh, write_params.hh, custom_structs.hh, common.hh), which
// supply geometry, sim_params, sim_manager, write_params, double_int and
// p_fatal_error.  It carries every member and method the reference's callers
// use:
//   stepper(sim_params&, p_class&), link, step, run      stepper.hh:211-376
//     set_dt, step_forward, rank, out, time, grid->cart
//   write_files                                           stepper.hh:579-610
//     tr->print(name[, binary]), output_contours, write_slice, write_chk_pt
//   dump_all                                              stepper.hh:612-632
//     o, write_slice
//   collect_macro_data                                    stepper.hh:378-475
//     div_u(double_int&, double_int&), total_momentum, total_energy,
//     total_dev_in_detF, compute_v_centroid, mgmt->n_obj, mgmt->avg_*
//   run_sim main                                          run_sim.cc:94-165
//     the constructor, initialize, public dt, init_iter, time, error,
//     watch.report, expper.watch.report, free
// The DEBUG-only run_diagnostics (stepper.hh:477-577) needs stress and
// curvature maxima that the library does not compute; it is not supported.
#ifndef FLUID_3D_GPU_HH
#define FLUID_3D_GPU_HH

#include <stdio.h>
#include <stdlib.h>
#include <vector>

#include "rmt3d.h"

class fluid_3d_gpu {
  public:
    // ---- members the stepper and run_sim read ----
    int rank = 0;                  // one GPU, one rank
    geometry *grid;                // stepper::run: MPI_Barrier(pr.grid->cart)
    sim_manager *mgmt;             // collect_macro_data: n_obj, avg_x .. avg_velz
    int o;                         // cells along z: dump_all's slice loop
    double time = 0;               // simulated time after the last step
    double dt = 0;                 // run_sim sets it before init_iter
    char *out = nullptr;           // status text the progress bar prints
    rmt3d_ctx *ctx = nullptr;

    // pr.tr->print(filename, binary): tracers::print (tracers.cc:110-139)
    struct tracer_shim {
        rmt3d_ctx *ctx = nullptr;
        void print(const char *fname, bool binary = true) {
            if (rmt3d_write_tracers(ctx, fname, binary ? 1 : 0)) p_fatal_error(rmt3d_last_error(ctx), 1);
        }
    };
    tracer_shim trs;
    tracer_shim *tr = &trs;

    // run_sim: f3d.watch.report() and f3d.expper.watch.report() print the
    // timer buckets (fluid_3d.cc:59); here the device's per-stage times
    struct stage_watch {
        rmt3d_ctx *ctx = nullptr;
        void report() {
            static const char *names[6] = {"Extrapolation", "Pressure Poisson", "Communication", "Stresses", "MAC",
                                           "Advection"};
            double ms[6] = {0, 0, 0, 0, 0, 0};
            if (rmt3d_stage_times(ctx, ms, 6) != 0) return;
            for (int i = 0; i < 6; i++) printf("# %-18s %12.6g s\n", names[i], ms[i] * 1e-3);
        }
    };
    stage_watch watch;
    struct { stage_watch watch; } expper;

    fluid_3d_gpu(geometry *gm, sim_manager *mgmt_, sim_params *spars, const char *cfg_path,
                 int precision = RMT3D_PREC_FP64)
        : grid(gm), mgmt(mgmt_), o(spars->sys_size[2]) {
        rmt3d_config cfg;
        std::vector<rmt3d_object> objs(spars->n_obj > 0 ? spars->n_obj : 1);
        // the library reads the same .cfg file the reference parsed
        if (rmt3d_parse_cfg(cfg_path, &cfg, objs.data(), (int)objs.size()) != 0)
            p_fatal_error("fluid_3d_gpu: rmt3d_parse_cfg failed", 1);
        if (rmt3d_create(&cfg, objs.data(), precision, &ctx) != 0)
            p_fatal_error("fluid_3d_gpu: rmt3d_create failed", 1);
        trs.ctx = ctx;
        watch.ctx = ctx;
        expper.watch.ctx = ctx;
        rmt3d_profile(ctx, 1);       // stage buckets for watch.report()
        outbuf[0] = '\0';
        out = outbuf;
    }
    ~fluid_3d_gpu() { free(); }
    void free() {
        if (ctx) rmt3d_destroy(ctx);
        ctx = nullptr;
        trs.ctx = watch.ctx = expper.watch.ctx = nullptr;
    }

    int initialize() {
        int e = rmt3d_initialize(ctx);
        if (e < 0) p_fatal_error(rmt3d_last_error(ctx), 1);
        dt = rmt3d_get_dt(ctx);
        time = rmt3d_get_time(ctx);
        return e;
    }
    // run_sim assigns f3d.dt = sim.dt first (run_sim.cc:105-106)
    void init_iter(int init_err) {
        rmt3d_set_dt(ctx, dt);
        if (rmt3d_init_iter(ctx, init_err) < 0) p_fatal_error(rmt3d_last_error(ctx), 1);
        time = rmt3d_get_time(ctx);
    }
    void set_dt(double d) {
        rmt3d_set_dt(ctx, d);
        dt = d;
    }
    int step_forward(int debug) {
        int e = rmt3d_step(ctx, debug);
        if (e < 0) p_fatal_error(rmt3d_last_error(ctx), 1);
        time = rmt3d_get_time(ctx);
        double mac = 0, fem = 0;
        rmt3d_diag(ctx, 1, &mac);
        rmt3d_diag(ctx, 2, &fem);
        snprintf(outbuf, sizeof outbuf, "MAC %d V-cycles, FEM %d V-cycles\n", (int)mac, (int)fem);
        return e;
    }

    // ---- output (stepper::write_files, dump_all) ----
    void write_slice(write_params wp, const char *filename) {
        if (rmt3d_write_slice(ctx, wp.dim, wp.point, wp.o_type, wp.obj_id, wp.format, filename))
            p_fatal_error(rmt3d_last_error(ctx), 1);
    }
    void output_contours(const char *filename) {
        if (rmt3d_write_contours(ctx, filename)) p_fatal_error(rmt3d_last_error(ctx), 1);
    }
    void write_chk_pt(int snum, int fnum, const char *chk_dirname) {
        if (rmt3d_write_checkpoint(ctx, snum, fnum, chk_dirname)) p_fatal_error(rmt3d_last_error(ctx), 1);
        // the reference compresses each file here (fluid_3d_chk_pt.cc:54-60)
        static const char *names[8] = {"vx", "vy", "vz", "p", "X", "Y", "Z", "OID"};
        char cmd[1024];
        for (const char *f : names) {
            snprintf(cmd, sizeof cmd, "rm -f %s/%s.xz; xz %s/%s", chk_dirname, f, chk_dirname, f);
            if (system(cmd)) printf("compression command %s went wrong!\n", cmd);
        }
    }

    // ---- diagnostics (stepper::collect_macro_data) ----
    double div_u() {
        double v = 0;
        rmt3d_diag(ctx, 0, &v);
        return v;
    }
    // collect_macro_data calls this first; the other three read its record
    double div_u(double_int &gmx, double_int &gmn) {
        if (rmt3d_macro_data(ctx, &macro)) p_fatal_error(rmt3d_last_error(ctx), 1);
        gmx.value = macro.divu_max;
        gmx.rank = 0;
        gmn.value = macro.divu_min;
        gmn.rank = 0;
        return macro.divu_l2;
    }
    void total_momentum(double (&mv)[3]) {
        for (int s = 0; s < 3; s++) mv[s] = macro.momentum[s];
    }
    double total_energy(double &pot, double &kin, double &elas, double &pow) {
        pot = macro.potential;
        kin = macro.kinetic;
        elas = macro.elastic;
        pow = macro.power;
        return macro.energy;
    }
    double total_dev_in_detF(double &max_dev) {
        max_dev = macro.max_dev_detF;
        return macro.dev_detF_l2;
    }
    void compute_v_centroid() {
        std::vector<double> c(6 * (mgmt->n_obj > 0 ? mgmt->n_obj : 1));
        if (rmt3d_centroids(ctx, c.data())) p_fatal_error(rmt3d_last_error(ctx), 1);
        for (int i = 0; i < mgmt->n_obj; i++) {
            mgmt->avg_x[i] = c[6 * i];
            mgmt->avg_y[i] = c[6 * i + 1];
            mgmt->avg_z[i] = c[6 * i + 2];
            mgmt->avg_velx[i] = c[6 * i + 3];
            mgmt->avg_vely[i] = c[6 * i + 4];
            mgmt->avg_velz[i] = c[6 * i + 5];
        }
    }
    // fluid_3d::error (fluid_3d.cc:2809-2861) compares with sim_ftest's
    // exact solution, which the library does not evaluate; run_sim calls it
    // only when has_exact is set
    void error(double (&err)[4]) {
        for (int i = 0; i < 4; i++) err[i] = 0;
        p_fatal_error("fluid_3d_gpu: error() against the exact solution is not supported", 1);
    }

  private:
    rmt3d_macro macro{};
    char outbuf[128];
};

#endif
