// C ABI of include/rmt3d.h over rmt::Sim.  No exception crosses the boundary.
#include <new>
#include <string>
#include <vector>
#include <algorithm>
#include "fluid3d.h"

struct rmt3d_ctx {
    rmt::Sim sim;
};

static thread_local std::string g_last;

extern "C" {

int rmt3d_create(const rmt3d_config *cfg, const rmt3d_object *objs, int precision, rmt3d_ctx **out) {
    if (!cfg || !out) return RMT3D_ERR_ARG;
    if (cfg->n_obj > 0 && !objs && cfg->sim_type != RMT3D_SIM_FTEST) return RMT3D_ERR_ARG;
    if (precision != RMT3D_PREC_FP64 && precision != RMT3D_PREC_MIXED) return RMT3D_ERR_ARG;
    rmt3d_ctx *c = new (std::nothrow) rmt3d_ctx;
    if (!c) return RMT3D_ERR_NOMEM;
    int rc;
    try {
        rc = c->sim.create(cfg, objs, precision);
    } catch (...) {
        rc = RMT3D_ERR_FATAL;
    }
    if (rc) {
        g_last = c->sim.err;
        delete c;
        *out = nullptr;
        return rc;
    }
    *out = c;
    return 0;
}

void rmt3d_destroy(rmt3d_ctx *ctx) { delete ctx; }

const char *rmt3d_last_error(const rmt3d_ctx *ctx) {
    if (!ctx) return g_last.c_str();
    return ctx->sim.err.c_str();
}

#define GUARD(expr)                     \
    do {                                \
        if (!ctx) return RMT3D_ERR_ARG; \
        try {                           \
            return (expr);              \
        } catch (...) {                 \
            return RMT3D_ERR_FATAL;     \
        }                               \
    } while (0)

int rmt3d_initialize(rmt3d_ctx *ctx) { GUARD(ctx->sim.initialize()); }
int rmt3d_init_iter(rmt3d_ctx *ctx, int init_err) { GUARD(ctx->sim.init_iter(init_err)); }
int rmt3d_prepare_run(rmt3d_ctx *ctx, int init_err, double *dt, int *total) {
    GUARD(ctx->sim.prepare_run(init_err, dt, total));
}
int rmt3d_step(rmt3d_ctx *ctx, int debug) { GUARD(ctx->sim.step(debug)); }
int rmt3d_set_dt(rmt3d_ctx *ctx, double dt) {
    if (!ctx) return RMT3D_ERR_ARG;
    ctx->sim.dt = dt;
    ctx->sim.cfg.dt = dt;
    return 0;
}
double rmt3d_get_dt(const rmt3d_ctx *ctx) { return ctx ? ctx->sim.dt : 0; }
double rmt3d_get_time(const rmt3d_ctx *ctx) { return ctx ? ctx->sim.time : 0; }
int rmt3d_set_time(rmt3d_ctx *ctx, double t) {
    if (!ctx) return RMT3D_ERR_ARG;
    ctx->sim.time = t;
    return 0;
}
int rmt3d_upload_state(rmt3d_ctx *ctx, const rmt3d_state *st) {
    if (!st) return RMT3D_ERR_ARG;
    GUARD(ctx->sim.upload(st));
}
int rmt3d_download_state(rmt3d_ctx *ctx, rmt3d_state *st) {
    if (!st) return RMT3D_ERR_ARG;
    GUARD(ctx->sim.download(st));
}
int rmt3d_band_count(rmt3d_ctx *ctx) { GUARD(ctx->sim.band_count()); }
int rmt3d_write_checkpoint(rmt3d_ctx *ctx, int snum, int fnum, const char *chk_dirname) {
    if (!chk_dirname) return RMT3D_ERR_ARG;
    GUARD(ctx->sim.write_checkpoint(snum, fnum, chk_dirname));
}
int rmt3d_write_contours(rmt3d_ctx *ctx, const char *filename) {
    if (!filename) return RMT3D_ERR_ARG;
    GUARD(ctx->sim.write_contours(filename));
}
int rmt3d_contour(rmt3d_ctx *ctx, int obj, float *pts, float *nrm, int *nvert, int *tri, int *ntri) {
    if (!ctx || !nvert || !ntri || obj < 1) return RMT3D_ERR_ARG;
    try {
        std::vector<float> p, n;
        std::vector<int> t;
        int e = ctx->sim.contour(obj, p, n, t);
        if (e) return e;
        int nv = (int)(p.size() / 3), nt = (int)(t.size() / 3);
        bool fits = pts && nrm && tri && *nvert >= nv && *ntri >= nt;
        int query = !pts && !nrm && !tri;
        *nvert = nv;
        *ntri = nt;
        if (query) return 0;
        if (!fits) return RMT3D_ERR_ARG;
        std::copy(p.begin(), p.end(), pts);
        std::copy(n.begin(), n.end(), nrm);
        std::copy(t.begin(), t.end(), tri);
        return 0;
    } catch (...) {
        return RMT3D_ERR_FATAL;
    }
}
int rmt3d_write_tracers(rmt3d_ctx *ctx, const char *filename, int binary) {
    if (!filename) return RMT3D_ERR_ARG;
    GUARD(ctx->sim.write_tracers(filename, binary));
}
int rmt3d_tracer_positions(rmt3d_ctx *ctx, float *xyz, int *n) {
    if (!ctx || !n) return RMT3D_ERR_ARG;
    try {
        std::vector<float> v;
        int e = ctx->sim.tracer_positions(v);
        if (e) return e;
        int cnt = (int)(v.size() / 3);
        if (!xyz || *n < cnt) { *n = cnt; return xyz ? RMT3D_ERR_ARG : 0; }
        std::copy(v.begin(), v.end(), xyz);
        *n = cnt;
        return 0;
    } catch (...) {
        return RMT3D_ERR_FATAL;
    }
}
int rmt3d_macro_data(rmt3d_ctx *ctx, rmt3d_macro *out) {
    if (!out) return RMT3D_ERR_ARG;
    GUARD(ctx->sim.macro_data(out));
}
int rmt3d_centroids(rmt3d_ctx *ctx, double *out) { GUARD(ctx->sim.centroids(out)); }
int rmt3d_slice(rmt3d_ctx *ctx, int dim, int point, int o_type, int obj_id, double *out, int *len) {
    GUARD(ctx->sim.slice(dim, point, o_type, obj_id, out, len));
}
int rmt3d_slice_noalloc(rmt3d_ctx *ctx, int dim, int point, int o_type, int obj_id, double *out, int *len) {
    GUARD(ctx->sim.slice_noalloc(dim, point, o_type, obj_id, out, len));
}
int rmt3d_write_slice(rmt3d_ctx *ctx, int dim, int point, int o_type, int obj_id, int format, const char *filename) {
    GUARD(ctx->sim.write_slice(dim, point, o_type, obj_id, format, filename));
}
int rmt3d_mg_solve(rmt3d_ctx *ctx, int which, const double *rhs, double *x, double tol, int *ncycles, double *resid) {
    if (!rhs || !x || (which != 0 && which != 1)) return RMT3D_ERR_ARG;
    GUARD(ctx->sim.mg_solve(which, rhs, x, tol, ncycles, resid));
}
int rmt3d_mg_dims(const rmt3d_ctx *ctx, int which, int dims[3]) {
    if (!ctx || !dims || (which != 0 && which != 1)) return RMT3D_ERR_ARG;
    const rmt::MGSolver &G = which ? ctx->sim.fem : ctx->sim.mac;
    for (int a = 0; a < 3; a++) dims[a] = G.L[0].d[a];
    return 0;
}
int rmt3d_diag(rmt3d_ctx *ctx, int which, double *out) {
    if (!ctx || !out) return RMT3D_ERR_ARG;
    try {
        rmt::Sim &S = ctx->sim;
        switch (which) {
        case 0: *out = S.div_u(); return 0;
        case 1: *out = S.mac.last_cycles; return 0;
        case 2: *out = S.fem.last_cycles; return 0;
        case 3: *out = S.mac.last_res; return 0;
        case 4: *out = S.fem.last_res; return 0;
        case 5: *out = S.dt_reg; return 0;
        case 6: *out = S.eps; return 0;
        case 8: *out = S.vmax; return 0;
        case 9: *out = S.K_stiff; return 0;
        case 10: *out = S.mac.nlev; return 0;
        case 11: *out = S.fem.nlev; return 0;
        case 12: *out = S.mac.tail_lev; return 0;
        case 13: *out = S.fem.tail_lev; return 0;
        case 14: *out = S.mac.tail_lds64 + 2 * S.mac.tail_lds32; return 0;
        case 15: *out = S.fem.tail_lds64 + 2 * S.fem.tail_lds32; return 0;
        case 16: *out = S.mac.npass[0]; return 0;
        case 17: *out = S.fem.npass[0]; return 0;
        case 18: *out = S.mac.fdm; return 0;
        case 19: *out = S.fem.fdm; return 0;
        case 24: *out = S.mac.smooth_bs[0]; return 0;
        case 25: *out = S.fem.smooth_bs[0]; return 0;
        case 26: *out = S.mac.fdm64.pb[0] + 10 * S.mac.fdm64.pb[1] + 100 * S.mac.fdm64.pb[2]; return 0;
        case 27: *out = S.fem.fdm64.pb[0] + 10 * S.fem.fdm64.pb[1] + 100 * S.fem.fdm64.pb[2]; return 0;
        case 31: *out = S.mac.mixed ? S.mac.fdm32.fused : S.mac.fdm64.fused; return 0;
        case 32: *out = S.fem.mixed ? S.fem.fdm32.fused : S.fem.fdm64.fused; return 0;
        case 33: *out = 10 * S.mac.v_down + S.mac.v_up; return 0;
        case 34: *out = 10 * S.fem.v_down + S.fem.v_up; return 0;
        case 30: *out = S.last_batches; return 0;
        case 28: *out = S.mac.tuned; return 0;
        case 29: *out = S.fem.tuned; return 0;
        case 20: case 21: case 22: {
            int c[3];
            int e = S.step_counters(c);
            if (e) return e;
            *out = c[which - 20];
            return 0;
        }
        }
    } catch (...) {
        return RMT3D_ERR_FATAL;
    }
    return RMT3D_ERR_ARG;
}
int rmt3d_stage_times(rmt3d_ctx *ctx, double *ms, int n) {
    if (!ms) return RMT3D_ERR_ARG;
    GUARD(ctx->sim.stage_times(ms, n));
}
void *rmt3d_stream(rmt3d_ctx *ctx) { return ctx ? (void *)ctx->sim.stream : nullptr; }
int rmt3d_profile(rmt3d_ctx *ctx, int enable) { GUARD(ctx->sim.profile(enable)); }
int rmt3d_profile_read(rmt3d_ctx *ctx, int which, double *ms, int *launches) {
    if (!ms || !launches) return RMT3D_ERR_ARG;
    GUARD(ctx->sim.profile_read(which, ms, launches));
}
int rmt3d_device_synchronize(rmt3d_ctx *ctx) {
    if (!ctx) return RMT3D_ERR_ARG;
    return hipDeviceSynchronize() == hipSuccess ? 0 : RMT3D_ERR_HIP;
}

}  // extern "C"
