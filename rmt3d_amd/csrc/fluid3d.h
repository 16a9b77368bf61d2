// Host side of the MI355X FSI step: the fluid_3d / sim_manager pair of the
// reference (F3D_Source/fluid_3d.hh:40-469, Sim/sim_manager.hh:101-320)
// re-expressed around device-resident structure-of-arrays state.  Every stage
// of step_forward is a kernel launch on the context's stream.
#pragma once
#include <hip/hip_runtime.h>
#include <mutex>
#include <string>
#include <vector>
#include "../../include/rmt3d.h"
#include "launch.h"

namespace rmt {

std::mutex &hip_setup_mutex();

// Event timing of one kernel class on the context stream: every launch of
// the class is bracketed by two events; durations are summed on collect.
struct KProf {
    static const int CAP = 4096;
    hipEvent_t *ev = nullptr;
    int n = 0;           // events recorded (pairs)
    int enabled = 0;
    void begin(hipStream_t s) {
        if (enabled && n < CAP) hipEventRecord(ev[2 * n], s);
    }
    void end(hipStream_t s) {
        if (enabled && n < CAP) { hipEventRecord(ev[2 * n + 1], s); n++; }
    }
    // the next event pair, for a launch that stamps its own start and end
    // (the kernel's execution alone, without the event records around it)
    bool pair(hipEvent_t &a, hipEvent_t &b) {
        if (!enabled || n >= CAP) return false;
        a = ev[2 * n];
        b = ev[2 * n + 1];
        n++;
        return true;
    }
};

// One projection problem: geometric multigrid behind mg_prob::run
// (Utilities/mg_probs.cc:3-83).
struct MGSolver {
    int fem = 0;
    int prd[3] = {0, 0, 0};
    int nlev = 0;
    MGLevel<double> L[16];
    MGLevel<float> Lf[16];      // FP32 hierarchy (mixed precision)
    bool mixed = false;
    double max_norm_tol = 0;
    double tol = 0, log_tol = 0;
    int v_up = 2, v_down = 2;
    int coarse_sweeps = 4;
    int tail_lev = 0;                      // first level run by the one-workgroup tail kernel
    int tail_lds64 = 0, tail_lds32 = 0;    // tail levels held in LDS
    int npass[16];                         // colour passes per smoothing launch, per level
    int smooth_bs[16] = {};                // workgroup size of single-pass launches (0: default)
    int tuned = 0;                         // 1 tuned by measurement, 2 taken from the cache
    // mg_pred (mg_probs.hh:37-78)
    int iP = 16, iV = 32;
    double V_avg = 32, V_var = 1e6, alpha = 0;
    int last_cycles = 0;
    double last_res = 0;
    KProf *prof = nullptr;     // times the finest-level smoother launches
    KProf *vprof = nullptr;    // times whole V-cycles (one defect-correction step when mixed)
    bool fdm = false;                      // coarsest level solved by fast diagonalisation
    double *d_fdm = nullptr;               // its bases, transposes and 1/eigenvalues (FP64)
    float *f_fdm = nullptr;                // their FP32 copy (mixed)
    FdmDev<double> fdm64{};
    FdmDev<float> fdm32{};
    double *d_pinv = nullptr;              // pseudo-inverse of the coarsest operator (or null)
    float *f_pinv = nullptr;               // its FP32 copy (mixed)
    unsigned long long *d_max = nullptr;
    size_t bytes = 0;                      // device bytes of the level arrays
    bool alloc_failed = false;
    unsigned long long *h_max = nullptr;   // pinned
    double tol_override = 0;               // > 0 replaces max_norm_tol

    void setup(int fem_, const int d[3], const int prd_[3], bool mixed_);
    // picks launch shapes by timing the candidates on the levels (or from
    // the per-process cache of that level shape)
    void tune(hipStream_t s);
    void tune_schedule(hipStream_t s);
    template <typename Cfg>
    std::vector<double> time_iterations(hipStream_t s, int ncand, Cfg configure);
    template <typename T>
    void tune_levels(hipStream_t s, MGLevel<T> *LV, FdmDev<T> &F, bool sweep_fixed);
    void release();
    double resid_max(hipStream_t s);
    // residual check in two halves: queue the max-norm and its read-back,
    // then wait for that read-back only
    void check_queue(hipStream_t s);
    double check_wait();
    hipEvent_t ev_check = nullptr;
    void vcycle64(hipStream_t s, int lev);
    void vcycle32(hipStream_t s, int lev, bool first = false, bool last = false);
    bool fuse_first() const;
    bool fuse_last() const;
    void cycle(hipStream_t s);              // one V-cycle (or one defect-correction step)
    void cycle_body(hipStream_t s);         // its launches
    hipGraphExec_t gexec = nullptr;         // the launches captured once
    bool graph_failed = false;
    bool no_graph = false;
    bool fuse_off = false;                  // RMT3D_MIXED_FUSE=0 at set-up: separate launches
    int fdm_fused_ok = 0;                   // the three-launch FDM fits: bit 0 FP64, bit 1 FP32
    void timed_cycle(hipStream_t s);
    double run(hipStream_t s);
};



class Sim {
  public:
    rmt3d_config cfg;
    std::vector<rmt3d_object> objs_in;
    int precision = RMT3D_PREC_FP64;
    std::string err;
    // a failed host-device copy outside a checked call, reported by check_status
    std::string hip_pending;

    // geometry and constants
    Geo g{};
    Phys P{};
    int n_obj = 0;
    double dt = 0, time = 0;
    int nt = 0;
    double dt_reg = 0, K_stiff = 0, vmax = 1, eps = 0;
    double gradphi[8][3];
    std::vector<FaceSpec> wall_faces;
    std::vector<double> hx, hy, hz;
    InitSpec init{};
    uint32_t nmask = 1;

    // device state
    hipStream_t stream = nullptr;
    ObjC *d_obj = nullptr;
    std::vector<ObjC> h_obj;
    Fluid U{};
    bool fields_written = false;
    int last_batches = 0;              // object batches of the last extrapolation       // a step, upload or checkpoint read has written the fields
    RMap R{};
    Band ext{}, tmp{};
    Scratch W{};
    double *d_wpow = nullptr;
    int *d_status = nullptr;
    int *d_solid_list = nullptr, *d_solid_n = nullptr;   // cells on the solid stress path
    int *h_status = nullptr;           // pinned
    int *h_int = nullptr;              // pinned scratch
    double *h_dbl = nullptr;           // pinned scratch
    unsigned long long *d_u64 = nullptr;
    double *d_partial = nullptr, *d_mean = nullptr;
    void *d_faces = nullptr;
    double *d_rhs_tmp = nullptr;       // FEM rhs for div_u
    double *d_diag = nullptr;          // diagnostics partials and results
    MGSolver mac, fem;
    std::vector<void *> allocs;
    // 0 FEM smoother colour pass (finest), 1 MAC smoother colour pass
    // (finest), 2 stress, 3 advection, 4 FEM V-cycle, 5 MAC V-cycle,
    // 6 checkpoint gather, 7 checkpoint scatter, 8 macro.dat partial sums,
    // 9 centroid partial sums, then the six stage buckets (extrapolation,
    // pressure Poisson, communication, stresses, MAC, advection)
    static const int NPROF_KERNEL = 10;
    static const int NPROF = 16;
    KProf kprof[NPROF];
    int profile(int enable);
    int profile_read(int which, double *ms, int *launches);
    int stage_times(double *ms, int n);

    size_t dev_bytes = 0;       // device bytes allocated by this context
    bool dalloc_failed = false;

    ~Sim();
    static double estimate_bytes(const rmt3d_config &c, const rmt3d_object *o, int prec);
    int create(const rmt3d_config *c, const rmt3d_object *o, int prec);
    int initialize();
    int init_iter(int init_err);
    int prepare_run(int init_err, double *dt_out, int *total);
    int step(int debug);
    double div_u();
    int upload(const rmt3d_state *st);
    int download(rmt3d_state *st);
    int band_count();
    int upload_band(const rmt3d_state *st);
    int step_counters(int out[3]);
    int mg_solve(int which, const double *rhs, double *x, double tol, int *ncyc, double *res);
    int write_checkpoint(int snum, int fnum, const char *dir);
    // tracers (Utilities/tracers.{hh,cc})
    bool trace = false;
    TrParams TQ{};
    TracerDev TR{};
    int init_tracers();
    int tracer_positions(std::vector<float> &xyz);
    int write_tracers(const char *filename, int binary);
    // contours (fluid_3d::output_contours, fluid_3d_io.cc:595-958)
    int contour(int cnum, std::vector<float> &pts, std::vector<float> &nrm, std::vector<int> &tri);
    int write_contours(const char *filename);
    int macro_data(rmt3d_macro *m);
    int centroids(double *out);
    int slice_dims(int dim, int t, int &d1, int &d2) const;
    int slice(int dim, int point, int t, int obj_id, double *out, int *len);
    int slice_noalloc(int dim, int point, int t, int obj_id, double *out, int *len);
    int slice_into(int dim, int point, int t, int obj_id, double *out, int *len, bool noalloc);
    int write_slice(int dim, int point, int t, int obj_id, int format, const char *filename);

  private:
    void *dalloc(size_t bytes);
    Maps maps() const;
    void fill_bc_cc();
    void fill_bc_cf();
    void comm_fluid(int flags);
    void comm_rm(int flags);
    void init_fluid();
    int init_from_checkpoint();
    void init_pres();
    int init_refmap();
    void init_extrapolate();
    void extrapolate();
    void cfl_recheck();
    void compute_grad_basis();
    int check_status(const char *where);
    int mac_solve();
    int pp_solve();
    int update_reference_map();
    void alloc_band(Band &B, int cap);
    void mark(int k);
    void mark_end(int k);
};

}  // namespace rmt
