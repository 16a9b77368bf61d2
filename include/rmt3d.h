/*
 * rmt3d.h -- C ABI of the MI355X-native RMT3D fluid-structure step.
 *
 * One context (rmt3d_ctx) is one simulation: the fields live in HBM, the
 * context owns them and its HIP stream.  A context is not thread-safe;
 * distinct contexts are independent (this is the serving unit).  No C++
 * exception crosses this boundary: every call returns an int status
 * (0 = ok, >0 = the reference's step error codes, <0 = library error) and
 * rmt3d_last_error() gives the message.
 *
 * Index order of every grid array handed across the boundary is the
 * reference's padded order: cell (i,j,k) of the (m+4)(n+4)(o+4) grid with
 * two ghost layers sits at i + (m+4)*(j + (n+4)*k)   (fluid_3d_utils.cc:89-91,
 * fluid_3d.cc:45 u_mem / rm_mem).
 */
#ifndef RMT3D_H
#define RMT3D_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- error codes of fluid_3d::step_forward (Utilities/defs.hh:47-50) ---- */
#define RMT3D_NORMAL_EXIT        0
#define RMT3D_REFMAP_UPDATE_ERR  1
#define RMT3D_PROJECTION_ERR     2
#define RMT3D_STEADY_SOLN_EXIT   100
/* library-side errors */
#define RMT3D_ERR_ARG           (-1)
#define RMT3D_ERR_HIP           (-2)
#define RMT3D_ERR_CFG           (-3)
#define RMT3D_ERR_FATAL         (-4)   /* a reference p_fatal_error path */
#define RMT3D_ERR_NOMEM         (-5)

/* ---- object types supported on this path (Utilities/defs.hh:62-75) ---- */
#define RMT3D_BASIC_SPHERE 0
#define RMT3D_SPHERE       1

/* ---- sim types (Execs/run_sim.cc:13-32) ---- */
#define RMT3D_SIM_FTEST   0
#define RMT3D_SIM_STEST   1
#define RMT3D_SIM_OBJECTS 2
#define RMT3D_SIM_MILLS   3

#define RMT3D_MAX_VEL_PROF 8
#define RMT3D_N_EXTRA_SPECS 10

/* Arithmetic precision of a context. */
#define RMT3D_PREC_FP64  0   /* every stage in FP64 (the parity path)           */
#define RMT3D_PREC_MIXED 1   /* FP32 multigrid preconditioner inside an FP64
                                defect-correction loop; all other stages FP64    */

/* One immersed object: the per-object keys of the .cfg
 * (object_list, basic_specs, extra_specs, srho, shear_mod;
 * Sim/sim_params.cc:275-314). */
typedef struct rmt3d_object {
    int    type;                           /* object_list              */
    double basic[4];                       /* centre x,y,z, radius     */
    double extra[RMT3D_N_EXTRA_SPECS];     /* extra_specs              */
    double srho;                           /* solid density            */
    double shear_mod;                      /* G                        */
} rmt3d_object;

/* POD mirror of the sim_params fields used on the step path
 * (Sim/sim_params.hh, defaults at Sim/sim_params.cc:10-41). */
typedef struct rmt3d_config {
    int    sys_size[3];          /* m n o (cells)                          */
    int    prd[3];               /* x_prd y_prd z_prd after BC consolidation */
    double xyz_bounds[6];        /* ax bx ay by az bz                      */
    int    bct[3][6];            /* bct[component][face]                   */
    double bcv[3][6];            /* bcv[component][face]                   */
    int    walls[6];
    double wall_pos[6];
    double wall_dist;
    double fmu, frho, fdt_pad;
    double dt;                   /* user dt (before CFL)                   */
    double T;
    int    frames;
    int    stop_short;
    int    num_iters;
    int    nlayers;
    int    max_extrap_rs;
    double wt_n, ex_visc_mult, ev_trans_mult, sdt_pad, dt_ex_pad, weight_fac;
    double gravity;
    int    godunov, impl, has_exact;
    int    sim_type;             /* RMT3D_SIM_*                            */
    int    vel_prof_num;
    double vel_profile[2 * RMT3D_MAX_VEL_PROF];
    int    dump_code, ntracers, debug_flag, chk_num, chk_step;
    int    n_obj;
    /* output and restart keys (Sim/sim_params.hh:20-84): carried through to
     * recover.cfg and the slice writer; defaults as sim_params.cc:12-20 */
    int    suppress, display, chkpt_freq, omp_num_thr, debug_obj_id;
    int    obj_body, data_fmt, out_flag, output_dim, output_ind;
    double current_time;         /* cur_time (restart time)                 */
    char   dirname[256];         /* output directory: the .cfg path with
                                    ".cfg" replaced by ".out"               */
    char   recover_dirname[256]; /* where chk.%05d directories are sought   */
    char   chk_dirname[256];     /* recover_dirname/chk.<chk_num>, set by the
                                    parser when its recover.cfg exists; ""
                                    otherwise (sim_params.cc:374-398)      */
} rmt3d_config;

typedef struct rmt3d_ctx rmt3d_ctx;

/* Host view of the state, in the reference's padded index order.
 * Every pointer may be NULL (that field is skipped).  Grid arrays have
 * (m+4)(n+4)(o+4) entries; the band arrays have nband entries. */
typedef struct rmt3d_state {
    double  *u, *v, *w;          /* field::vel                  (fields.hh:36)  */
    double  *p, *dp, *q;         /* nodal p, dp; MAC q          (fields.hh:38-40)*/
    double  *dvel;               /* 3 x Ng, component-major     (fields.hh:42)  */
    double  *fvel;               /* 18 x Ng, face*3+comp major  (fields.hh:46)  */
    double  *sigma;              /* 9 x Ng, lower-face*3+comp   (fields.hh:55)  */
    uint32_t *rm_c;              /* ref_map::c                  (fields.hh:287) */
    double  *rm_x;               /* 3 x Ng                                      */
    double  *rm_xpred;           /* 3 x Ng                                      */
    /* band ("extraps" multimap, extrap_mmap.hh:14-169), sorted by cell,
     * entries of one cell in multimap order */
    int      nband;              /* in: capacity of the arrays; out: count     */
    int32_t *band_cell;
    uint32_t *band_c;
    double  *band_x;             /* 3 x nband                                   */
    double  *band_xpred;         /* 3 x nband                                   */
} rmt3d_state;

/* ---------------------------------------------------------------------- */
/* .cfg parsing: sim_params::sim_params(const char*) (Sim/sim_params.cc:10-410).
 * Fills cfg and up to max_obj objects; returns 0, or RMT3D_ERR_CFG on an
 * unknown keyword / unreadable file (the reference exits(1) there). */
int rmt3d_parse_cfg(const char *path, rmt3d_config *cfg,
                    rmt3d_object *objs, int max_obj);
/* Defaults of sim_params (Sim/sim_params.cc:10-41) without a file. */
void rmt3d_default_config(rmt3d_config *cfg);

/* fluid_3d::fluid_3d(geometry*, sim_manager*, sim_params*) together with
 * sim_manager_alloc (run_sim.cc:13-32, fluid_3d.cc:12-77).  precision is one
 * of RMT3D_PREC_*.  Allocates every device buffer of the context. */
int rmt3d_create(const rmt3d_config *cfg, const rmt3d_object *objs,
                 int precision, rmt3d_ctx **out);
void rmt3d_destroy(rmt3d_ctx *ctx);
const char *rmt3d_last_error(const rmt3d_ctx *ctx);

/* fluid_3d::initialize (fluid_3d.cc:192-233): returns init_err (1 = from
 * scratch) or a negative library error.  With cfg.chk_num > 0 and a
 * non-empty cfg.chk_dirname the state is read from the checkpoint files
 * there (fluid_3d::initialize_from_chk_point, fluid_3d_chk_pt.cc:74-163):
 * 0 = resumed at cfg.current_time; 1 = a file is missing, 2 = a file is
 * empty, and in both cases the simulation is initialised from scratch with
 * chk_num reset to 0, as the reference does.  Files are read uncompressed:
 * where only "<name>.xz" exists the call fails with RMT3D_ERR_ARG (the
 * reference runs `xz -d` on it first). */
int rmt3d_initialize(rmt3d_ctx *ctx);
/* The run_sim.cc:98-106 sequence between initialize() and run(): clamp
 * spars.dt to dt_reg, build the stepper's dt (stepper.hh:211-225), then
 * fluid_3d::init_iter (fluid_3d.cc:235-287).  Returns the stepper dt in *dt. */
int rmt3d_prepare_run(rmt3d_ctx *ctx, int init_err, double *dt, int *total_steps);
/* fluid_3d::init_iter alone (fluid_3d.cc:235-287). */
int rmt3d_init_iter(rmt3d_ctx *ctx, int init_err);
/* fluid_3d::step_forward (fluid_3d.cc:429-577): returns 0, 1, 2 or <0. */
int rmt3d_step(rmt3d_ctx *ctx, int debug);
/* fluid_3d::set_dt (fluid_3d.hh:452-455) and the public dt / time members. */
int rmt3d_set_dt(rmt3d_ctx *ctx, double dt);
double rmt3d_get_dt(const rmt3d_ctx *ctx);
double rmt3d_get_time(const rmt3d_ctx *ctx);
int rmt3d_set_time(rmt3d_ctx *ctx, double t);

/* State transfer in the reference's index order (the checkpoint path,
 * fluid_3d_chk_pt.cc:74-163 reads the same fields).  upload replaces the
 * fields whose pointers are non-NULL; download fills them.  A band upload
 * needs all four band arrays (cells ascending, the entries of one cell in
 * multimap order) and replaces the whole band, each entry as a fresh
 * extrap_mmap insertion (extrap_mmap.cc:3-60): face values and det F zero,
 * face normals not yet computed. */
int rmt3d_upload_state(rmt3d_ctx *ctx, const rmt3d_state *st);
int rmt3d_download_state(rmt3d_ctx *ctx, rmt3d_state *st);
/* Number of band entries currently held (to size rmt3d_state arrays). */
int rmt3d_band_count(rmt3d_ctx *ctx);

/* ---------------------------------------------------------------------- */
/* Checkpoint and slice files (the reference's wire formats). */

/* fluid_3d::write_chk_pt (fluid_3d_chk_pt.cc:3-72): writes the eight files
 * vx vy vz p X Y Z OID (write_params::default_names 0,1,2,7,9,10,11,12) of
 * the current state into the existing directory chk_dirname, each o raw
 * FP64 z-slices of (m+1)(n+1) values in the reference's slice buffer layout,
 * then recover.cfg with chk_num = fnum, chk_step = snum, current_time = the
 * context's time.  Files are left uncompressed (the reference pipes each
 * through `xz`).  Also records fnum / snum / time in the context's config. */
int rmt3d_write_checkpoint(rmt3d_ctx *ctx, int snum, int fnum, const char *chk_dirname);
/* sim_params::write_params (Sim/sim_params.cc:489-606): writes
 * <dir>/recover.cfg from cfg and its n_obj objects. */
int rmt3d_write_params(const rmt3d_config *cfg, const rmt3d_object *objs, const char *dir);
/* fluid_3d::slice (fluid_3d_io.cc:98-418) on one GPU: the global slice at
 * index `point` (0 <= point < cells along dim) normal to dim (0 x, 1 y,
 * 2 z) of output type o_type (0..21, IO/write_params.hh:6-27; obj_id -1
 * picks each point's object by minimum phi).  out receives d1*d2 values,
 * d1 fastest, where (d1, d2) are the two other dimensions in increasing
 * order (FEM node counts for the corner fields 3,4,5,7,18); *len is the
 * capacity on entry and the count on exit. */
int rmt3d_slice(rmt3d_ctx *ctx, int dim, int point, int o_type, int obj_id, double *out, int *len);
/* rmt3d_slice without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Same arguments, results and errors as rmt3d_slice, except
 * that the slice is staged in a fixed per-GPU device scratch of 2^19 values
 * (4 MiB, held until the process ends): a slice of more values returns
 * RMT3D_ERR_ARG.  Calls on one GPU share that scratch, so calls that overlap
 * on one GPU (from several threads or contexts) race; serialise them. */
int rmt3d_slice_noalloc(rmt3d_ctx *ctx, int dim, int point, int o_type, int obj_id, double *out, int *len);
/* fluid_3d::write_slice (fluid_3d_io.cc:3-65): format 0 matrix text
 * (%17.14g), 1 coordinate text, 2 gnuplot binary (float). */
int rmt3d_write_slice(rmt3d_ctx *ctx, int dim, int point, int o_type, int obj_id, int format,
                      const char *filename);

/* Multigrid solve behind mg_prob::run (Utilities/mg_probs.cc:3-83):
 * which = 0 MAC (cell-centred 7-point, simp_mac, mg_probs.hh:497-530),
 * which = 1 FEM (nodal 27-point, simp_fem, mg_probs.hh:458-495).
 * rhs and x are dense arrays in i + j*M + k*M*N order of that problem's
 * unknowns (MAC: m x n x o; FEM: m or m+1 per dimension).  x is the initial
 * guess on entry and the solution on exit.  tol <= 0 selects the reference's
 * max-norm tolerance.  *ncycles receives the number of V-cycles run, *resid
 * the final max-norm residual. */
int rmt3d_mg_solve(rmt3d_ctx *ctx, int which, const double *rhs, double *x,
                   double tol, int *ncycles, double *resid);
/* Unknown counts of the two problems (for sizing rmt3d_mg_solve arrays). */
int rmt3d_mg_dims(const rmt3d_ctx *ctx, int which, int dims[3]);

/* Diagnostics: which = 0 -> div_u L2 (fluid_3d_utils.cc:507-653),
 * 1 -> number of V-cycles in the last MAC solve, 2 -> last FEM solve,
 * 3 -> last MAC residual, 4 -> last FEM residual, 5 -> dt_reg,
 * 6 -> eps (transition zone width), 8 -> vmax (sim_manager), 9 -> K_stiff;
 * multigrid layout (no counterpart in the reference): 10/11 -> MAC/FEM
 * level count, 12/13 -> first level run by the one-workgroup tail kernel,
 * 14/15 -> tail levels held in LDS (bit 0 FP64 hierarchy, bit 1 FP32),
 * 16/17 -> colour passes per smoothing launch on the finest level,
 * 18/19 -> 1 when the coarsest level is solved by fast diagonalisation;
 * branch counters of the last rmt3d_step (no counterpart in the reference):
 * 20 -> collision_stress contributions (lower face x object pair, both
 * stress passes), 21 -> primary maps redrawn into the band, 22 -> band maps
 * redrawn onto the primary grid (redraw_boundary, fluid_3d_ex.cc:336-394);
 * launch shapes picked by the multigrid autotuner at rmt3d_create (no
 * counterpart in the reference): 24/25 -> MAC/FEM finest-level colour-pass
 * workgroup size (0 = built-in default), 26/27 -> MAC/FEM fast-diagonalisation
 * outputs per work-item, pb_x + 10 pb_y + 100 pb_z, 28/29 -> 0 not tuned,
 * 1 tuned by measurement, 2 taken from the process's cache of that shape;
 * 30 -> object batches of the last band extrapolation (objects whose
 * extrapolation boxes are disjoint share one sequence of launches);
 * 31/32 -> MAC/FEM fast-diagonalisation form in the V-cycles' precision
 * (0 one launch per contraction; three MFMA launches, x-y planes, z
 * pencils, planes: 1 with each plane and its bases staged in LDS, 2 with
 * direct operands, one workgroup per 16-wide output block), picked by the
 * autotuner among the forms that fit the level;
 * 33/34 -> MAC/FEM V-cycle smoothing sweeps, 10 v_down + v_up (2 + 2 on the
 * FP64 path, the reference's default; picked by the autotuner on the mixed
 * path, whose FP32 V-cycle only preconditions the FP64 iteration). */
int rmt3d_diag(rmt3d_ctx *ctx, int which, double *out);

/* Tracers (Utilities/tracers.{hh,cc}): with dump_code & 1 and ntracers > 0
 * rmt3d_initialize places cfg.ntracers tracers (uniform random positions
 * from GSL's taus generator seeded with 1, or those of tr.<chk_num> in the
 * checkpoint directory) and every rmt3d_step advects them with the
 * tricubic interpolant of the velocity (improved Euler, slowed near no-slip
 * walls).  rmt3d_write_tracers is tracers::print (tracers.cc:110-139):
 * binary = an int count then 3 floats per tracer, else "%g %g %g" lines;
 * a tracer that has left the box is written as 0 0 0.
 * rmt3d_tracer_positions gives the same floats (*n: capacity in, count
 * out; xyz NULL queries the count). */
int rmt3d_write_tracers(rmt3d_ctx *ctx, const char *filename, int binary);
int rmt3d_tracer_positions(rmt3d_ctx *ctx, float *xyz, int *n);

/* Contours (fluid_3d::output_contours, fluid_3d_io.cc:595-958): the zero
 * level set of each object's phi over the lattice of cell centres,
 * triangulated by marching cubes, written as the ctr file: int n_obj, then
 * per object int nvert, int ntri and, if nvert > 0, nvert float xyz
 * positions, nvert float normals and ntri int vertex triples.  Vertex
 * positions, normals and order follow the reference; the triangle table is
 * derived by this library (rmt3d_amd/csrc/gen_mc_table.py) with the
 * reference's corner and edge numbering and its triangle count per box, but
 * its own split of each polygon into triangles.  rmt3d_contour returns
 * object obj's (1-based) arrays; with all three arrays NULL it only sets
 * the counts. */
int rmt3d_write_contours(rmt3d_ctx *ctx, const char *filename);
int rmt3d_contour(rmt3d_ctx *ctx, int obj, float *pts, float *nrm, int *nvert, int *tri, int *ntri);

/* The macro.dat quantities of stepper::collect_macro_data
 * (Sim/stepper.hh:378-440), reduced on the device.  Sums run in a fixed
 * blocked order (256 blocks of consecutive cells, a tree within each block,
 * the block partials added in order), not the reference's serial order, so
 * they agree with it to rounding. */
typedef struct rmt3d_macro {
    double divu_max, divu_min;   /* MAC divergence extremes (fluid_3d_utils.cc:572-594)  */
    double divu_l2;              /* the FEM right-hand-side norm (:507-560)              */
    double momentum[3];          /* total_momentum (:783-800)                            */
    double energy, kinetic, elastic, potential, power; /* total_energy (:806-870)        */
    double max_dev_detF, dev_detF_l2;                   /* total_dev_in_detF (:1032-1085) */
} rmt3d_macro;
int rmt3d_macro_data(rmt3d_ctx *ctx, rmt3d_macro *out);
/* fluid_3d::compute_v_centroid (fluid_3d_utils.cc:1290-1346) followed by
 * sim_manager::compute_avgs (sim_manager.hh:223-246): for each object,
 * out[6i..6i+5] = centroid x, y, z and average velocity x, y, z, weighted by
 * the Heaviside solid fraction of the primary cells.  As in the reference,
 * the averages are kept: the wall repulsion of later steps reads the
 * average velocity (sim_manager.cc:212-216).  out may be NULL. */
int rmt3d_centroids(rmt3d_ctx *ctx, double *out);

/* Per-stage device time of the steps run while rmt3d_profile is enabled
 * (ms, from event pairs on the context stream), in the order of the
 * reference's timer buckets "Extrapolation, Pressure Poisson, Communication,
 * Stresses, MAC" (fluid_3d.cc:59) followed by "Advection".  n <= 6. */
int rmt3d_stage_times(rmt3d_ctx *ctx, double *ms, int n);

/* Stream the context's kernels run on (a hipStream_t). */
void *rmt3d_stream(rmt3d_ctx *ctx);

/* Event timing on the context stream (measurement only; no counterpart in
 * the reference, whose timer buckets are wall clocks, Utilities/timer.hh).
 * enable: 0 off; 1 stage buckets and whole V-cycles; 2 also the kernel
 * classes 0-3 (the multigrid V-cycles then run as individual launches
 * instead of replaying their captured graph).  Enabling resets the counts.
 * which: 0 finest-level FEM smoother colour pass,
 * 1 finest-level MAC smoother colour pass, 2 compute_stress, 3 advection,
 * 4 one FEM V-cycle, 5 one MAC V-cycle (with the mixed path, one FP64
 * defect-correction step around an FP32 V-cycle), 6 the checkpoint gather
 * of rmt3d_write_checkpoint, 7 the checkpoint scatter of a resuming
 * rmt3d_initialize, 8 the partial sums of rmt3d_macro_data, 9 those of
 * rmt3d_centroids (6-9 are timed from level 1). */
int rmt3d_profile(rmt3d_ctx *ctx, int enable);
int rmt3d_profile_read(rmt3d_ctx *ctx, int which, double *ms, int *launches);
/* hipDeviceSynchronize on the context's device. */
int rmt3d_device_synchronize(rmt3d_ctx *ctx);

/* ---------------------------------------------------------------------- */
/* Serving: many independent solves on one GPU.  No counterpart in the
 * reference, which runs one solve per process (run_sim.cc:39-165); a request
 * is exactly that run (create, initialize, prepare_run, `steps` steps) on its
 * own context and stream.  Requests are admitted first-in first-out when
 * their estimated device footprint fits the budget beside the running ones;
 * up to max_active run concurrently, one worker thread each, and their
 * kernels overlap on the device.  A request's stage order is the single-solve
 * order, so its result is the same alone and in the mix. */
typedef struct rmt3d_server rmt3d_server;

#define RMT3D_REQ_QUEUED  0
#define RMT3D_REQ_RUNNING 1
#define RMT3D_REQ_DONE    2
#define RMT3D_REQ_FAILED  3

typedef struct rmt3d_request_info {
    int       status;            /* RMT3D_REQ_*                                  */
    int       rc;                /* 0, or the step / library code that ended it  */
    int       steps_done;
    long long cell_steps;        /* m*n*o per completed step                     */
    double    t_submit, t_admit, t_done;  /* seconds since server creation; -1 = not yet */
    double    device_bytes;      /* held while running (estimate while queued)   */
    double    sim_time;          /* simulated time reached                       */
} rmt3d_request_info;

/* Device bytes rmt3d_create would allocate for this configuration. */
int rmt3d_estimate_bytes(const rmt3d_config *cfg, const rmt3d_object *objs,
                         int precision, double *bytes);
/* max_active in 1..16 (worker threads, one stream each). */
int rmt3d_server_create(double hbm_budget_bytes, int max_active, rmt3d_server **out);
/* Queues a request; RMT3D_ERR_ARG if its estimate exceeds the whole budget. */
int rmt3d_server_submit(rmt3d_server *srv, const rmt3d_config *cfg,
                        const rmt3d_object *objs, int precision, int steps, int *req);
/* Blocks until the request is done or failed (timeout_s < 0: no limit);
 * returns 0 when it ended, 1 on timeout.  Do not release a request that
 * another thread waits on. */
int rmt3d_server_wait(rmt3d_server *srv, int req, double timeout_s, rmt3d_request_info *info);
int rmt3d_server_info(rmt3d_server *srv, int req, rmt3d_request_info *info);
const char *rmt3d_server_error(rmt3d_server *srv, int req);
/* Final u, v, w, p, rm_c of an ended request (Ng entries each, padded order). */
int rmt3d_server_result(rmt3d_server *srv, int req, rmt3d_state *st);
/* Frees an ended request's host copy of its fields. */
int rmt3d_server_release(rmt3d_server *srv, int req);
/* Peak device bytes held and peak concurrent requests so far; *now is the
 * server clock in seconds. */
int rmt3d_server_stats(rmt3d_server *srv, double *peak_bytes, int *peak_active, double *now);
/* Cancels queued requests, waits for running ones, frees everything. */
void rmt3d_server_destroy(rmt3d_server *srv);

#ifdef __cplusplus
}
#endif
#endif
