/*
 * ORACLE -- test infrastructure only.
 *
 * A plain-C, FP64, single-threaded restatement of RMT3D's per-timestep
 * fluid-structure step (reference: F3D_Source/fluid_3d.cc, fluid_3d_ex.cc,
 * fields.hh, Extrapolation/extrap_helper.{hh,cc}, Utilities/mg_probs.{hh,cc},
 * Sim/sim_manager.{hh,cc}, Sim/object.{hh,cc}).  Every function cites the
 * reference lines it follows.  It is the checker of the HIP path: only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it,
 * and never as the thing measured or shipped.
 *
 * Layout: the reference's AoS records, in the reference's padded index order
 * (two ghost layers, i fastest), so that a field of the oracle and of the HIP
 * path can be compared element by element after a download.
 *
 * The band ("extraps", extrap_mmap) is held as a cell-sorted entry list with
 * per-cell start/count.  Within one cell the entries keep the multimap's
 * insertion order (object order for fresh extrapolations, a redrawn primary
 * appended last).
 *
 * The multigrid solves replace the external PGMG library, which is absent:
 * the stopping rule, iteration predictor and insurance cycles of mg_prob::run
 * are the reference's; the V-cycle itself (smoother, transfers, coarse solve)
 * is this project's, identical to the HIP implementation, so PGMG's own
 * iterates are "parity unpinned" and only the converged discrete solution is
 * compared.
 */
#ifndef RMT3D_ORACLE_H
#define RMT3D_ORACLE_H

#include <stdint.h>
#include "../include/rmt3d.h"

/* ref_map record (fields.hh:285-311); sigma and act_power are unused on the
 * path and dropped, debug-only fields are not compiled in the reference. */
typedef struct {
    uint32_t c;
    double x[3], xpred[3];
    double nvd[3][3];
    double fx[6][3];
    double gphi[3][3];
    double detF[3];
    double energy[3];
} orc_rmap;

/* band entry = ref_map + its padded cell index */
typedef struct {
    int cell;
    orc_rmap r;
} orc_bentry;

typedef struct {
    int n, cap;
    orc_bentry *e;
    int *start;            /* Ng */
    int *cnt;              /* Ng */
} orc_band;

/* field record (fields.hh:34-55) */
typedef struct {
    double vel[3];
    double p, dp, q;
    double dvel[3];
    double nvd[3][3];
    double fvel[6][3];
    double sigma[3][3];
} orc_field;

/* face_gm (fields.hh:661-740) */
typedef struct {
    int sgn, n;
    int bct[3];
    int t[2];
    int len4[3];
    int len[3];
    int ghost_cf[2];
    int u0_off;
    double bcv[3];
    double h;
} orc_face;

/* a level with at most this many unknowns ends the hierarchy and is solved
 * exactly (pseudo-inverse) */
#define ORC_MG_DIRECT_NN 160
/* The hierarchy instead ends at the first coarse level (below the finest)
 * with more than ORC_MG_SMALL_NN and at most ORC_MG_FDM_NN unknowns, solved
 * exactly by fast diagonalisation */
#define ORC_MG_SMALL_NN 6144
#define ORC_MG_FDM_NN 300000
#define ORC_MG_FDM_DMAX 512      /* and at most this many per dimension */

/* Moore-Penrose pseudo-inverse of a symmetric n x n matrix (row-major) */
void orc_sym_pinv(int n, const double *A, double *P);

/* multigrid level */
typedef struct {
    int d[3];              /* unknowns per dimension */
    int nn;
    double *x, *b, *r;
} orc_mglev;

typedef struct {
    int fem;               /* 0 MAC, 1 FEM */
    int prd[3];
    int nlev;
    orc_mglev lev[16];
    double max_norm_tol;
    double tol, log_tol;
    int v_up, v_down;
    int coarse_sweeps;
    /* pseudo-inverse of the coarsest operator (nn x nn, symmetric) when that
     * level has at most ORC_MG_DIRECT_NN unknowns, else NULL */
    double *pinv;
    /* fast-diagonalisation solve of the coarsest level (fdm = 1): per
     * dimension a, the 1D eigenbasis S[a] (d x d, S[q*d + p] = mode p at
     * point q) and its transpose ST[a]; dinv = 1 / eigenvalue of each
     * tensor-product mode (0 for the null mode) */
    int fdm;
    double *fS[3], *fST[3], *fdinv;
    /* mg_pred (mg_probs.hh:37-78) */
    int iP, iV;
    double V_avg, V_var, alpha;
    int last_cycles;
    double last_res;
} orc_mg;

/* sl_mat (sim_manager.hh:15-74) */
typedef struct {
    double rho, invrho, G, c, ex_visc_mult, ev_trans_mult, dt_pad, dt_ex_pad, ex_mu;
} orc_slmat;

typedef struct {
    int type;
    double c[3], R;
    double eig[3];         /* sphere rm0 scale (object.hh:95-114) */
    double omega0[3];
} orc_obj;

/* tracer (Utilities/tracers.hh:31-56) */
typedef struct {
    int alive;
    int p_crd[3];
    double rp0[3], rpf[3], v0[3], vf[3];
} orc_tracer;

typedef struct orc_sim {
    rmt3d_config cfg;
    int n_obj;
    orc_obj *obj;
    orc_slmat *sm;

    /* geometry */
    int m, n, o, m4, n4, o4, mn4, Ng, G0;
    int prd[3];
    int fm, fn, fo;        /* FEM node counts */
    double dx, dy, dz, dxsp, dysp, dzsp;
    double *lx, *ly, *lz;  /* padded coordinates (fluid_3d.cc:136-138) */

    /* sim_manager */
    double eps, eps_inv, tf1, tf2, tf3, dt_reg, K_stiff, vmax;
    double lxL, lyL, lzL;
    int walls[6];
    double wall_pos[6], wall_dist, winv, gravity, min_wall_acc, wall_acc_mult;
    double *avg_vel;       /* 3 x n_obj, all zero on the path */

    /* fluid_3d */
    double dt, time;
    int nt;
    double gradphi[8][3];
    orc_face faces[6];
    int neigh_face[6];     /* -1 wall, else periodic */

    orc_field *u;
    orc_rmap *rm;
    orc_band ext, tmp;

    /* extrap_helper (extrap_helper.hh:22-443) */
    uint32_t *mask;
    uint32_t nmask, cur_mask;
    int *list, nlist, *next, nnext;
    int n_layers;

    orc_mg mac, fem;
    double mac_res, fem_res;
    int mac_cycles, fem_cycles;

    /* branch counters of the current step (reset by orc_step): face and
     * object-pair contributions of collision_stress, primary maps moved to
     * the band and band maps moved to the primary grid by redraw_boundary */
    long n_coll, n_redraw_out, n_redraw_in;

    /* tracers */
    int trace, ntr;
    orc_tracer *tr;
} orc_sim;

/* orc_core.c */
orc_sim *orc_create(const rmt3d_config *cfg, const rmt3d_object *objs);
void orc_destroy(orc_sim *S);
int orc_initialize(orc_sim *S);
int orc_init_iter(orc_sim *S, int init_err);
int orc_prepare_run(orc_sim *S, int init_err, double *dt, int *total);
int orc_step(orc_sim *S, int debug);
double orc_div_u(orc_sim *S);
double orc_blocked_sum(const double *a, int n, double scale);
double orc_phi(const orc_sim *S, int obj_index, const double xi[3]);
double orc_heaviside(const orc_sim *S, double phi);
void orc_compute_stress(orc_sim *S);
void orc_collision_face(orc_sim *S, int eid, int F, double fs[3]);

void orc_init_extrapolate(orc_sim *S);
void orc_fill_pp_rhs(orc_sim *S, double *rp);

/* orc_io.c: checkpoint and slice formats */
int orc_init_from_chk_point(orc_sim *S);
double orc_slice_point(orc_sim *S, const double *r0, int o_type, int obj_id, int i, int j, int k);
int orc_slice(orc_sim *S, int dim, int point, int o_type, int obj_id, double *g_val);
void orc_chk_images(orc_sim *S, double *img);
void orc_chk_restore(orc_sim *S, const double *img);
int orc_chk_read(const char *dir, int o, size_t L, double *img);
int orc_write_chk_pt(orc_sim *S, const char *dir);

/* orc_tracers.c */
int orc_tracers_init(orc_sim *S);
void orc_tracers_predict(orc_sim *S);
void orc_tracers_advect(orc_sim *S);
void orc_tracer_positions(const orc_sim *S, float *xyz);

/* orc_diag.c: macro.dat / centroid.dat reductions */
void orc_macro(orc_sim *S, double *out);
void orc_centroids(orc_sim *S, double *out);

/* orc_extrap.c */
void orc_band_init(orc_band *B, int Ng);
void orc_band_free(orc_band *B);
void orc_band_reset(orc_band *B, int Ng);
void orc_band_copy(orc_band *dst, const orc_band *src, int Ng);
orc_rmap *orc_get_refmap(orc_sim *S, int ind, unsigned oid);
orc_rmap *orc_get_refmap_pos(orc_sim *S, int ind, unsigned oid, const int pos[3]);
orc_rmap *orc_get_refmap_prev(orc_sim *S, int ind, unsigned oid);
void orc_extrapolate(orc_sim *S);
void orc_redraw_boundary(orc_sim *S);
int orc_least_square_fit(int npt, const double *pos, const double *val,
                         const double *w, double wsum, double out_avgpos[3],
                         double out_avgval[3], double coeff[9], double *det);

/* orc_mg.c */
void orc_mg_setup(orc_mg *G, int fem, const int d[3], const int prd[3]);
void orc_mg_free(orc_mg *G);
double orc_mg_run(orc_mg *G);
double orc_mg_resid_max(orc_mg *G, int lev);
void orc_mg_vcycle(orc_mg *G);
void orc_mg_apply(const orc_mg *G, int lev, const double *x, double *y);

/* orc_comm.c: ghost fills */
void orc_face_setup(orc_sim *S);
void orc_set_bc_cc(orc_sim *S);
void orc_set_bc_cf(orc_sim *S);
#define ORC_COMM_VEL    1
#define ORC_COMM_PRES   2
#define ORC_COMM_FVEL   4
#define ORC_COMM_C      8
#define ORC_COMM_X      16
#define ORC_COMM_XPRED  32
#define ORC_COMM_FX     64
#define ORC_COMM_DVEL   2048
void orc_communicate(orc_sim *S, int flags);
void orc_communicate_band_layer(orc_sim *S, unsigned c);
void orc_communicate_band_fx(orc_sim *S);
void orc_comm_mask(orc_sim *S);
int orc_periodic_src(const orc_sim *S, int i, int j, int k, int corner, int *src);

#endif
