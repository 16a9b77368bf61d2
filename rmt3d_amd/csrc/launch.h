// Host-side declarations of the kernel launchers (one translation unit per
// stage group: k_fluid.hip, k_mg.hip, k_band.hip).
#pragma once
#include <string>
#include "dev.h"

namespace rmt {

struct FaceSpec {
    int f;
    int bct[3];
    double bcv[3];
};

struct InitSpec {
    int has_exact;
    int vel_prof_num;
    double vel_profile[16];
    int type[64];
    double eig[64][3];
    double omega[64][3];
};

// ---- k_fluid.hip ----
// list: Ng ints of scratch for the solid cells, count: one int
void launch_stress(hipStream_t s, const Geo &g, const Phys &P, const ObjC *O, const Fluid &U, const Maps &M,
                   int *list, int *count);
// list/count: the solid-path cells gathered by launch_stress of this step
void launch_advect(hipStream_t s, const Geo &g, const Phys &P, const ObjC *O, const Fluid &U, const Maps &M,
                   const int *list, const int *count, int *status);
void launch_upwind(hipStream_t s, const Geo &g, const Fluid &U, const Maps &M, int *status);
// the wall faces' ghost fills, in order (set_ghost_cc / set_ghost_cf)
void launch_bc_cc_faces(hipStream_t s, const Geo &g, const FaceSpec *f, int nf, const Fluid &U);
void launch_bc_cf_faces(hipStream_t s, const Geo &g, const FaceSpec *f, int nf, const Fluid &U);
void launch_periodic(hipStream_t s, const Geo &g, double *const *d, int nd, uint32_t *const *u, int nu, int corner);
void launch_periodic_band_fx(hipStream_t s, const Geo &g, const Band &B);
void launch_mac_rhs(hipStream_t s, const Geo &g, const Fluid &U, double *b, double fac);
void launch_mac_extract(hipStream_t s, const Geo &g, const Fluid &U, const double *x);
void launch_sub_qgrad(hipStream_t s, const Geo &g, const Fluid &U, double f2);
// list/count: the solid-path cells gathered by the launch_stress just before
void launch_ustar(hipStream_t s, const Geo &g, const Phys &P, const ObjC *O, const Fluid &U, const Maps &M,
                  const int *list, const int *count);
void launch_pp_rhs(hipStream_t s, const Geo &g, const Fluid &U, const double gradphi[8][3], const FaceSpec *faces,
                   int nwall, void *faces_dev, double *r);
size_t face_struct_size();
void launch_pp_extract(hipStream_t s, const Geo &g, const Fluid &U, const double *x, const double *mean);
void launch_sub_pgrad(hipStream_t s, const Geo &g, const Fluid &U, double dt);
void launch_divu_terms(hipStream_t s, const Geo &g, const double *r, double *out);
// zero_rest: also clear the cell record's other fields (not needed while
// they still hold the zeros of their allocation)
void launch_init_fluid(hipStream_t s, const Geo &g, const Phys &P, const ObjC *O, const InitSpec &is, const Fluid &U,
                       unsigned long long *vmax_bits, int zero_rest = 1);
void launch_init_pres(hipStream_t s, const Geo &g, const Phys &P, const Fluid &U);
void launch_init_refmap(hipStream_t s, const Geo &g, const Phys &P, const ObjC *O, const InitSpec &is, const RMap &R,
                        int *status);

// ---- k_mg.hip ----
template <typename T>
struct MGLevel {
    int d[3];
    int nn;
    T *x, *b, *r;
};

// a level with at most this many unknowns ends the hierarchy and is solved
// exactly with the pseudo-inverse of its operator
constexpr int MG_DIRECT_NN = 160;

// ... unless the hierarchy first reaches a coarse level (below the finest)
// with more than MG_TAIL_NN and at most MG_FDM_NN unknowns: that level ends
// it and is solved exactly by fast diagonalisation
constexpr int MG_FDM_NN = 300000;
constexpr int MG_FDM_DMAX = 512;      // and at most this many per dimension
// the fast-diagonalisation operands of one level: per dimension a, the 1D
// eigenbasis S[a] (S[q*d + p] = mode p at point q) and its transpose ST[a];
// dinv[e] = 1 / eigenvalue of the tensor-product mode e, 0 for the null mode
template <typename T>
struct FdmDev {
    const T *S[3], *ST[3], *dinv;
    int pb[3] = {8, 8, 8};   // outputs per work-item of each dimension's passes (2, 4 or 8)
    // FP32 only, the three MFMA launches (planes, pencils, planes): 1 with
    // each plane and its bases staged in LDS, 2 with direct operands (one
    // workgroup per 16-wide output block); 0: one launch per contraction
    int fused = 0;
    int forms = 0;           // bit 0: form 1 fits this level, bit 1: form 2
    // form 2: S and ST zero-padded to X x X, X = d rounded up to 16
    const T *SP[3] = {nullptr, nullptr, nullptr}, *STP[3] = {nullptr, nullptr, nullptr};
};
// LDS bytes of the staged FP32 three-launch form for a level of dimensions
// d, and enabling the MFMA forms: returns the forms that fit (bit 0 staged,
// bit 1 direct: every extent <= 128, unless RMT3D_FDM_DIRECT=0)
size_t mg_fdm_fused_bytes(const int d[3]);
int mg_fdm_fused_enable(const int d[3]);
// host: the 1D basis of one dimension (fem 1: Q1 nodes, 0: MAC cells) and
// its eigenvalues
void mg_fdm_basis(int fem, int d, int prd, double *S, double *lam);
// x = S D^+ S^T b on level L (six contractions through L.r and L.x, in
// six launches or, FP32 with F.fused, three on MFMA)
template <typename T>
void launch_mg_fdm(hipStream_t s, const MGLevel<T> &L, const FdmDev<T> &F);
// one of those contractions: along dimension a, pb outputs per work-item
template <typename T>
void launch_mg_fdm_pass(hipStream_t s, const MGLevel<T> &L, int a, int pb, const T *W, const T *dinv, const T *in,
                        T *out);

// the levels below a size threshold, processed by one workgroup
constexpr int MG_TAIL_MAXLEV = 8;       // a level of <= 6144 unknowns coarsens at most 4 more times
constexpr int MG_TAIL_NN = 6144;       // largest level handed to the tail kernel
constexpr size_t MG_TAIL_LDS_MAX = 160 * 1024 - 1024;   // dynamic LDS of the tail (gfx950: 160 KiB per workgroup)
template <typename T>
struct MGTail {
    MGLevel<T> L[MG_TAIL_MAXLEV];
    int n;
};

// np consecutive colour passes in one launch: one workgroup per tile, the
// tile and a halo of np nodes staged in LDS (k_sweep)
struct SweepPlan {
    int nt[3];     // tiles per dimension
    int ts[3];     // tile extent per dimension (the last tile may be shorter)
    int np;        // colour passes per launch = halo width
    int c0;        // first colour of the launch
    int block;     // work-items per workgroup
    int region;    // unknowns of the largest tile plus halo
    int lds;       // bytes of LDS: unknowns and right-hand sides of that region
};
// the plan for np passes on a level of dims d whose unknowns take esize bytes
SweepPlan mg_sweep_plan(int fem, const int d[3], const int prd[3], int esize, int np);
// allows k_sweep the dynamic LDS that its largest plan can ask for
void mg_sweep_enable();
// reads the unknowns from xin and writes every unknown to xout (xin != xout)
template <typename T>
void launch_mg_sweep(hipStream_t s, int fem, const MGLevel<T> &L, const int prd[3], const SweepPlan &plan,
                     const T *xin, T *xout, hipEvent_t start = nullptr, hipEvent_t stop = nullptr);
template <typename T>
void launch_mg_copy(hipStream_t s, const T *a, T *b, int n);

// smooth one colour (MAC: red-black SOR, FEM: four colours by (j,k) parity);
// non-null start/stop events are stamped with the kernel's own start and end
template <typename T>
// x64 (FP32 levels of the mixed path only): also add each updated node's
// value to the FP64 array x64 (the final sweep of the correction)
void launch_mg_smooth(hipStream_t s, int fem, const MGLevel<T> &L, const int prd[3], int color,
                      hipEvent_t start = nullptr, hipEvent_t stop = nullptr, int bs = 0, double *x64 = nullptr);
template <typename T>
void launch_mg_resid(hipStream_t s, int fem, const MGLevel<T> &L, const int prd[3]);
template <typename T>
void launch_mg_restrict(hipStream_t s, int fem, const MGLevel<T> &F, const MGLevel<T> &C, const int prd[3]);
// the residual of F restricted into C.b (C.x = 0): fused for MAC levels,
// which then leave F.r untouched
template <typename T>
void launch_mg_resid_restrict(hipStream_t s, int fem, const MGLevel<T> &F, const MGLevel<T> &C, const int prd[3]);
template <typename T>
void launch_mg_prolong(hipStream_t s, int fem, const MGLevel<T> &F, const MGLevel<T> &C, const int prd[3],
                       int assign = 0);
// the rest of a V-cycle from L[0] (finest of the tail) to L[n-1] (coarsest);
// pinv (n_c x n_c, symmetric) solves the coarsest level exactly, or null for
// csweeps smoothing sweeps there
template <typename T>
void launch_mg_tail(hipStream_t s, int fem, const MGLevel<T> *L, int n, const int prd[3], int vdown, int vup,
                    int csweeps, int lds, const T *pinv);
// LDS bytes the tail needs for levels L[0..n-1]; enable returns 1 when the
// kernel may use that much dynamic LDS on the current device
template <typename T>
size_t mg_tail_lds_bytes(const MGLevel<T> *L, int n);
template <typename T>
int mg_tail_enable_lds(size_t bytes);
// FP64 residual r64 = b - A x64 with max|r| accumulated into *maxbits
void launch_mg_resid64_max(hipStream_t s, int fem, const MGLevel<double> &L, const int prd[3],
                           unsigned long long *maxbits);
// mixed precision helpers
// bf = (float) r64; xf (when not null) = 0, or with first_pass the first
// colour pass of the V-cycle from zero (nodes of colour 0), zero elsewhere
void launch_resid_to_f32(hipStream_t s, int fem, const MGLevel<double> &L, const int prd[3], float *bf, float *xf,
                         int first_pass = 0);
void launch_cast_d2f(hipStream_t s, const double *a, float *b, int n);
void launch_axpy_f2d(hipStream_t s, double *x, const float *e, int n);
void launch_zero_f(hipStream_t s, float *a, int n);

// ---- reductions / scans (k_band.hip) ----
void launch_sum_blocks(hipStream_t s, const double *a, int n, double scale, double *partial);
void launch_finish_sum(hipStream_t s, const double *partial, double *out);
void launch_zero_u64(hipStream_t s, unsigned long long *p, int n);
void launch_zero_i32(hipStream_t s, int *p, int n);

// ---- band / extrapolation (k_band.hip) ----
// objects whose extrapolation boxes are pairwise disjoint, run together
constexpr int BATCH_MAX = 64;
struct BoxBatch {
    int n;                       // objects in the batch
    int total;                   // cells of all their boxes
    int oid[BATCH_MAX];          // object ids (1-based)
    int lo[BATCH_MAX][3], ext[BATCH_MAX][3];
    int start[BATCH_MAX + 1];    // first work-item of each box
};

struct Scratch {
    uint32_t *mask;
    unsigned char *flag;
    signed char *lid;
    double *val[3];
    int *list[2];
    int *cnt;            // 2 counters (list sizes)
    // staging
    int *st_cell;
    uint32_t *st_c;
    double *st_x[3], *st_xpred[3];
    double *st_en[3], *st_detF[3];   // face energies and det F carried by a redraw
    int *st_n;
    int st_cap;
    int *perm;
    int *fill;
    int *scan_tmp;       // block sums
    int *err;            // corrector problem count
    int *box;            // per object: the extrapolation's index box (8 ints)
    int *h_box;          // its pinned host copy
    BoxBatch *batch;     // batches of objects with disjoint boxes (device)
    BoxBatch *h_batch;   // the same, pinned host
    Band tband;          // previous band (get_refmap_prev)
};

// returns the number of object batches it ran, or -1 when copying the
// object boxes between host and device failed
int band_extrapolate(hipStream_t s, const Geo &g, const Phys &P, const ObjC *O, RMap R, Band B, Scratch &W,
                     int nlayers, int max_rs, const double *wpow_dev, uint32_t &nmask, int *status, int n_obj);
// carry: take each entry's face energies and det F from the staging (a
// redraw moves existing maps) instead of marking them not computed (fresh
// extrapolated maps)
void band_build_from_staging(hipStream_t s, const Geo &g, Band B, Scratch &W, int nlayers, int *status,
                             int carry = 0);
void launch_corrector(hipStream_t s, const Geo &g, const Phys &P, const ObjC *O, RMap R, Band B, Band T, int *err);
void launch_revert(hipStream_t s, const Geo &g, RMap R, Band B, Band T);
void launch_full_update(hipStream_t s, const Geo &g, RMap R, Band B);
void launch_redraw(hipStream_t s, const Geo &g, const Phys &P, const ObjC *O, RMap R, Band B, Scratch &W, int nlayers,
                   int *status);
void launch_xpred(hipStream_t s, const Geo &g, const Phys &P, const Fluid &U, RMap R);
void launch_band_x_eq_xpred(hipStream_t s, Band B);

// ---- k_io.hip ----
// slice of output type t (write_params numbering) normal to dim at point;
// out holds d1*d2 values; r0 is the pressure-Poisson rhs (type 18 only)
void launch_slice(hipStream_t s, const Geo &g, const ObjC *O, int nobj, const Fluid &U, const Maps &M,
                  const double *r0, int t, int obj_id, int dim, int point, int d1, int d2, double *out);
// The allocation-free slice's fixed device scratch: at most SLICE_SCRATCH_MAX
// values (4 MiB per GPU, a slice of up to 724 x 724 points) on devices
// 0 .. SLICE_SCRATCH_DEVICES-1.  Returns this device's copy (looked up once
// per device, with that device current), or null with err set.
constexpr size_t SLICE_SCRATCH_MAX = size_t(1) << 19;
constexpr int SLICE_SCRATCH_DEVICES = 64;
double *slice_scratch(std::string &err);
// the eight checkpoint files as 8 x o x (m+1)(n+1) doubles; start/stop
// (may be null) are recorded at the kernel's own start and end
void launch_chk_pack(hipStream_t s, const Geo &g, const ObjC *O, int nobj, const Fluid &U, const Maps &M,
                     double *out, hipEvent_t start, hipEvent_t stop);
void launch_chk_unpack(hipStream_t s, const Geo &g, const Fluid &U, const RMap &R, const double *in,
                       hipEvent_t start, hipEvent_t stop);
void launch_chk_reset(hipStream_t s, const Geo &g, const Fluid &U, const RMap &R);

// ---- k_diag.hip ----
// macro.dat sums into out[0..6] (momentum 3, kinetic, potential, elastic,
// squared det F deviation) and maxima out[7..9] (div max, -div min, max
// squared det F deviation, -1 when no solid cell)
void launch_macro(hipStream_t s, const Geo &g, const Fluid &U, const Maps &M, double gravity, double az,
                  double *scratch, double *out, hipEvent_t start = nullptr, hipEvent_t stop = nullptr);
// per object 7 sums (H, u H, v H, w H, x H, y H, z H) into out[7*o + q]
void launch_centroids(hipStream_t s, const Geo &g, const Phys &P, const ObjC *O, const Fluid &U, const RMap &R,
                      int nobj, double *scratch, double *out, hipEvent_t start = nullptr, hipEvent_t stop = nullptr);
size_t diag_scratch_doubles(int nobj);

// exclusive scan (k_band.hip): bsum needs n / 1024 + 1 ints
void launch_exclusive_scan(hipStream_t s, const int *in, int *out, int *bsum, int n, int *total);

// ---- k_contour.hip ----
// one object's zero contour (fluid_3d::send_contour_data / output_contour,
// fluid_3d_io.cc:595-945): phi per padded cell, then per node of the
// (m+1)(n+1)(o+1) lattice of cell centres its vertex and triangle counts
struct CtBufs {
    double *phi;                  // Ng
    int *vcnt, *voff, *tcnt, *toff, *vidx, *bsum, *tot;
    float *pts, *nrm;             // 3 per vertex
    int *tri;                     // 3 per triangle
    const double *gx, *gy, *gz;   // cell-centre coordinates, m+1 / n+1 / o+1 entries
};
void launch_contour_count(hipStream_t s, const Geo &g, const ObjC *O, int nobj, const Maps &M, int cnum,
                          const CtBufs &B);
void launch_contour_emit(hipStream_t s, const Geo &g, const CtBufs &B);

// ---- k_tracers.hip ----
// tracer constants: cell size, the padded grid's lowest cell centre (low),
// the box (lo, hi), padded node counts, periodicity and the boundary type
// and value of each velocity component on its own low / high face
struct TrParams {
    double h[3], low[3], lo[3], hi[3];
    int lim[3];
    int prd[3];
    int bct[3][2];
    double bcv[3][2];
};
// tracer state, one entry per tracer id (structure of arrays)
struct TracerDev {
    int n;
    int *alive;
    int *p[3];
    double *rp0[3], *rpf[3], *v0[3], *vf[3];
};
void launch_tracer_predict(hipStream_t s, const TrParams &Q, const TracerDev &T, const Fluid &U, double dt);
void launch_tracer_advect(hipStream_t s, const TrParams &Q, const TracerDev &T, const Fluid &U, double dt);

// ---- chk_pt.cpp (host) ----
extern const char *const CHK_NAMES[8];
// writes the eight files from their images (8 x o x L doubles)
int chk_write_files(const char *dir, const double *img, int o, size_t L, std::string &err);
// reads them back as fluid_3d::initialize_from_chk_point does: 0, 1 (a file
// cannot be opened), 2 (a slice read returned nothing), <0 library error
int chk_read_files(const char *dir, double *img, int o, size_t L, std::string &err);
// save_matrix / save_text / save_gnuplot (fluid_3d_io.cc:57-90)
int slice_write_file(const char *filename, int format, const double *v, int d1, int d2, const double *a1,
                     const double *a2, int m);

}  // namespace rmt
