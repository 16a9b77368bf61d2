// Device-side views and per-cell helpers of the MI355X FSI step.
//
// Layout in HBM (structure of arrays, one array per component, padded grid
// of (m+4)(n+4)(o+4) cells in the reference's i-fastest order):
//   fluid      vel[3] p dp q dvel[3] fvel[18] sigma[9]          (field)
//   primary    c, x[3], xpred[3], fx[18], gphi[9], detF[3], en[3] (ref_map)
//   band       entry arrays of the same ref_map fields + cell index, sorted
//              by cell, with per-cell start/count (extrap_mmap)
// The kernels that need a cell's maps look them up exactly like
// fluid_3d::get_refmap (fluid_3d_utils.cc:31-84): primary if c == oid, else
// the first band entry of that cell whose object bits equal oid.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "detmath.h"

namespace rmt {

constexpr unsigned OMASK = 1023u;   // defs.hh:13
constexpr int LDIG = 10;            // defs.hh:12
constexpr int MAXOBJ = 256;

struct Geo {
    int m, n, o, m4, n4, o4, mn4, Ng, G0;
    int prd[3];
    int fm, fn, fo;                 // FEM node counts
    double dx, dy, dz, dxsp, dysp, dzsp;
    const double *lx, *ly, *lz;     // padded cell-centre coordinates
};

// sim_manager constants used inside kernels
struct Phys {
    double eps, tf1, tf2, tf3;
    double fmu;
    double dt, time;
    double K_stiff, gravity;
    double min_wall_acc, wall_acc_mult, wall_dist, winv;
    int walls[6];
    double wall_pos[6];
    int sim_type;
    int n_obj;
    // branch counters of the current step: [0] collision contributions
    // (face x object pair), [1] primary maps redrawn into the band, [2] band
    // maps redrawn onto the primary grid
    int *counters;
};

// per-object constants (sl_mat + sphere geometry), device array
struct ObjC {
    double c[3], R;
    double G, ex_mu, ev_trans_mult;
    double avg_vel[3];
};

struct Fluid {
    double *vel[3];
    double *p, *dp, *q;
    double *dvel[3];
    double *fvel[18];
    double *sigma[9];
};

struct RMap {
    uint32_t *c;
    double *phi;         // level-set value of x for its object (NaN: no object), per stress pass
    double *x[3], *xpred[3];
    double *fx[18];
    double *gphi[9];
    double *detF[3], *en[3];
};

struct Band {
    int *start;          // Ng
    int *cnt;            // Ng
    int *cell;           // cap
    uint32_t *c;         // cap
    double *phi;         // cap, as RMap::phi
    double *x[3], *xpred[3];
    double *fx[18];
    double *gphi[9];
    double *detF[3], *en[3];
    int *n;              // device count
    int cap;
};

// A map reference: primary cell index in [0,Ng), band entry e as Ng+e, -1 none.
struct Maps {
    RMap r;
    Band b;
    int Ng;
    __device__ __forceinline__ bool is_prim(int ref) const { return ref < Ng; }
    __device__ __forceinline__ double x(int ref, int c) const {
        return ref < Ng ? r.x[c][ref] : b.x[c][ref - Ng];
    }
    __device__ __forceinline__ double &xr(int ref, int c) const {
        return ref < Ng ? r.x[c][ref] : b.x[c][ref - Ng];
    }
    __device__ __forceinline__ double &xpredr(int ref, int c) const {
        return ref < Ng ? r.xpred[c][ref] : b.xpred[c][ref - Ng];
    }
    __device__ __forceinline__ double &fxr(int ref, int k) const {
        return ref < Ng ? r.fx[k][ref] : b.fx[k][ref - Ng];
    }
    __device__ __forceinline__ double &gphir(int ref, int k) const {
        return ref < Ng ? r.gphi[k][ref] : b.gphi[k][ref - Ng];
    }
    __device__ __forceinline__ double &detFr(int ref, int k) const {
        return ref < Ng ? r.detF[k][ref] : b.detF[k][ref - Ng];
    }
    __device__ __forceinline__ double &enr(int ref, int k) const {
        return ref < Ng ? r.en[k][ref] : b.en[k][ref - Ng];
    }
    __device__ __forceinline__ uint32_t c(int ref) const {
        return ref < Ng ? r.c[ref] : b.c[ref - Ng];
    }
    __device__ __forceinline__ double phi(int ref) const {
        return ref < Ng ? r.phi[ref] : b.phi[ref - Ng];
    }
    // fluid_3d::get_refmap (fluid_3d_utils.cc:31-42)
    __device__ __forceinline__ int get(int ind, unsigned oid) const {
        if (ind < 0 || ind >= Ng) return -1;
        if (r.c[ind] == oid) return ind;
        int s = b.start[ind], e = s + b.cnt[ind];
        for (int a = s; a < e; a++)
            if ((b.c[a] & OMASK) == oid) return Ng + a;
        return -1;
    }
};

__device__ __forceinline__ double dmin(double a, double b) { return (a < b) ? a : b; }
__device__ __forceinline__ double dsign(double x) { return (x >= 0) ? 1.0 : -1.0; }

// fluid_3d::del_lim / del_f / mono_diff (fluid_3d.cc:2310-2340)
__device__ __forceinline__ double del_lim(double u0, double u1, double u2) {
    double t1 = 2 * (u1 - u0), t2 = 2 * (u2 - u1);
    if (t1 * t2 > 0) return dmin(fabs(t1), fabs(t2));
    return 0;
}
__device__ __forceinline__ double del_f(double u0, double u1, double u2) {
    double dl = del_lim(u0, u1, u2);
    double cd = 0.5 * (u2 - u0);
    return dsign(cd) * dmin(fabs(cd), dl);
}
__device__ __forceinline__ double mono_diff(double u0, double u1, double u2, double u3, double u4) {
    double dl = del_lim(u1, u2, u3);
    double cd = 0.5 * (u3 - u1);
    double nd = (del_f(u0, u1, u2) + del_f(u2, u3, u4)) / 6.0;
    return dsign(cd) * dmin(dl, fabs(4.0 / 3.0 * cd - nd));
}

// basic_sphere::phi (object.cc:76-82) via sim_manager::phi (sim_manager.cc:97-100)
__device__ __forceinline__ double obj_phi(const ObjC *O, int nobj, int oi, double x0, double x1, double x2) {
    if (oi < 0 || oi >= nobj) return __builtin_nan("");
    const ObjC &q = O[oi];
    double sq = 0;
    sq += (x0 - q.c[0]) * (x0 - q.c[0]);
    sq += (x1 - q.c[1]) * (x1 - q.c[1]);
    sq += (x2 - q.c[2]) * (x2 - q.c[2]);
    return sqrt(sq) - q.R;
}

// sim_manager::heaviside (sim_manager.hh:259-268)
__device__ __forceinline__ double heaviside(const Phys &P, double phi) {
    if (phi >= P.eps) return 0;
    if (phi <= -P.eps) return 1;
    double tmp = 0.5 - phi * P.tf1 - det_sin(P.tf3 * phi) * P.tf2;
    if (tmp < 0) tmp = 0.;
    else if (tmp > 1) tmp = 1;
    return tmp;
}
__device__ __forceinline__ double tderiv_func_in(const Phys &P, double x) {
    double tmp = 0;
    if (x > -P.eps && x < P.eps) tmp = 0.5 + 0.5 * det_cos(x * P.tf3);
    return tmp;
}
__device__ __forceinline__ double coll_func(const Phys &P, double x) { return x > P.eps ? 0 : (0.5 - x * P.tf1); }

// Slot of this lane in an append to a device list: one atomic per wavefront
// for all its appending lanes (-1 for a lane that does not append).  Every
// active lane of the wavefront must reach the call.
__device__ __forceinline__ int wave_append(bool pred, int *counter) {
    unsigned long long bal = __ballot(pred);
    if (!bal) return -1;
    int lane = __lane_id();
    int first = __ffsll((long long)bal) - 1;
    int base = 0;
    if (lane == first) base = atomicAdd(counter, __popcll(bal));
    base = __shfl(base, first, 64);
    return pred ? base + __popcll(bal & ((1ull << lane) - 1)) : -1;
}

// The same for a whole workgroup (blockDim a multiple of 64, at most 1024):
// one atomic per workgroup.  Every thread of the workgroup must reach the
// call; sh is __shared__ int[17].
__device__ __forceinline__ int block_append(bool pred, int *counter, int *sh) {
    unsigned long long bal = __ballot(pred);
    int lane = __lane_id(), w = threadIdx.x >> 6, nw = blockDim.x >> 6;
    if (lane == 0) sh[w] = __popcll(bal);
    __syncthreads();
    if (threadIdx.x == 0) {
        int run = 0;
        for (int a = 0; a < nw; a++) {
            int c = sh[a];
            sh[a] = run;
            run += c;
        }
        sh[16] = run ? atomicAdd(counter, run) : 0;
    }
    __syncthreads();
    int slot = pred ? sh[16] + sh[w] + __popcll(bal & ((1ull << lane) - 1)) : -1;
    __syncthreads();
    return slot;
}

__device__ __forceinline__ int symind(int i, int j) {
    if (i > j) { int t = j; j = i; i = t; }
    return j + i * (5 - i) / 2;
}

// device status word bits
constexpr int ST_FATAL_NVD = 1;       // normal_derivatives_refmap missing cells
constexpr int ST_FATAL_UPWIND = 2;    // upwinding neighbour map missing
constexpr int ST_FATAL_EXTRAP = 4;    // extrapolation failed
constexpr int ST_FATAL_REDRAW = 8;    // two deep maps
constexpr int ST_BAND_OVERFLOW = 16;  // band capacity exceeded
constexpr int ST_FATAL_INIT = 32;     // overlapping primaries

// words (128 bytes apart) taking the per-wavefront maxima of |u| at init
constexpr int INIT_VMAX_SLOTS = 32;

}  // namespace rmt
