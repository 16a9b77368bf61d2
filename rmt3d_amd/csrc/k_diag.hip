// Diagnostics reductions of the stepper's macro.dat / centroid.dat output:
// fluid_3d::total_momentum, total_energy, total_dev_in_detF, the MAC part of
// div_u and compute_v_centroid (fluid_3d_utils.cc:507-1346).  Each quantity
// is a per-cell term over the interior cells in the reference's loop order
// (i fastest), summed in a fixed blocked order: 2048 blocks of consecutive
// cells, 256 work-items per block striding the block's range, a tree
// within the block; then 256 work-items each add 8 consecutive block
// partials in order, and a tree adds those.  The same order is the
// oracle's, so FP64 results agree bit for bit; against the reference's
// serial sum they agree to rounding.  Maxima and minima are exact in any
// order.
#include <hip/hip_ext.h>
#include "launch.h"

namespace rmt {

namespace {

constexpr int NBLK = 2048, NTHR = 256, NFIN = NBLK / NTHR;

__device__ __forceinline__ int IXd(const Geo &g, int i, int j, int k) { return i + g.m4 * (j + g.n4 * k); }

// the interior cell of dense index e (i + m*(j + n*k)) as a padded index
__device__ __forceinline__ int cell_of(const Geo &g, int e, int &k) {
    int mn = g.m * g.n;
    k = e / mn;
    int r = e - k * mn;
    int j = r / g.m, i = r - j * g.m;
    return IXd(g, i + 2, j + 2, k + 2);
}

template <int Q>
__device__ __forceinline__ void block_tree(double (*sh)[NTHR], double *v) {
    int t = threadIdx.x;
    for (int q = 0; q < Q; q++) sh[q][t] = v[q];
    __syncthreads();
    for (int off = NTHR / 2; off > 0; off >>= 1) {
        if (t < off)
            for (int q = 0; q < Q; q++) sh[q][t] += sh[q][t + off];
        __syncthreads();
    }
}

}  // namespace

// Terms per cell: 0-2 velocity (total_momentum), 3 kinetic energy
// 0.5 u.u, 4 potential energy g (z - az) of solid cells, 5 elastic energy
// and 6 the squared face-averaged |det F - 1| of solid cells, each over
// the cell's three lower faces and the three lower faces of its upper
// neighbours' maps of the same object (fluid_3d_utils.cc:826-856,
// 1041-1066); the active-stress power of these objects is zero.  Maxima:
// the MAC divergence from the face velocities (max and -min), and the
// squared det F deviation.
__global__ void __launch_bounds__(NTHR) k_macro_partials(Geo g, Fluid U, Maps M, double gravity, double az,
                                                         double *part, double *maxes) {
    __shared__ double sh[7][NTHR];
    __shared__ double mx[3][NTHR];
    int N = g.m * g.n * g.o, b = blockIdx.x, t = threadIdx.x;
    int per = (N + NBLK - 1) / NBLK;
    int s0 = b * per, e0 = min(s0 + per, N);
    double v[7] = {0, 0, 0, 0, 0, 0, 0};
    double dmax = -1e200, dmin = 1e200, devmax = -1;
    const double dhsp[3] = {g.dxsp, g.dysp, g.dzsp};
    const int strides[3] = {1, g.m4, g.mn4};
    for (int e = s0 + t; e < e0; e += NTHR) {
        int k;
        int c = cell_of(g, e, k);
        double u0 = U.vel[0][c], u1 = U.vel[1][c], u2 = U.vel[2][c];
        v[0] += u0;
        v[1] += u1;
        v[2] += u2;
        double kin = 0;
        kin += 0.5 * u0 * u0;
        kin += 0.5 * u1 * u1;
        kin += 0.5 * u2 * u2;
        v[3] += kin;
        unsigned oid = M.r.c[c] & OMASK;
        if (oid > 0) {
            if (gravity != 0) v[4] += gravity * (g.lz[k + 2] - az);
            double el = 0, dd = 0;
            for (int s = 0; s < 3; s++) {
                el += M.r.en[s][c];
                dd += fabs(M.r.detF[s][c] - 1);
                int rp = M.get(c + strides[s], oid);
                if (rp >= 0) {
                    el += M.enr(rp, s);
                    dd += fabs(M.detFr(rp, s) - 1);
                }
            }
            el /= 6.;
            dd /= 6.;
            v[5] += el;
            double sq = dd * dd;
            v[6] += sq;
            if (sq > devmax) devmax = sq;
        }
        double div = 0;
        for (int f = 0; f < 3; f++) div += dhsp[f] * (U.fvel[(2 * f + 1) * 3 + f][c] - U.fvel[(2 * f) * 3 + f][c]);
        if (div > dmax) dmax = div;
        if (div < dmin) dmin = div;
    }
    block_tree<7>(sh, v);
    if (t < 7) part[t * NBLK + b] = sh[t][0];
    mx[0][t] = dmax;
    mx[1][t] = -dmin;
    mx[2][t] = devmax;
    __syncthreads();
    for (int off = NTHR / 2; off > 0; off >>= 1) {
        if (t < off)
            for (int q = 0; q < 3; q++) mx[q][t] = fmax(mx[q][t], mx[q][t + off]);
        __syncthreads();
    }
    if (t < 3) maxes[t * NBLK + b] = mx[t][0];
}

// Terms per cell and object o (primary object id o+1): the solid fraction
// H(phi), the velocity times H and the cell-centre coordinates times H
// (fluid_3d_utils.cc:1294-1316).  blockIdx.y is the object.
__global__ void __launch_bounds__(NTHR) k_centroid_partials(Geo g, Phys P, const ObjC *O, Fluid U, RMap R,
                                                            double *part) {
    __shared__ double sh[7][NTHR];
    int N = g.m * g.n * g.o, b = blockIdx.x, t = threadIdx.x, o = blockIdx.y;
    int per = (N + NBLK - 1) / NBLK;
    int s0 = b * per, e0 = min(s0 + per, N);
    double v[7] = {0, 0, 0, 0, 0, 0, 0};
    for (int e = s0 + t; e < e0; e += NTHR) {
        int k;
        int c = cell_of(g, e, k);
        if ((int)(R.c[c] & OMASK) != o + 1) continue;
        double phiv = obj_phi(O, P.n_obj, o, R.x[0][c], R.x[1][c], R.x[2][c]);
        double sf = heaviside(P, phiv);
        int mn = g.m * g.n, r = e - k * mn, j = r / g.m, i = r - j * g.m;
        v[0] += sf;
        v[1] += U.vel[0][c] * sf;
        v[2] += U.vel[1][c] * sf;
        v[3] += U.vel[2][c] * sf;
        v[4] += g.lx[i + 2] * sf;
        v[5] += g.ly[j + 2] * sf;
        v[6] += g.lz[k + 2] * sf;
    }
    block_tree<7>(sh, v);
    if (t < 7) part[((size_t)o * 7 + t) * NBLK + b] = sh[t][0];
}

// The block partials of one quantity per workgroup (blockIdx.x): work-item
// t adds partials 8t..8t+7 in order, then a tree; the maxima (after the nq
// sums) reduce the same way.
__global__ void __launch_bounds__(NTHR) k_diag_finish(const double *part, int nq, const double *maxes, double *out) {
    __shared__ double sh[NTHR];
    int q = blockIdx.x, t = threadIdx.x;
    bool mx = q >= nq;
    const double *p = mx ? maxes + (size_t)(q - nq) * NBLK : part + (size_t)q * NBLK;
    double s = p[NFIN * t];
    for (int i = 1; i < NFIN; i++) s = mx ? fmax(s, p[NFIN * t + i]) : s + p[NFIN * t + i];
    sh[t] = s;
    __syncthreads();
    for (int off = NTHR / 2; off > 0; off >>= 1) {
        if (t < off) sh[t] = mx ? fmax(sh[t], sh[t + off]) : sh[t] + sh[t + off];
        __syncthreads();
    }
    if (t == 0) out[q] = sh[0];
}

void launch_macro(hipStream_t s, const Geo &g, const Fluid &U, const Maps &M, double gravity, double az,
                  double *scratch, double *out, hipEvent_t start, hipEvent_t stop) {
    double *part = scratch, *maxes = scratch + 7 * NBLK;
    hipExtLaunchKernelGGL(k_macro_partials, dim3(NBLK), dim3(NTHR), 0, s, start, stop, 0, g, U, M, gravity, az, part,
                          maxes);
    hipLaunchKernelGGL(k_diag_finish, dim3(10), dim3(NTHR), 0, s, (const double *)part, 7, (const double *)maxes,
                       out);
}

void launch_centroids(hipStream_t s, const Geo &g, const Phys &P, const ObjC *O, const Fluid &U, const RMap &R,
                      int nobj, double *scratch, double *out, hipEvent_t start, hipEvent_t stop) {
    if (nobj <= 0) return;
    hipExtLaunchKernelGGL(k_centroid_partials, dim3(NBLK, nobj), dim3(NTHR), 0, s, start, stop, 0, g, P, O, U, R,
                          scratch);
    int nq = 7 * nobj;
    hipLaunchKernelGGL(k_diag_finish, dim3(nq), dim3(NTHR), 0, s, (const double *)scratch, nq,
                       (const double *)nullptr, out);
}

size_t diag_scratch_doubles(int nobj) {
    size_t a = 10 * (size_t)NBLK, b = 7 * (size_t)NBLK * (nobj > 0 ? nobj : 1);
    return (a > b ? a : b) + 16 + 7 * (size_t)(nobj > 0 ? nobj : 1);
}

}  // namespace rmt
