// Checkpoint and slice output on the device: the per-point values of
// fluid_3d::copy_slice_to_buf (fluid_3d_io.cc:257-418) gathered straight into
// the reference's global slice layout, and the checkpoint read-back of
// fluid_3d::copy_chk_slice_to_buf / copy_chk_to_field (fluid_3d_chk_pt.cc:
// 303-352) scattered into the padded fields.  One thread per output point;
// every access is a plain coalesced load or store (the work is HBM-bound
// byte movement).
#include <hip/hip_ext.h>

#include <mutex>
#include <string>

#include "launch.h"

namespace rmt {

namespace {

constexpr unsigned LMASK = 63u << LDIG;   // defs.hh:14

__device__ __forceinline__ int IXp(const Geo &g, int i, int j, int k) { return i + g.m4 * (j + g.n4 * k); }

// ref_map::phi (fields.hh:430-434) of a map reference (primary cell or band
// entry), on x (pred = 0) or xpred (pred = 1)
__device__ __forceinline__ double map_phi(const ObjC *O, int nobj, const Maps &M, int ref, int pred) {
    unsigned o = M.c(ref) & OMASK;
    if (o == 0) return __builtin_nan("");
    if (pred) return obj_phi(O, nobj, (int)o - 1, M.xpredr(ref, 0), M.xpredr(ref, 1), M.xpredr(ref, 2));
    return obj_phi(O, nobj, (int)o - 1, M.x(ref, 0), M.x(ref, 1), M.x(ref, 2));
}

// fluid_3d::min_phi_id (fluid_3d_io.cc:420-432)
__device__ __forceinline__ int min_phi_id(const ObjC *O, int nobj, const Maps &M, int c) {
    int lo = (int)(M.r.c[c] & OMASK), id = -1;
    double mphi = 1e30;
    if (lo > 0) { id = lo; mphi = map_phi(O, nobj, M, c, 0); }
    int s = M.b.start[c], e = s + M.b.cnt[c];
    for (int a = s; a < e; a++) {
        double t = map_phi(O, nobj, M, M.Ng + a, 0);
        if (t < mphi) { mphi = t; id = (int)(M.b.c[a] & OMASK); }
    }
    return id;
}

// The solid-field branch of copy_slice_to_buf (fluid_3d_io.cc:341-412): the
// map that supplies the value for object obj_id at padded cell c, or -1:
// the primary when its object matches, else the first band entry of that
// object in the cell's multimap order.
__device__ __forceinline__ int pick_map(const Maps &M, int c, int obj_id) {
    if ((int)(M.r.c[c] & OMASK) == obj_id) return c;
    int s = M.b.start[c], e = s + M.b.cnt[c];
    for (int a = s; a < e; a++)
        if ((int)(M.b.c[a] & OMASK) == obj_id) return M.Ng + a;
    return -1;
}

}  // namespace

// Value of output type t at local cell/node (i,j,k) (padded index c), as
// copy_slice_to_buf computes it.  r0 is the pressure-Poisson right-hand
// side (type 18 only).
__device__ __forceinline__ double slice_value(const Geo &g, const ObjC *O, int nobj, const Fluid &U, const Maps &M,
                                              const double *r0, int t, int obj_id, int i, int j, int k) {
    int c = IXp(g, i + 2, j + 2, k + 2);
    int sm4 = g.m4, lo = c - g.mn4;
    if (t < 3) return U.vel[t][c];
    if (t == 3) {
        // wx = d(vz)/dy - d(vy)/dz at the node
        const double *a = U.vel[2], *b = U.vel[1];
        double v = 0.25 * g.dysp * (a[c] - a[c - sm4] + a[c - 1] - a[c - 1 - sm4] + a[lo] - a[lo - sm4] + a[lo - 1] -
                                    a[lo - 1 - sm4]) -
                   0.25 * g.dzsp * (b[c] - b[lo] + b[c - 1] - b[lo - 1] + b[c - sm4] - b[lo - sm4] + b[c - 1 - sm4] -
                                    b[lo - 1 - sm4]);
        return v;
    }
    if (t == 4) {
        // wy = d(vx)/dz - d(vz)/dx
        const double *a = U.vel[0], *b = U.vel[2];
        double v = 0.25 * g.dzsp * (a[c] - a[lo] + a[c - 1] - a[lo - 1] + a[c - sm4] - a[lo - sm4] + a[c - 1 - sm4] -
                                    a[lo - 1 - sm4]) -
                   0.25 * g.dxsp * (b[c] - b[c - 1] + b[c - sm4] - b[c - sm4 - 1] + b[lo] - b[lo - 1] + b[lo - sm4] -
                                    b[lo - sm4 - 1]);
        return v;
    }
    if (t == 5) {
        // wz = d(vy)/dx - d(vx)/dy
        const double *a = U.vel[1], *b = U.vel[0];
        double v = 0.25 * g.dxsp * (a[c] - a[c - 1] + a[c - sm4] - a[c - sm4 - 1] + a[lo] - a[lo - 1] + a[lo - sm4] -
                                    a[lo - sm4 - 1]) -
                   0.25 * g.dysp * (b[c] - b[c - sm4] + b[c - 1] - b[c - 1 - sm4] + b[lo] - b[lo - sm4] + b[lo - 1] -
                                    b[lo - 1 - sm4]);
        return v;
    }
    if (t == 6) {
        double s = 0;
        for (int a = 0; a < 3; a++) s += U.vel[a][c] * U.vel[a][c];
        return sqrt(s);
    }
    if (t == 7) return U.p[c];
    if (t == 18) return r0[i + g.fm * j + g.fm * g.fn * k];
    if (t == 19)
        // fvel[face*3 + component]: faces 1-0 (x), 3-2 (y), 5-4 (z)
        return g.dxsp * (U.fvel[3][c] - U.fvel[0][c]) + g.dysp * (U.fvel[10][c] - U.fvel[7][c]) +
               g.dzsp * (U.fvel[17][c] - U.fvel[14][c]);
    if (t == 20) return 0.;
    if (t == 21) return 100.;
    int id = obj_id;
    if (id == -1) {
        id = min_phi_id(O, nobj, M, c);
        if (id == -1) return t == 12 ? 0. : 100.;
    }
    int ref = pick_map(M, c, id);
    if (ref < 0) return t == 12 ? 0. : 100.;
    bool prim = ref < M.Ng;
    uint32_t cc = M.c(ref);
    switch (t) {
        case 8: return map_phi(O, nobj, M, ref, 0);
        case 9: case 10: case 11: return M.x(ref, t - 9);
        case 12: return prim ? (double)(cc & OMASK) : 0.;
        case 13: return (double)((cc & LMASK) >> LDIG);
        case 14: return map_phi(O, nobj, M, ref, 1);
        default: return M.xpredr(ref, t - 15);
    }
}

// Global slice normal to dim at index point: out[a + b*d1] for the two
// remaining dimensions in increasing order (fluid_3d::slice and
// copy_slice_from_buf, fluid_3d_io.cc:98-255, 479-504).
__global__ void k_slice(Geo g, const ObjC *O, int nobj, Fluid U, Maps M, const double *r0, int t, int obj_id, int dim,
                        int point, int d1, int d2, double *out) {
    int a = blockIdx.x * blockDim.x + threadIdx.x, b = blockIdx.y;
    if (a >= d1 || b >= d2) return;
    int i, j, k;
    if (dim == 0) { i = point; j = a; k = b; }
    else if (dim == 1) { i = a; j = point; k = b; }
    else { i = a; j = b; k = point; }
    out[a + (size_t)b * d1] = slice_value(g, O, nobj, U, M, r0, t, obj_id, i, j, k);
}

// The eight checkpoint files of fluid_3d::write_chk_pt (fluid_3d_chk_pt.cc:
// 3-72): out[f][k][e] for f over vx vy vz p X Y Z OID, z-slice k < o and
// e < L = (m+1)(n+1).  The reference reuses one slice buffer, zeroed once,
// for all fields and slices and writes only its leading part: cell fields
// fill e < m*n (row stride m), the pressure fills e < fm*fn (row stride fm).
// So a cell field written before the pressure carries zeros past m*n, and
// one written after it carries the pressure's last slice (k = o-1) there.
__global__ void __launch_bounds__(256) k_chk_pack(Geo g, const ObjC *O, int nobj, Fluid U, Maps M, size_t L,
                                                   double *out) {
    int e = blockIdx.x * blockDim.x + threadIdx.x, k = blockIdx.y;
    if (e >= (int)L) return;
    size_t fs = (size_t)g.o * L, base = (size_t)k * L + e;
    int mn = g.m * g.n, fmn = g.fm * g.fn;
    double v[8];
    if (e < mn) {
        int c = IXp(g, e % g.m + 2, e / g.m + 2, k + 2);
        for (int a = 0; a < 3; a++) v[a] = U.vel[a][c];
        // X, Y, Z and OID of the object with the least phi (min_phi_id, then
        // the solid branch of copy_slice_to_buf), selected once per cell
        int id = min_phi_id(O, nobj, M, c);
        int ref = id < 0 ? -1 : pick_map(M, c, id);
        if (ref < 0) {
            v[4] = v[5] = v[6] = 100.;
            v[7] = 0.;
        } else {
            for (int a = 0; a < 3; a++) v[4 + a] = M.x(ref, a);
            v[7] = ref < M.Ng ? (double)(M.r.c[c] & OMASK) : 0.;
        }
    } else {
        v[0] = v[1] = v[2] = 0.;
        double tail = 0.;
        if (e < fmn) tail = U.p[IXp(g, e % g.fm + 2, e / g.fm + 2, g.o - 1 + 2)];
        v[4] = v[5] = v[6] = v[7] = tail;
    }
    v[3] = e < fmn ? U.p[IXp(g, e % g.fm + 2, e / g.fm + 2, k + 2)] : 0.;
    for (int f = 0; f < 8; f++) out[f * fs + base] = v[f];
}

// Checkpoint read-back (fluid_3d_chk_pt.cc:303-352 with init_from_slice
// :165-301 on one rank): cell (i,j,k<o) takes in[f][k][i + j*m]; the node
// (i<fm, j<fn, k<o) takes its pressure from in[3][k][i + j*m] -- the row
// stride of the cell fields, which is the reference reader's (the writer
// used fm, so the two differ when x is walled).  x and xpred both take X/Y/Z,
// c takes OID.
__global__ void k_chk_unpack(Geo g, Fluid U, RMap R, size_t L, const double *in) {
    int e = blockIdx.x * blockDim.x + threadIdx.x, k = blockIdx.y;
    int fmn = g.fm * g.fn;
    if (e >= fmn) return;
    size_t fs = (size_t)g.o * L, sl = (size_t)k * L;
    {
        int i = e % g.fm, j = e / g.fm;
        U.p[IXp(g, i + 2, j + 2, k + 2)] = in[3 * fs + sl + i + (size_t)j * g.m];
    }
    if (e < g.m * g.n) {
        int i = e % g.m, j = e / g.m;
        int c = IXp(g, i + 2, j + 2, k + 2);
        size_t s = sl + e;
        for (int a = 0; a < 3; a++) U.vel[a][c] = in[a * fs + s];
        for (int a = 0; a < 3; a++) {
            double x = in[(4 + a) * fs + s];
            R.x[a][c] = x;
            R.xpred[a][c] = x;
        }
        R.c[c] = (uint32_t)in[7 * fs + s];
    }
}

// The constructor state of field and ref_map (fields.hh:57-70, 326-357) on
// every padded cell, before a checkpoint is read into it.
__global__ void k_chk_reset(Geo g, Fluid U, RMap R) {
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= g.Ng) return;
    for (int a = 0; a < 3; a++) { U.vel[a][c] = 0; U.dvel[a][c] = 0; }
    U.p[c] = 0; U.dp[c] = 0; U.q[c] = 0;
    for (int a = 0; a < 18; a++) U.fvel[a][c] = 0;
    for (int a = 0; a < 9; a++) U.sigma[a][c] = 0;
    R.c[c] = 0;
    for (int a = 0; a < 3; a++) { R.x[a][c] = 0; R.xpred[a][c] = 0; R.detF[a][c] = 0; R.en[a][c] = 0; }
    for (int a = 0; a < 18; a++) R.fx[a][c] = 0;
    for (int a = 0; a < 9; a++) R.gphi[a][c] = -1.e200;
}

// Device scratch of the allocation-free slice (rmt3d_slice_noalloc): one
// fixed array per GPU, resident until the process ends.  Host code never
// passes the array's name to a kernel (on the host it is only a stand-in for
// the symbol); it asks hipGetSymbolAddress for this device's copy once, with
// that device current, and uses the returned address from then on.  Every
// call on one GPU shares its array, so calls that overlap on one GPU race.
static __device__ __align__(256) double scratch[SLICE_SCRATCH_MAX];

double *slice_scratch(std::string &err) {
    static std::mutex mu;
    static double *ptr[SLICE_SCRATCH_DEVICES] = {};
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= SLICE_SCRATCH_DEVICES) {
        (void)hipGetLastError();
        err = "slice_noalloc: no usable current device";
        return nullptr;
    }
    std::lock_guard<std::mutex> lk(mu);
    if (!ptr[dev]) {
        void *p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            (void)hipGetLastError();
            err = std::string("slice_noalloc: hipGetSymbolAddress failed: ") + hipGetErrorString(e);
            return nullptr;
        }
        ptr[dev] = (double *)p;
    }
    return ptr[dev];
}

void launch_slice(hipStream_t s, const Geo &g, const ObjC *O, int nobj, const Fluid &U, const Maps &M,
                  const double *r0, int t, int obj_id, int dim, int point, int d1, int d2, double *out) {
    dim3 grid((d1 + 255) / 256, d2);
    hipLaunchKernelGGL(k_slice, grid, dim3(256), 0, s, g, O, nobj, U, M, r0, t, obj_id, dim, point, d1, d2, out);
}

void launch_chk_pack(hipStream_t s, const Geo &g, const ObjC *O, int nobj, const Fluid &U, const Maps &M,
                     double *out, hipEvent_t start, hipEvent_t stop) {
    size_t L = (size_t)(g.m + 1) * (g.n + 1);
    dim3 grid((unsigned)((L + 255) / 256), g.o);
    hipExtLaunchKernelGGL(k_chk_pack, grid, dim3(256), 0, s, start, stop, 0, g, O, nobj, U, M, L, out);
}

void launch_chk_unpack(hipStream_t s, const Geo &g, const Fluid &U, const RMap &R, const double *in,
                       hipEvent_t start, hipEvent_t stop) {
    size_t L = (size_t)(g.m + 1) * (g.n + 1);
    int fmn = g.fm * g.fn;
    dim3 grid((fmn + 255) / 256, g.o);
    hipExtLaunchKernelGGL(k_chk_unpack, grid, dim3(256), 0, s, start, stop, 0, g, U, R, L, in);
}

void launch_chk_reset(hipStream_t s, const Geo &g, const Fluid &U, const RMap &R) {
    hipLaunchKernelGGL(k_chk_reset, dim3((g.Ng + 255) / 256), dim3(256), 0, s, g, U, R);
}

}  // namespace rmt
