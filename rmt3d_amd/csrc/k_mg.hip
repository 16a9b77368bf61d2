// Geometric multigrid kernels for the two projection problems, templated on
// the arithmetic type (double for the parity path, float for the FP32
// preconditioner of the mixed path).
//
// Operators restate simp_mac / simp_fem with stencil_set::set_partials
// (Utilities/mg_probs.hh:281-343, 458-530) in closed form: MAC 7-point with
// the missing wall face folded into the centre; FEM 27-point Q1 with wall
// entries halved per wall direction whose offset is zero.  Periodic
// dimensions wrap.  Smoothers: red-black (MAC) and four-colour by (j,k)
// parity (FEM: whole x-rows per colour) Gauss-Seidel, one launch per colour; inside a colour no two unknowns are
// neighbours, so the update order is free and the result is deterministic.
#include "dev.h"
#include "launch.h"

namespace rmt {

__device__ __forceinline__ int nbw(int i, int di, int d, int prd) {
    int j = i + di;
    if (j < 0) return prd ? j + d : -1;
    if (j >= d) return prd ? j - d : -1;
    return j;
}

struct Prd { int p[3]; };

template <typename T>
__device__ __forceinline__ T mac_off(const MGLevel<T> &L, const Prd &P, const T *x, int i, int j, int k, T &dg) {
    int D0 = L.d[0], D1 = L.d[1], D2 = L.d[2];
    T s = 0;
    int miss = 0, ii;
    ii = nbw(i, -1, D0, P.p[0]); if (ii < 0) miss++; else s += x[ii + D0 * (j + D1 * k)];
    ii = nbw(i, 1, D0, P.p[0]);  if (ii < 0) miss++; else s += x[ii + D0 * (j + D1 * k)];
    ii = nbw(j, -1, D1, P.p[1]); if (ii < 0) miss++; else s += x[i + D0 * (ii + D1 * k)];
    ii = nbw(j, 1, D1, P.p[1]);  if (ii < 0) miss++; else s += x[i + D0 * (ii + D1 * k)];
    ii = nbw(k, -1, D2, P.p[2]); if (ii < 0) miss++; else s += x[i + D0 * (j + D1 * ii)];
    ii = nbw(k, 1, D2, P.p[2]);  if (ii < 0) miss++; else s += x[i + D0 * (j + D1 * ii)];
    dg = T(-6.0 + miss);
    return s;
}

template <typename T>
__device__ __forceinline__ T fem_off(const MGLevel<T> &L, const Prd &P, const T *x, int i, int j, int k, T &dg) {
    int D0 = L.d[0], D1 = L.d[1], D2 = L.d[2];
    bool w0 = !P.p[0] && (i == 0 || i == D0 - 1);
    bool w1 = !P.p[1] && (j == 0 || j == D1 - 1);
    bool w2 = !P.p[2] && (k == 0 || k == D2 - 1);
    T s = 0;
    dg = 0;
    for (int ok = -1; ok <= 1; ok++) {
        int kk = nbw(k, ok, D2, P.p[2]);
        if (kk < 0) continue;
        for (int oj = -1; oj <= 1; oj++) {
            int jj = nbw(j, oj, D1, P.p[1]);
            if (jj < 0) continue;
            for (int oi = -1; oi <= 1; oi++) {
                int ii = nbw(i, oi, D0, P.p[0]);
                if (ii < 0) continue;
                int nz = (oi != 0) + (oj != 0) + (ok != 0);
                if (nz == 1) continue;
                T w = (nz == 0) ? T(32) : (nz == 2) ? T(-2) : T(-1);
                T f = 1;
                if (w0 && oi == 0) f *= 2;
                if (w1 && oj == 0) f *= 2;
                if (w2 && ok == 0) f *= 2;
                w /= f;
                if (nz == 0) dg = w;
                else s += w * x[ii + D0 * (jj + D1 * kk)];
            }
        }
    }
    return s;
}

template <typename T>
__global__ void k_mac_smooth(MGLevel<T> L, Prd P, int color) {
    int j = blockIdx.y, k = blockIdx.z;
    int i = 2 * (blockIdx.x * blockDim.x + threadIdx.x) + ((j + k + color) & 1);
    if (i >= L.d[0]) return;
    int c = i + L.d[0] * (j + L.d[1] * k);
    T dg;
    T s = mac_off(L, P, L.x, i, j, k, dg);
    L.x[c] = (L.b[c] - s) / dg;
}

// interior node (no wall, no wrap): the 20 non-zero neighbours in the order
// of fem_off's (k, j, i) loop, so the sum is bit-identical to the general path
template <typename T>
__device__ __forceinline__ T fem_off_interior(const T *x, int c, int D0, int D01) {
    T s = 0;
    const T *p = x + c - D01;
    s += T(-1) * p[-D0 - 1]; s += T(-2) * p[-D0]; s += T(-1) * p[-D0 + 1];
    s += T(-2) * p[-1];      s += T(-2) * p[1];
    s += T(-1) * p[D0 - 1];  s += T(-2) * p[D0];  s += T(-1) * p[D0 + 1];
    p = x + c;
    s += T(-2) * p[-D0 - 1]; s += T(-2) * p[-D0 + 1];
    s += T(-2) * p[D0 - 1];  s += T(-2) * p[D0 + 1];
    p = x + c + D01;
    s += T(-1) * p[-D0 - 1]; s += T(-2) * p[-D0]; s += T(-1) * p[-D0 + 1];
    s += T(-2) * p[-1];      s += T(-2) * p[1];
    s += T(-1) * p[D0 - 1];  s += T(-2) * p[D0];  s += T(-1) * p[D0 + 1];
    return s;
}

// four colours by (j,k) parity: one colour is a set of whole x-rows
template <typename T>
__global__ void k_fem_smooth(MGLevel<T> L, Prd P, int color) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    int j = 2 * blockIdx.y + (color & 1);
    int k = 2 * blockIdx.z + ((color >> 1) & 1);
    int D0 = L.d[0], D1 = L.d[1], D2 = L.d[2];
    if (i >= D0 || j >= D1 || k >= D2) return;
    int c = i + D0 * (j + D1 * k);
    T s, dg;
    if (i >= 1 && i <= D0 - 2 && j >= 1 && j <= D1 - 2 && k >= 1 && k <= D2 - 2) {
        s = fem_off_interior(L.x, c, D0, D0 * D1);
        dg = T(32);
    } else {
        s = fem_off(L, P, L.x, i, j, k, dg);
    }
    L.x[c] = (L.b[c] - s) / dg;
}

template <typename T>
__device__ __forceinline__ T apply_off(const MGLevel<T> &L, const Prd &P, int fem, int i, int j, int k, int c, T &dg) {
    int D0 = L.d[0], D1 = L.d[1], D2 = L.d[2];
    if (fem && i >= 1 && i <= D0 - 2 && j >= 1 && j <= D1 - 2 && k >= 1 && k <= D2 - 2) {
        dg = T(32);
        return fem_off_interior(L.x, c, D0, D0 * D1);
    }
    return fem ? fem_off(L, P, L.x, i, j, k, dg) : mac_off(L, P, L.x, i, j, k, dg);
}

template <typename T>
__global__ void k_resid(MGLevel<T> L, Prd P, int fem) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    int j = blockIdx.y, k = blockIdx.z;
    if (i >= L.d[0]) return;
    int c = i + L.d[0] * (j + L.d[1] * k);
    T dg;
    T s = apply_off(L, P, fem, i, j, k, c, dg);
    L.r[c] = L.b[c] - (s + dg * L.x[c]);
}

// FP64 residual and its max norm: grid-stride blocks of 256, one atomic per
// block (non-negative doubles order as their bit patterns; a NaN compares
// above every finite value)
__global__ void k_resid64_max(MGLevel<double> L, Prd P, int fem, unsigned long long *maxbits) {
    __shared__ unsigned long long sh[4];
    unsigned long long v = 0;
    int D0 = L.d[0], D1 = L.d[1];
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < L.nn; c += gridDim.x * blockDim.x) {
        int i = c % D0, j = (c / D0) % D1, k = c / (D0 * D1);
        double dg;
        double s = apply_off(L, P, fem, i, j, k, c, dg);
        double r = L.b[c] - (s + dg * L.x[c]);
        L.r[c] = r;
        unsigned long long b = (unsigned long long)__double_as_longlong(fabs(r));
        v = b > v ? b : v;
    }
    for (int off = 32; off > 0; off >>= 1) {
        unsigned long long o = __shfl_xor(v, off, 64);
        v = o > v ? o : v;
    }
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long m = sh[0];
        for (int w = 1; w < (int)(blockDim.x >> 6); w++) m = sh[w] > m ? sh[w] : m;
        atomicMax(maxbits, m);
    }
}

template <typename T>
__global__ void k_restrict(MGLevel<T> F, MGLevel<T> C, Prd P, int fem) {
    int I = blockIdx.x * blockDim.x + threadIdx.x;
    int J = blockIdx.y, K = blockIdx.z;
    if (I >= C.d[0]) return;
    T s = 0;
    if (!fem) {
        for (int c = 0; c < 2; c++)
            for (int b = 0; b < 2; b++)
                for (int a = 0; a < 2; a++)
                    s += F.r[(2 * I + a) + F.d[0] * ((2 * J + b) + F.d[1] * (2 * K + c))];
        s *= T(0.5);
    } else {
        for (int dc = -1; dc <= 1; dc++) {
            int kk = nbw(2 * K, dc, F.d[2], P.p[2]);
            if (kk < 0) continue;
            for (int db = -1; db <= 1; db++) {
                int jj = nbw(2 * J, db, F.d[1], P.p[1]);
                if (jj < 0) continue;
                for (int da = -1; da <= 1; da++) {
                    int ii = nbw(2 * I, da, F.d[0], P.p[0]);
                    if (ii < 0) continue;
                    T w = (da ? T(0.5) : T(1.0)) * (db ? T(0.5) : T(1.0)) * (dc ? T(0.5) : T(1.0));
                    s += w * F.r[ii + F.d[0] * (jj + F.d[1] * kk)];
                }
            }
        }
        s *= T(0.5);
    }
    int ci = I + C.d[0] * (J + C.d[1] * K);
    C.b[ci] = s;
    C.x[ci] = 0;
}

template <typename T>
__device__ __forceinline__ void cc_w(int f, int D, int prd, int &i0, int &i1, T &w0, T &w1) {
    int I = f >> 1, a = f & 1;
    int o = a ? I + 1 : I - 1;
    int oo = nbw(I, o - I, D, prd);
    i0 = I; w0 = T(0.75);
    i1 = (oo < 0) ? I : oo; w1 = T(0.25);
}
template <typename T>
__device__ __forceinline__ void vx_w(int f, int D, int prd, int &i0, int &i1, T &w0, T &w1) {
    int I = f >> 1;
    if ((f & 1) == 0) { i0 = I; i1 = I; w0 = T(1.0); w1 = T(0.0); return; }
    int J = I + 1;
    if (J >= D) J = prd ? J - D : D - 1;
    i0 = I; i1 = J; w0 = T(0.5); w1 = T(0.5);
}

template <typename T>
__global__ void k_prolong(MGLevel<T> F, MGLevel<T> C, Prd P, int fem) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    int j = blockIdx.y, k = blockIdx.z;
    if (i >= F.d[0]) return;
    int i0, i1, j0, j1, k0, k1;
    T wi0, wi1, wj0, wj1, wk0, wk1;
    if (fem) {
        vx_w(k, C.d[2], P.p[2], k0, k1, wk0, wk1);
        vx_w(j, C.d[1], P.p[1], j0, j1, wj0, wj1);
        vx_w(i, C.d[0], P.p[0], i0, i1, wi0, wi1);
    } else {
        cc_w(k, C.d[2], P.p[2], k0, k1, wk0, wk1);
        cc_w(j, C.d[1], P.p[1], j0, j1, wj0, wj1);
        cc_w(i, C.d[0], P.p[0], i0, i1, wi0, wi1);
    }
    const T *e = C.x;
    int D0 = C.d[0], D1 = C.d[1];
#define E(a, b, c) e[(a) + D0 * ((b) + D1 * (c))]
    T v = wk0 * (wj0 * (wi0 * E(i0, j0, k0) + wi1 * E(i1, j0, k0)) + wj1 * (wi0 * E(i0, j1, k0) + wi1 * E(i1, j1, k0))) +
          wk1 * (wj0 * (wi0 * E(i0, j0, k1) + wi1 * E(i1, j0, k1)) + wj1 * (wi0 * E(i0, j1, k1) + wi1 * E(i1, j1, k1)));
#undef E
    F.x[i + F.d[0] * (j + F.d[1] * k)] += v;
}

// coarsest level: coloured Gauss-Seidel sweeps inside one workgroup, on an
// LDS copy when it fits
template <typename T>
__global__ void k_coarse(MGLevel<T> L, Prd P, int fem, int sweeps, int use_lds) {
    extern __shared__ unsigned char smem[];
    int n = L.nn;
    int ncol = fem ? 4 : 2;
    MGLevel<T> W = L;
    if (use_lds) {
        T *xs = (T *)smem, *bs = xs + n;
        for (int c = threadIdx.x; c < n; c += blockDim.x) { xs[c] = L.x[c]; bs[c] = L.b[c]; }
        W.x = xs;
        W.b = bs;
        __syncthreads();
    }
    for (int sw = 0; sw < sweeps; sw++) {
        for (int col = 0; col < ncol; col++) {
            for (int c = threadIdx.x; c < n; c += blockDim.x) {
                int i = c % W.d[0], j = (c / W.d[0]) % W.d[1], k = c / (W.d[0] * W.d[1]);
                int my = fem ? ((j & 1) + 2 * (k & 1)) : ((i + j + k) & 1);
                if (my != col) continue;
                T dg;
                T s = fem ? fem_off(W, P, W.x, i, j, k, dg) : mac_off(W, P, W.x, i, j, k, dg);
                W.x[c] = (W.b[c] - s) / dg;
            }
            __syncthreads();
        }
    }
    if (use_lds)
        for (int c = threadIdx.x; c < n; c += blockDim.x) L.x[c] = W.x[c];
}

__global__ void k_cast_d2f(const double *a, float *b, int n) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) b[t] = (float)a[t];
}
__global__ void k_axpy_f2d(double *x, const float *e, int n) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) x[t] += (double)e[t];
}
__global__ void k_zero_f(float *a, int n) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) a[t] = 0.f;
}

static inline Prd mkprd(const int p[3]) { Prd P; P.p[0] = p[0]; P.p[1] = p[1]; P.p[2] = p[2]; return P; }

template <typename T>
void launch_mg_smooth(hipStream_t s, int fem, const MGLevel<T> &L, const int prd[3], int color) {
    Prd P = mkprd(prd);
    if (!fem) {
        int hx = (L.d[0] + 1) / 2;
        dim3 gr((hx + 63) / 64, L.d[1], L.d[2]);
        hipLaunchKernelGGL(k_mac_smooth<T>, gr, dim3(64), 0, s, L, P, color);
    } else {
        int hy = (L.d[1] + 1) / 2, hz = (L.d[2] + 1) / 2;
        dim3 gr((L.d[0] + 63) / 64, hy, hz);
        hipLaunchKernelGGL(k_fem_smooth<T>, gr, dim3(64), 0, s, L, P, color);
    }
}
template <typename T>
void launch_mg_resid(hipStream_t s, int fem, const MGLevel<T> &L, const int prd[3]) {
    dim3 gr((L.d[0] + 63) / 64, L.d[1], L.d[2]);
    hipLaunchKernelGGL(k_resid<T>, gr, dim3(64), 0, s, L, mkprd(prd), fem);
}
void launch_mg_resid64_max(hipStream_t s, int fem, const MGLevel<double> &L, const int prd[3],
                           unsigned long long *maxbits) {
    int nb = (L.nn + 255) / 256;
    if (nb > 1024) nb = 1024;
    hipLaunchKernelGGL(k_resid64_max, dim3(nb), dim3(256), 0, s, L, mkprd(prd), fem, maxbits);
}
template <typename T>
void launch_mg_restrict(hipStream_t s, int fem, const MGLevel<T> &F, const MGLevel<T> &C, const int prd[3]) {
    dim3 gr((C.d[0] + 63) / 64, C.d[1], C.d[2]);
    hipLaunchKernelGGL(k_restrict<T>, gr, dim3(64), 0, s, F, C, mkprd(prd), fem);
}
template <typename T>
void launch_mg_prolong(hipStream_t s, int fem, const MGLevel<T> &F, const MGLevel<T> &C, const int prd[3]) {
    dim3 gr((F.d[0] + 63) / 64, F.d[1], F.d[2]);
    hipLaunchKernelGGL(k_prolong<T>, gr, dim3(64), 0, s, F, C, mkprd(prd), fem);
}
template <typename T>
void launch_mg_coarse(hipStream_t s, int fem, const MGLevel<T> &L, const int prd[3], int sweeps) {
    size_t lds = 2 * sizeof(T) * (size_t)L.nn;
    int use = lds <= 48 * 1024;
    hipLaunchKernelGGL(k_coarse<T>, dim3(1), dim3(256), use ? lds : 0, s, L, mkprd(prd), fem, sweeps, use);
}
void launch_cast_d2f(hipStream_t s, const double *a, float *b, int n) {
    hipLaunchKernelGGL(k_cast_d2f, dim3((n + 255) / 256), dim3(256), 0, s, a, b, n);
}
void launch_axpy_f2d(hipStream_t s, double *x, const float *e, int n) {
    hipLaunchKernelGGL(k_axpy_f2d, dim3((n + 255) / 256), dim3(256), 0, s, x, e, n);
}
void launch_zero_f(hipStream_t s, float *a, int n) {
    hipLaunchKernelGGL(k_zero_f, dim3((n + 255) / 256), dim3(256), 0, s, a, n);
}

#define INST(T)                                                                                              \
    template void launch_mg_smooth<T>(hipStream_t, int, const MGLevel<T> &, const int *, int);              \
    template void launch_mg_resid<T>(hipStream_t, int, const MGLevel<T> &, const int *);                    \
    template void launch_mg_restrict<T>(hipStream_t, int, const MGLevel<T> &, const MGLevel<T> &, const int *); \
    template void launch_mg_prolong<T>(hipStream_t, int, const MGLevel<T> &, const MGLevel<T> &, const int *);  \
    template void launch_mg_coarse<T>(hipStream_t, int, const MGLevel<T> &, const int *, int);
INST(double)
INST(float)

}  // namespace rmt
