// Geometric multigrid kernels for the two projection problems, templated on
// the arithmetic type (double for the parity path, float for the FP32
// preconditioner of the mixed path).
//
// Operators restate simp_mac / simp_fem with stencil_set::set_partials
// (Utilities/mg_probs.hh:281-343, 458-530) in closed form: MAC 7-point with
// the missing wall face folded into the centre; FEM 27-point Q1 with wall
// entries halved per wall direction whose offset is zero.  Periodic
// dimensions wrap.  Smoothers: red-black SOR (MAC, omega 1.25) and
// four-colour Gauss-Seidel by (j,k) parity (FEM: whole x-rows per colour),
// one launch per colour; inside a colour no two unknowns are
// neighbours, so the update order is free and the result is deterministic.
#include <hip/hip_ext.h>

#include <vector>

#include "dev.h"
#include "launch.h"
#include <type_traits>

namespace rmt {

__device__ __forceinline__ int nbw(int i, int di, int d, int prd) {
    int j = i + di;
    if (j < 0) return prd ? j + d : -1;
    if (j >= d) return prd ? j - d : -1;
    return j;
}

struct Prd { int p[3]; };

// A missing (wall) neighbour contributes +0 at its place in the sum instead
// of being skipped: the running sum starts at +0 and is never -0, so adding
// a zero leaves it bit-identical, and every lane of a wavefront executes the
// same instructions.
template <typename T>
__device__ __forceinline__ T mac_off(const MGLevel<T> &L, const Prd &P, const T *x, int i, int j, int k, T &dg) {
    int D0 = L.d[0], D1 = L.d[1], D2 = L.d[2];
    int c = i + D0 * (j + D1 * k);
    T s = 0;
    int miss = 0, ii;
    bool v;
#define MAC_TERM(expr, idx)                       \
    ii = (expr);                                  \
    v = ii >= 0;                                  \
    miss += !v;                                   \
    {                                             \
        T xv = x[v ? (idx) : c];                  \
        s += v ? xv : T(0);                       \
    }
    MAC_TERM(nbw(i, -1, D0, P.p[0]), ii + D0 * (j + D1 * k))
    MAC_TERM(nbw(i, 1, D0, P.p[0]), ii + D0 * (j + D1 * k))
    MAC_TERM(nbw(j, -1, D1, P.p[1]), i + D0 * (ii + D1 * k))
    MAC_TERM(nbw(j, 1, D1, P.p[1]), i + D0 * (ii + D1 * k))
    MAC_TERM(nbw(k, -1, D2, P.p[2]), i + D0 * (j + D1 * ii))
    MAC_TERM(nbw(k, 1, D2, P.p[2]), i + D0 * (j + D1 * ii))
#undef MAC_TERM
    dg = T(-6.0 + miss);
    return s;
}

// The wall halving of a weight is a multiplication by 1/2 per wall dimension
// with zero offset: exact, so it equals the division by 2, 4 or 8 of the
// closed form.  A missing neighbour enters with weight 0 at a valid address
// (see mac_off).  rdg returns 1/dg, exact because dg is 32 times a power of
// two.
template <typename T>
__device__ __forceinline__ T fem_off(const MGLevel<T> &L, const Prd &P, const T *x, int i, int j, int k, T &dg,
                                     T &rdg) {
    int D0 = L.d[0], D1 = L.d[1], D2 = L.d[2];
    int c = i + D0 * (j + D1 * k);
    int ia[3] = {nbw(i, -1, D0, P.p[0]), i, nbw(i, 1, D0, P.p[0])};
    int ja[3] = {nbw(j, -1, D1, P.p[1]), j, nbw(j, 1, D1, P.p[1])};
    int ka[3] = {nbw(k, -1, D2, P.p[2]), k, nbw(k, 1, D2, P.p[2])};
    bool w0 = !P.p[0] && (i == 0 || i == D0 - 1);
    bool w1 = !P.p[1] && (j == 0 || j == D1 - 1);
    bool w2 = !P.p[2] && (k == 0 || k == D2 - 1);
    T sx = w0 ? T(0.5) : T(1), sy = w1 ? T(0.5) : T(1), sz = w2 ? T(0.5) : T(1);
    T s = 0;
#pragma unroll
    for (int ok = 0; ok < 3; ok++) {
#pragma unroll
        for (int oj = 0; oj < 3; oj++) {
            int row = D0 * (ja[oj] + D1 * ka[ok]);
            bool vjk = ja[oj] >= 0 && ka[ok] >= 0;
#pragma unroll
            for (int oi = 0; oi < 3; oi++) {
                const int nz = (oi != 1) + (oj != 1) + (ok != 1);
                if (nz < 2) continue;               // centre and face terms
                bool v = vjk && ia[oi] >= 0;
                T w = (nz == 2) ? T(-2) : T(-1);
                if (oi == 1) w *= sx;
                if (oj == 1) w *= sy;
                if (ok == 1) w *= sz;
                T xv = x[v ? ia[oi] + row : c];
                s += (v ? w : T(0)) * xv;
            }
        }
    }
    dg = T(32) * sx * sy * sz;
    rdg = T(0.03125);
    if (w0) rdg *= T(2);
    if (w1) rdg *= T(2);
    if (w2) rdg *= T(2);
    return s;
}
template <typename T>
__device__ __forceinline__ T fem_off(const MGLevel<T> &L, const Prd &P, const T *x, int i, int j, int k, T &dg) {
    T rdg;
    return fem_off(L, P, x, i, j, k, dg, rdg);
}

// quotient of non-negative integers (a < 2^25, d >= 2) from a float
// reciprocal, corrected by the exact remainder (the estimate is off by at
// most 2 in that range)
__device__ __forceinline__ int qdiv(int a, int d, float rd) {
    int q = (int)((float)a * rd);
    int r = a - q * d;
    while (r < 0) { q--; r += d; }
    while (r >= d) { q++; r -= d; }
    return q;
}

// red-black SOR: x <- (1 - w) x + w (b - s) / dg, w = 1.25
constexpr double MAC_OMEGA = 1.25;

template <typename T>
__device__ __forceinline__ T mac_smooth_at(const MGLevel<T> &L, const Prd &P, int i, int j, int k) {
    int c = i + L.d[0] * (j + L.d[1] * k);
    T dg;
    T s = mac_off(L, P, L.x, i, j, k, dg);
    T xn = (L.b[c] - s) / dg;
    T v = T(1 - MAC_OMEGA) * L.x[c] + T(MAC_OMEGA) * xn;
    L.x[c] = v;
    return v;
}

// one thread per node of the colour: t -> (ih, j, k), i = 2 ih + parity.
// x64 (the final sweep of the mixed path's FP32 V-cycle on the finest
// level): this pass is the colour's last update of the correction, which is
// added to the FP64 solution at once, x64 += (double) x.
template <typename T>
__global__ void k_mac_smooth(MGLevel<T> L, Prd P, int color, int hx, float rhx, float rd1, double *x64) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    int row = qdiv(t, hx, rhx);
    int ih = t - row * hx;
    int k = qdiv(row, L.d[1], rd1);
    int j = row - k * L.d[1];
    if (k >= L.d[2]) return;
    int i = 2 * ih + ((j + k + color) & 1);
    if (i >= L.d[0]) return;
    const int c = i + L.d[0] * (j + L.d[1] * k);
    T v;
    if ((L.d[0] & 1) == 0) {
        // Even rows: the node and its row partner c ^ 1 (the other colour,
        // which this pass leaves unchanged) are written as one aligned pair,
        // so that the stores fill whole memory sectors instead of every
        // other word of them.  Only this work-item writes the partner, and
        // with the value it already holds.
        T dg;
        T sm = mac_off(L, P, L.x, i, j, k, dg);
        T xn = (L.b[c] - sm) / dg;
        v = T(1 - MAC_OMEGA) * L.x[c] + T(MAC_OMEGA) * xn;
        const T o = L.x[c ^ 1];
        typedef T pair_t __attribute__((ext_vector_type(2)));
        pair_t pv;
        pv[0] = (c & 1) ? o : v;
        pv[1] = (c & 1) ? v : o;
        *reinterpret_cast<pair_t *>(L.x + (c & ~1)) = pv;
        if (x64) {
            // the FP64 solution the same way: the pair read and written back
            // whole, only this node's element changed
            typedef double pair64_t __attribute__((ext_vector_type(2)));
            pair64_t *xp = reinterpret_cast<pair64_t *>(x64 + (c & ~1));
            pair64_t q = *xp;
            if (c & 1) q[1] += (double)v;
            else q[0] += (double)v;
            *xp = q;
        }
        return;
    } else {
        v = mac_smooth_at(L, P, i, j, k);
    }
    if (x64) x64[c] += (double)v;
}

template <typename T>
__device__ __forceinline__ T fem_smooth_at(const MGLevel<T> &L, const Prd &P, int i, int j, int k) {
    int c = i + L.d[0] * (j + L.d[1] * k);
    T dg, rdg;
    T s = fem_off(L, P, L.x, i, j, k, dg, rdg);
    T v = (L.b[c] - s) * rdg;
    L.x[c] = v;
    return v;
}

// Interior node (no wall, no wrap): the 20 non-zero neighbours at fixed
// offsets, in fem_off's (k, j, i) order with its weights, so the sum is
// bit-identical to the general path.
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

// Work-item order of one FEM colour (rows j = jc mod 2, k = kc mod 2):
//   [0, NI)          interior nodes of interior rows (fast path)
//   [NI, NI+NA)      the two end nodes (i = 0, D0-1) of interior rows
//   [NI+NA, +NB)     every node of the colour's boundary rows
// so that a wavefront runs either the fast path or the general one.
struct FemColour {
    int plain;                // 1: no interior rows; t -> row t / D0 of all colour rows
    int jI0, njI, kI0, nkI;   // interior rows: j = jI0 + 2a (a < njI), k = kI0 + 2b
    int hyc;                  // rows of the colour per k: j = jc + 2a, a < hyc
    int kc;
    int nks, ks0, ks1;        // boundary k values of the colour
    int njs, js0, js1;        // boundary j values of the colour
    int NI, NA, NB;
    // float reciprocals of the divisors of the work-item -> node map
    float rD0, rw, rnjI, rhyc, rnkI;
};

template <typename T>
__device__ __forceinline__ bool fem_colour_node(const MGLevel<T> &L, const FemColour &F, int jc, int t, int &i, int &j,
                                                int &k) {
    int D0 = L.d[0];
    if (F.plain) {
        int row = qdiv(t, D0, F.rD0);
        i = t - row * D0;
        int b = qdiv(row, F.hyc, F.rhyc);
        j = jc + 2 * (row - b * F.hyc);
        k = F.kc + 2 * b;
        return false;
    }
    if (t < F.NI) {
        int w = D0 - 2;
        int row = qdiv(t, w, F.rw);
        i = 1 + (t - row * w);
        int b = qdiv(row, F.njI, F.rnjI);
        j = F.jI0 + 2 * (row - b * F.njI);
        k = F.kI0 + 2 * b;
        return true;
    }
    t -= F.NI;
    if (t < F.NA) {
        int row = t >> 1;
        i = (t & 1) ? D0 - 1 : 0;
        int b = qdiv(row, F.njI, F.rnjI);
        j = F.jI0 + 2 * (row - b * F.njI);
        k = F.kI0 + 2 * b;
        return false;
    }
    t -= F.NA;
    int rb = qdiv(t, D0, F.rD0);
    i = t - rb * D0;
    int n1 = F.nks * F.hyc;
    if (rb < n1) {
        int a = qdiv(rb, F.hyc, F.rhyc);
        k = a ? F.ks1 : F.ks0;
        j = jc + 2 * (rb - a * F.hyc);
    } else {
        rb -= n1;
        int a = qdiv(rb, F.nkI, F.rnkI);
        j = a ? F.js1 : F.js0;
        k = F.kI0 + 2 * (rb - a * F.nkI);
    }
    return false;
}

// The launch issues the boundary-capable work-items (end nodes and boundary
// rows, several times the cost of an interior node) first and the interior
// after them, so that the expensive workgroups start early instead of
// forming the tail of the launch: work-item t takes the node of rank
// (t + NI) mod n of fem_colour_node's order.
// x64: as in k_mac_smooth, the final sweep adds each node's correction to
// the FP64 solution
template <typename T>
__global__ void k_fem_smooth(MGLevel<T> L, Prd P, int color, FemColour F, double *x64) {
    int n = F.NI + F.NA + F.NB;
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    t += F.NI;
    if (t >= n) t -= n;
    int i, j, k;
    bool interior = fem_colour_node(L, F, color & 1, t, i, j, k);
    int c = i + L.d[0] * (j + L.d[1] * k);
    T v;
    if (interior) {
        T s = fem_off_interior(L.x, c, L.d[0], L.d[0] * L.d[1]);
        v = (L.b[c] - s) * T(0.03125);
        L.x[c] = v;
    } else {
        v = fem_smooth_at(L, P, i, j, k);
    }
    if (x64) x64[c] += (double)v;
}

// host side: the work-item layout of one colour of a level
static __host__ __device__ FemColour fem_colour(const int d[3], int color) {
    FemColour F{};
    int jc = color & 1, kc = (color >> 1) & 1;
    int D0 = d[0], D1 = d[1], D2 = d[2];
    F.kc = kc;
    F.hyc = (D1 - jc + 1) / 2;
    int hzc = (D2 - kc + 1) / 2;
    // interior = no wall and no wrap in that dimension
    F.jI0 = jc ? 1 : 2;
    F.njI = (F.jI0 <= D1 - 2) ? (D1 - 2 - F.jI0) / 2 + 1 : 0;
    F.kI0 = kc ? 1 : 2;
    F.nkI = (F.kI0 <= D2 - 2) ? (D2 - 2 - F.kI0) / 2 + 1 : 0;
    F.rD0 = 1.0f / D0;
    F.rw = 1.0f / (D0 > 2 ? D0 - 2 : 1);
    F.rnjI = 1.0f / (F.njI > 0 ? F.njI : 1);
    F.rhyc = 1.0f / (F.hyc > 0 ? F.hyc : 1);
    F.rnkI = 1.0f / (F.nkI > 0 ? F.nkI : 1);
    if (D0 < 3 || F.njI == 0 || F.nkI == 0) {
        F.plain = 1;
        F.NB = D0 * F.hyc * hzc;
        return F;
    }
    if (kc == 0) F.ks0 = 0, F.nks = 1;
    if ((D2 - 1) % 2 == kc) {
        if (F.nks) F.ks1 = D2 - 1;
        else F.ks0 = D2 - 1;
        F.nks++;
    }
    if (jc == 0) F.js0 = 0, F.njs = 1;
    if ((D1 - 1) % 2 == jc) {
        if (F.njs) F.js1 = D1 - 1;
        else F.js0 = D1 - 1;
        F.njs++;
    }
    F.NI = (D0 - 2) * F.njI * F.nkI;
    F.NA = 2 * F.njI * F.nkI;
    F.NB = (F.nks * F.hyc + F.njs * F.nkI) * D0;
    return F;
}

template <typename T>
__device__ __forceinline__ T apply_off(const MGLevel<T> &L, const Prd &P, int fem, int i, int j, int k, int c, T &dg) {
    (void)c;
    return fem ? fem_off(L, P, L.x, i, j, k, dg) : mac_off(L, P, L.x, i, j, k, dg);
}

template <typename T>
__device__ __forceinline__ void resid_at(const MGLevel<T> &L, const Prd &P, int fem, int i, int j, int k) {
    int c = i + L.d[0] * (j + L.d[1] * k);
    T dg;
    T s = apply_off(L, P, fem, i, j, k, c, dg);
    L.r[c] = L.b[c] - (s + dg * L.x[c]);
}

// one thread per node of a level, t -> (i, j, k) with float reciprocals of
// the two fastest dimensions
#define FLAT_IJK(D0, D1, r0, r1)              \
    int t = blockIdx.x * blockDim.x + threadIdx.x; \
    int row = qdiv(t, D0, r0);                 \
    int i = t - row * (D0);                    \
    int k = qdiv(row, D1, r1);                 \
    int j = row - k * (D1);

// Node order of a residual pass.  MAC levels, and FEM levels with an extent
// below 4, are flat (x fastest).  Other FEM levels take the interior box
// (1..D-2 in every dimension: no wall, no wrap) first, flat, so that its
// nodes take the fast path together and neighbouring rows are read while
// cached, then the boundary: the planes k = 0 and k = D2-1, then per inner
// k the rows j = 0 and j = D1-1 and the two end nodes of every inner row.
struct NodeMap {
    int fem, plain, total;
    int D0, D1, D2;
    float r0, r1;            // flat: reciprocals of D0, D1
    int NI, wi, wj;          // interior count and extents in x and y
    float rwi, rwj;
    int S;                   // boundary nodes per inner k: 2 D0 + 2 (D1 - 2)
    float rS;
};

static __host__ __device__ NodeMap node_map(const int d[3], int fem) {
    NodeMap M{};
    M.fem = fem;
    M.D0 = d[0];
    M.D1 = d[1];
    M.D2 = d[2];
    M.total = d[0] * d[1] * d[2];
    M.r0 = 1.0f / d[0];
    M.r1 = 1.0f / d[1];
    M.plain = !fem || d[0] < 4 || d[1] < 4 || d[2] < 4;
    if (M.plain) return M;
    M.wi = d[0] - 2;
    M.wj = d[1] - 2;
    M.NI = M.wi * M.wj * (d[2] - 2);
    M.rwi = 1.0f / M.wi;
    M.rwj = 1.0f / M.wj;
    M.S = 2 * d[0] + 2 * (d[1] - 2);
    M.rS = 1.0f / M.S;
    return M;
}

// returns true for an interior FEM node.  Ranks are rotated by NI, so
// that the boundary nodes (the general stencil) come first in a launch and
// the interior box after them
template <typename T>
__device__ __forceinline__ bool map_node(const MGLevel<T> &L, const NodeMap &M, int t, int &i, int &j, int &k) {
    const int D0 = M.D0, D1 = M.D1;
    if (!M.plain) {
        t += M.NI;
        if (t >= M.total) t -= M.total;
    }
    if (M.plain) {
        int row = qdiv(t, D0, M.r0);
        i = t - row * D0;
        k = qdiv(row, D1, M.r1);
        j = row - k * D1;
        return false;
    }
    if (t < M.NI) {
        int row = qdiv(t, M.wi, M.rwi);
        i = 1 + (t - row * M.wi);
        int kk = qdiv(row, M.wj, M.rwj);
        j = 1 + (row - kk * M.wj);
        k = 1 + kk;
        return true;
    }
    t -= M.NI;
    const int plane = D0 * D1;
    if (t < 2 * plane) {
        int hi = t >= plane;
        int r = t - hi * plane;
        k = hi ? M.D2 - 1 : 0;
        j = qdiv(r, D0, M.r0);
        i = r - j * D0;
        return false;
    }
    t -= 2 * plane;
    int kk = qdiv(t, M.S, M.rS);
    k = 1 + kk;
    int r = t - kk * M.S;
    if (r < 2 * D0) {
        int hi = r >= D0;
        j = hi ? D1 - 1 : 0;
        i = r - hi * D0;
    } else {
        r -= 2 * D0;
        j = 1 + (r >> 1);
        i = (r & 1) ? D0 - 1 : 0;
    }
    return false;
}

// r = b - A x at one node; the interior FEM form is fem_off's sum
template <typename T>
__device__ __forceinline__ T resid_value(const MGLevel<T> &L, const Prd &P, int fem, bool interior, int i, int j, int k,
                                         int c) {
    if (interior) {
        T s = fem_off_interior(L.x, c, L.d[0], L.d[0] * L.d[1]);
        return L.b[c] - (s + T(32) * L.x[c]);
    }
    T dg;
    T s = apply_off(L, P, fem, i, j, k, c, dg);
    return L.b[c] - (s + dg * L.x[c]);
}

// FEM residual with the interior marched along z.  Work-items [0, NB)
// take the boundary nodes one each (the general stencil, map_node's
// order); the rest take a column segment (i, j, k0 .. k0 + RES_KC - 1) of
// the interior box each and walk it in k, holding the 3 x 3 neighbourhoods
// of planes k - 1, k and k + 1 in registers: a node costs nine loads of x
// (the new plane) instead of twenty-one.  The terms are summed in
// fem_off_interior's order, so every residual is the same bit for bit.
constexpr int RES_KC = 8;

struct MarchMap {
    NodeMap M;
    int NB;      // boundary nodes
    int nkc;     // z segments of the interior box
    int items;   // NB + wi * wj * nkc
};

static MarchMap march_map(const int d[3]) {
    MarchMap A;
    A.M = node_map(d, 1);
    A.NB = A.M.total - A.M.NI;
    A.nkc = (d[2] - 2 + RES_KC - 1) / RES_KC;
    A.items = A.NB + A.M.wi * A.M.wj * A.nkc;
    return A;
}

// calls emit(c, j, k, r) for every node of work-item t
template <typename T, typename E>
__device__ __forceinline__ void resid_march(const MGLevel<T> &L, const Prd &P, const MarchMap &A, int t, E emit) {
    if (t < A.NB) {
        int i, j, k;
        map_node(L, A.M, t, i, j, k);
        int c = i + L.d[0] * (j + L.d[1] * k);
        emit(c, j, k, resid_value(L, P, 1, false, i, j, k, c));
        return;
    }
    const NodeMap &M = A.M;
    int u = t - A.NB;
    int row = qdiv(u, M.wi, M.rwi);
    const int i = 1 + (u - row * M.wi);
    int kc = qdiv(row, M.wj, M.rwj);
    const int j = 1 + (row - kc * M.wj);
    const int k0 = 1 + kc * RES_KC;
    const int D0 = L.d[0], D01 = L.d[0] * L.d[1];
    const int nk = min(RES_KC, L.d[2] - 1 - k0);
    const T *x = L.x;
    // q[a][b][e]: plane k - 1 + a, row j - 1 + b, column i - 1 + e
    T q[3][3][3];
    int c = i + D0 * j + D01 * k0;
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
        for (int b = 0; b < 3; b++)
#pragma unroll
            for (int e = 0; e < 3; e++) q[a][b][e] = x[c + (a - 1) * D01 + (b - 1) * D0 + (e - 1)];
#pragma unroll
    for (int kk = 0; kk < RES_KC; kk++) {
        if (kk >= nk) break;
#pragma unroll
        for (int b = 0; b < 3; b++)
#pragma unroll
            for (int e = 0; e < 3; e++) q[2][b][e] = x[c + D01 + (b - 1) * D0 + (e - 1)];
        const T bc = L.b[c];
        T s = 0;
        s += T(-1) * q[0][0][0]; s += T(-2) * q[0][0][1]; s += T(-1) * q[0][0][2];
        s += T(-2) * q[0][1][0]; s += T(-2) * q[0][1][2];
        s += T(-1) * q[0][2][0]; s += T(-2) * q[0][2][1]; s += T(-1) * q[0][2][2];
        s += T(-2) * q[1][0][0]; s += T(-2) * q[1][0][2];
        s += T(-2) * q[1][2][0]; s += T(-2) * q[1][2][2];
        s += T(-1) * q[2][0][0]; s += T(-2) * q[2][0][1]; s += T(-1) * q[2][0][2];
        s += T(-2) * q[2][1][0]; s += T(-2) * q[2][1][2];
        s += T(-1) * q[2][2][0]; s += T(-2) * q[2][2][1]; s += T(-1) * q[2][2][2];
        emit(c, j, k0 + kk, bc - (s + T(32) * q[1][1][1]));
#pragma unroll
        for (int b = 0; b < 3; b++)
#pragma unroll
            for (int e = 0; e < 3; e++) {
                q[0][b][e] = q[1][b][e];
                q[1][b][e] = q[2][b][e];
            }
        c += D01;
    }
}

template <typename T>
__global__ void k_resid_march(MGLevel<T> L, Prd P, MarchMap A) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= A.items) return;
    resid_march(L, P, A, t, [&](int c, int, int, T r) { L.r[c] = r; });
}

// k_resid64_max for FEM levels with the interior marched along z
__global__ void k_resid64_max_march(MGLevel<double> L, Prd P, MarchMap A, unsigned long long *maxbits) {
    __shared__ unsigned long long sh[4];
    unsigned long long v = 0;
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < A.items; t += gridDim.x * blockDim.x)
        resid_march(L, P, A, t, [&](int c, int, int, double r) {
            L.r[c] = r;
            unsigned long long b = (unsigned long long)__double_as_longlong(fabs(r));
            v = b > v ? b : v;
        });
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
__global__ void k_resid(MGLevel<T> L, Prd P, NodeMap M) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= M.total) return;
    int i, j, k;
    bool in = map_node(L, M, t, i, j, k);
    int c = i + L.d[0] * (j + L.d[1] * k);
    L.r[c] = resid_value(L, P, M.fem, in, i, j, k, c);
}

// FP64 residual and its max norm: grid-stride blocks of 256, one atomic per
// block (non-negative doubles order as their bit patterns; a NaN compares
// above every finite value)
__global__ void k_resid64_max(MGLevel<double> L, Prd P, NodeMap M, unsigned long long *maxbits) {
    __shared__ unsigned long long sh[4];
    unsigned long long v = 0;
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < M.total; t += gridDim.x * blockDim.x) {
        int i, j, k;
        bool in = map_node(L, M, t, i, j, k);
        int c = i + L.d[0] * (j + L.d[1] * k);
        double r = resid_value(L, P, M.fem, in, i, j, k, c);
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
__device__ __forceinline__ void restrict_at(const MGLevel<T> &F, const MGLevel<T> &C, const Prd &P, int fem,
                                            int I, int J, int K) {
    T s = 0;
    if (!fem) {
        for (int c = 0; c < 2; c++)
            for (int b = 0; b < 2; b++)
                for (int a = 0; a < 2; a++)
                    s += F.r[(2 * I + a) + F.d[0] * ((2 * J + b) + F.d[1] * (2 * K + c))];
        s *= T(0.5);
    } else {
        // full weighting; a fine node outside the grid enters with weight 0
        // at a valid address (see mac_off)
        int c0 = 2 * I + F.d[0] * (2 * J + F.d[1] * 2 * K);
#pragma unroll
        for (int dc = -1; dc <= 1; dc++) {
            int kk = nbw(2 * K, dc, F.d[2], P.p[2]);
#pragma unroll
            for (int db = -1; db <= 1; db++) {
                int jj = nbw(2 * J, db, F.d[1], P.p[1]);
#pragma unroll
                for (int da = -1; da <= 1; da++) {
                    int ii = nbw(2 * I, da, F.d[0], P.p[0]);
                    bool v = ii >= 0 && jj >= 0 && kk >= 0;
                    T w = (da ? T(0.5) : T(1.0)) * (db ? T(0.5) : T(1.0)) * (dc ? T(0.5) : T(1.0));
                    T rv = F.r[v ? ii + F.d[0] * (jj + F.d[1] * kk) : c0];
                    s += (v ? w : T(0)) * rv;
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
__global__ void k_restrict(MGLevel<T> F, MGLevel<T> C, Prd P, int fem, float r0, float r1) {
    FLAT_IJK(C.d[0], C.d[1], r0, r1)
    if (k >= C.d[2]) return;
    restrict_at(F, C, P, fem, i, j, k);
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
__device__ __forceinline__ void prolong_at(const MGLevel<T> &F, const MGLevel<T> &C, const Prd &P, int fem,
                                           int i, int j, int k, bool assign = false) {
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
    T *xf = F.x + i + F.d[0] * (j + F.d[1] * k);
    if (assign) *xf = v;
    else *xf += v;
}

// ASSIGN: the fine x is known to be zero (no sweep before the coarse
// correction), so the interpolated correction is stored instead of added
// (no read of x; +0 instead of -0 where the correction is -0)
template <typename T, bool ASSIGN>
__global__ void k_prolong(MGLevel<T> F, MGLevel<T> C, Prd P, int fem, float r0, float r1) {
    FLAT_IJK(F.d[0], F.d[1], r0, r1)
    if (k >= F.d[2]) return;
    prolong_at(F, C, P, fem, i, j, k, ASSIGN);
}

// (i, j, k) of a flat index e of dims (a, b, .) with float reciprocals ra, rb
__device__ __forceinline__ void unflat(int e, int a, int b, float ra, float rb, int &i, int &j, int &k) {
    int row = qdiv(e, a, ra);
    i = e - row * a;
    k = qdiv(row, b, rb);
    j = row - k * b;
}

// ---- several colour passes per launch ---------------------------------------
// One workgroup owns a box of the level (its tile) and stages the tile plus a
// halo of np nodes in LDS.  Colour pass q of the np (colours c0, c0+1, ...)
// updates the nodes of its colour within Chebyshev distance np-1-q of the
// tile.  Every node a pass updates then reads neighbours that are either of
// another colour not yet updated (loaded values, correct) or of an earlier
// colour updated at a distance one larger (correct), so each tile node ends
// bit-identical to np single-colour launches over the whole level.  Halo
// nodes are recomputed by every tile that holds them; only tile nodes are
// stored, all of them, into xout: the tiles load their halos from xin while
// others store, so the two must be distinct arrays.  A periodic dimension
// wraps when loading, so the region may hold a node more than once; each
// copy is updated from copies of its own neighbours and stays correct.
constexpr int SWEEP_BATCH = 8;

template <typename T, int FEM>
__global__ __launch_bounds__(1024) void k_sweep(MGLevel<T> L, Prd P, SweepPlan S, const T *__restrict__ xin,
                                                T *__restrict__ xout) {
    extern __shared__ unsigned char smem[];
    T *xs = (T *)smem;
    constexpr int NCOL = FEM ? 4 : 2;
    const int D[3] = {L.d[0], L.d[1], L.d[2]};
    const int np = S.np;
    int tid = blockIdx.x;
    int tix[3];
    tix[0] = tid % S.nt[0];
    tid /= S.nt[0];
    tix[1] = tid % S.nt[1];
    tix[2] = tid / S.nt[1];
    // tile [t0, t1) and region [r0, r0 + rn) per dimension, in unwrapped
    // coordinates (a periodic region may start below 0 or end beyond D)
    int t0[3], t1[3], r0[3], rn[3];
#pragma unroll
    for (int a = 0; a < 3; a++) {
        t0[a] = tix[a] * S.ts[a];
        t1[a] = min(t0[a] + S.ts[a], D[a]);
        if (P.p[a]) {
            r0[a] = t0[a] - np;
            rn[a] = t1[a] - t0[a] + 2 * np;
        } else {
            r0[a] = max(t0[a] - np, 0);
            rn[a] = min(t1[a] + np, D[a]) - r0[a];
        }
    }
    const int R0 = rn[0], R01 = rn[0] * rn[1], RN = R01 * rn[2];
    // the region's unknowns and right-hand sides, side by side in LDS (the
    // passes then wait on no global load)
    T *bs = xs + S.region;
    {
        // SWEEP_BATCH elements per work-item in flight before the LDS stores,
        // so the global load latencies overlap instead of adding up
        float ra = 1.0f / rn[0], rb = 1.0f / rn[1];
        const int stride = blockDim.x;
        for (int e0 = threadIdx.x; e0 < RN; e0 += SWEEP_BATCH * stride) {
            T xv[SWEEP_BATCH], bv[SWEEP_BATCH];
#pragma unroll
            for (int u = 0; u < SWEEP_BATCH; u++) {
                int e = e0 + u * stride;
                if (e < RN) {
                    int li, lj, lk;
                    unflat(e, rn[0], rn[1], ra, rb, li, lj, lk);
                    int gi = r0[0] + li, gj = r0[1] + lj, gk = r0[2] + lk;
                    gi += (gi < 0) ? D[0] : ((gi >= D[0]) ? -D[0] : 0);
                    gj += (gj < 0) ? D[1] : ((gj >= D[1]) ? -D[1] : 0);
                    gk += (gk < 0) ? D[2] : ((gk >= D[2]) ? -D[2] : 0);
                    int gc = gi + D[0] * (gj + D[1] * gk);
                    xv[u] = xin[gc];
                    bv[u] = L.b[gc];
                }
            }
#pragma unroll
            for (int u = 0; u < SWEEP_BATCH; u++) {
                int e = e0 + u * stride;
                if (e < RN) {
                    xs[e] = xv[u];
                    bs[e] = bv[u];
                }
            }
        }
    }
    __syncthreads();
    for (int q = 0; q < np; q++) {
        const int col = (S.c0 + q) % NCOL;
        const int dist = np - 1 - q;
        // the pass's box [u0, u1) per dimension
        int u0[3], u1[3];
#pragma unroll
        for (int a = 0; a < 3; a++) {
            u0[a] = P.p[a] ? t0[a] - dist : max(t0[a] - dist, 0);
            u1[a] = P.p[a] ? t1[a] + dist : min(t1[a] + dist, D[a]);
        }
        if (FEM) {
            // rows (j, k) of the colour: j = jc, k = kc mod 2; every i
            int jc = col & 1, kc = col >> 1;
            int j0 = u0[1] + ((u0[1] - jc) & 1), k0 = u0[2] + ((u0[2] - kc) & 1);
            int nx = u1[0] - u0[0];
            int nj = u1[1] > j0 ? (u1[1] - j0 + 1) / 2 : 0;
            int nk = u1[2] > k0 ? (u1[2] - k0 + 1) / 2 : 0;
            int n = nx * nj * nk;
            float ra = 1.0f / nx, rb = 1.0f / max(nj, 1);
            for (int e = threadIdx.x; e < n; e += blockDim.x) {
                int ii, a, b;
                unflat(e, nx, max(nj, 1), ra, rb, ii, a, b);
                int g[3] = {u0[0] + ii, j0 + 2 * a, k0 + 2 * b};
                int le = (g[0] - r0[0]) + R0 * ((g[1] - r0[1]) + rn[1] * (g[2] - r0[2]));
                int w[3];
                bool inner = true;
#pragma unroll
                for (int d = 0; d < 3; d++) {
                    w[d] = g[d] + ((g[d] < 0) ? D[d] : ((g[d] >= D[d]) ? -D[d] : 0));
                    inner = inner && (P.p[d] || (g[d] >= 1 && g[d] <= D[d] - 2));
                }
                T bv = bs[le];
                if (inner) {
                    T s = fem_off_interior(xs, le, R0, R01);
                    xs[le] = (bv - s) * T(0.03125);
                } else {
                    // fem_off with wall tests on the wrapped global index and
                    // addresses in the region (a periodic neighbour is the
                    // next region node)
                    bool wl[3];
                    bool ok[3][3];
#pragma unroll
                    for (int d = 0; d < 3; d++) {
                        wl[d] = !P.p[d] && (w[d] == 0 || w[d] == D[d] - 1);
                        ok[d][0] = P.p[d] || w[d] >= 1;
                        ok[d][1] = true;
                        ok[d][2] = P.p[d] || w[d] <= D[d] - 2;
                    }
                    T sx = wl[0] ? T(0.5) : T(1), sy = wl[1] ? T(0.5) : T(1), sz = wl[2] ? T(0.5) : T(1);
                    T s = 0;
#pragma unroll
                    for (int okk = 0; okk < 3; okk++) {
#pragma unroll
                        for (int oj = 0; oj < 3; oj++) {
#pragma unroll
                            for (int oi = 0; oi < 3; oi++) {
                                const int nz = (oi != 1) + (oj != 1) + (okk != 1);
                                if (nz < 2) continue;
                                bool v = ok[0][oi] && ok[1][oj] && ok[2][okk];
                                T wt = (nz == 2) ? T(-2) : T(-1);
                                if (oi == 1) wt *= sx;
                                if (oj == 1) wt *= sy;
                                if (okk == 1) wt *= sz;
                                T xv = xs[v ? le + (oi - 1) + R0 * (oj - 1) + R01 * (okk - 1) : le];
                                s += (v ? wt : T(0)) * xv;
                            }
                        }
                    }
                    T rdg = T(0.03125);
                    if (wl[0]) rdg *= T(2);
                    if (wl[1]) rdg *= T(2);
                    if (wl[2]) rdg *= T(2);
                    xs[le] = (bv - s) * rdg;
                }
            }
        } else {
            // red-black: i of parity (j + k + col) in each (j, k) row
            int nj = u1[1] - u0[1], nk = u1[2] - u0[2];
            int hx = (u1[0] - u0[0] + 1) / 2;
            int n = hx * nj * nk;
            float ra = 1.0f / hx, rb = 1.0f / nj;
            for (int e = threadIdx.x; e < n; e += blockDim.x) {
                int ih, a, b;
                unflat(e, hx, nj, ra, rb, ih, a, b);
                int gj = u0[1] + a, gk = u0[2] + b;
                int gi = u0[0] + 2 * ih + ((u0[0] + gj + gk + col) & 1);
                if (gi >= u1[0]) continue;
                int g[3] = {gi, gj, gk};
                int le = (g[0] - r0[0]) + R0 * ((g[1] - r0[1]) + rn[1] * (g[2] - r0[2]));
                int w[3];
#pragma unroll
                for (int d = 0; d < 3; d++) w[d] = g[d] + ((g[d] < 0) ? D[d] : ((g[d] >= D[d]) ? -D[d] : 0));
                // mac_off's order: i-1, i+1, j-1, j+1, k-1, k+1
                const int off[3] = {1, R0, R01};
                T s = 0;
                int miss = 0;
#pragma unroll
                for (int d = 0; d < 3; d++) {
#pragma unroll
                    for (int sg = -1; sg <= 1; sg += 2) {
                        bool v = P.p[d] || (sg < 0 ? w[d] >= 1 : w[d] <= D[d] - 2);
                        miss += !v;
                        T xv = xs[v ? le + sg * off[d] : le];
                        s += v ? xv : T(0);
                    }
                }
                T dg = T(-6.0 + miss);
                T bv = bs[le];
                T xn = (bv - s) / dg;
                xs[le] = T(1 - MAC_OMEGA) * xs[le] + T(MAC_OMEGA) * xn;
            }
        }
        __syncthreads();
    }
    // store every node of the tile
    {
        int n0 = t1[0] - t0[0], n1 = t1[1] - t0[1], n2 = t1[2] - t0[2];
        int n = n0 * n1 * n2;
        float ra = 1.0f / n0, rb = 1.0f / n1;
        for (int e = threadIdx.x; e < n; e += blockDim.x) {
            int ii, jj, kk;
            unflat(e, n0, n1, ra, rb, ii, jj, kk);
            int gi = t0[0] + ii, gj = t0[1] + jj, gk = t0[2] + kk;
            int le = (gi - r0[0]) + R0 * ((gj - r0[1]) + rn[1] * (gk - r0[2]));
            xout[gi + D[0] * (gj + D[1] * gk)] = xs[le];
        }
    }
}

template <typename T>
__global__ void k_copy(const T *a, T *b, int n) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) b[t] = a[t];
}

// ---- the lower part of a V-cycle in one workgroup --------------------------
// Every level from the first one small enough down to the coarsest is
// processed by a single workgroup: the same per-node operations as the
// per-level kernels, in the same colour order, with workgroup barriers in
// place of kernel boundaries, so the result is identical and the launch cost
// of the small levels is paid once per V-cycle.

// Per tail level, computed on the host: the FEM colour layouts, the float
// reciprocals of the index maps and the level's place in LDS, so that a
// phase starts from kernel-argument constants instead of recomputing them.
struct TailLevel {
    FemColour F[4];
    float rd0, rd1;           // 1 / D0, 1 / D1
    int hx;                   // MAC: nodes of one colour per x-row (rounded up)
    float rhx;
    int off;                  // element offset of the level's x, b, r in LDS
};
template <typename T>
struct TailArgs {
    MGTail<T> t;
    TailLevel I[MG_TAIL_MAXLEV];
    const T *pinv;            // pseudo-inverse of the coarsest operator, or null
};

template <typename T>
__device__ void wg_smooth(const MGLevel<T> &L, const TailLevel &I, const Prd &P, int fem, int col) {
    int D0 = L.d[0], D1 = L.d[1], D2 = L.d[2];
    if (fem) {
        // the colour's interior nodes first, as in k_fem_smooth
        const FemColour &F = I.F[col];
        int n = F.NI + F.NA + F.NB;
        for (int t = threadIdx.x; t < n; t += blockDim.x) {
            int i, j, k;
            bool in = fem_colour_node(L, F, col & 1, t, i, j, k);
            int c = i + D0 * (j + D1 * k);
            if (in) {
                T s = fem_off_interior(L.x, c, D0, D0 * D1);
                L.x[c] = (L.b[c] - s) * T(0.03125);
            } else {
                fem_smooth_at(L, P, i, j, k);
            }
        }
    } else {
        int hx = I.hx, n = hx * D1 * D2;
        for (int t = threadIdx.x; t < n; t += blockDim.x) {
            int ih, j, k;
            unflat(t, hx, D1, I.rhx, I.rd1, ih, j, k);
            int i = 2 * ih + ((j + k + col) & 1);
            if (i < D0) mac_smooth_at(L, P, i, j, k);
        }
    }
    __syncthreads();
}

// IN_LDS: every tail level lives in LDS (x, b, r of each level packed in
// that order at the offset the host computed); the finest tail level's x and
// b are loaded from HBM on entry and its x stored back on exit.  The level
// views are rebuilt from the shared base at each use, so the compiler sees
// LDS pointers and issues LDS instructions.  Otherwise the levels stay in
// HBM.
template <typename T, bool IN_LDS>
__device__ __forceinline__ MGLevel<T> tail_level(const TailArgs<T> &A, unsigned char *smem, int l) {
    MGLevel<T> L;
    L.d[0] = A.t.L[l].d[0]; L.d[1] = A.t.L[l].d[1]; L.d[2] = A.t.L[l].d[2]; L.nn = A.t.L[l].nn;
    if (IN_LDS) {
        T *base = (T *)smem + A.I[l].off;
        L.x = base; L.b = base + L.nn; L.r = base + 2 * L.nn;
    } else {
        L.x = A.t.L[l].x; L.b = A.t.L[l].b; L.r = A.t.L[l].r;
    }
    return L;
}

template <typename T, bool IN_LDS>
__global__ __launch_bounds__(1024) void k_tail(TailArgs<T> A, Prd P, int fem, int vdown, int vup, int csweeps) {
    extern __shared__ unsigned char smem[];
    const int ncol = fem ? 4 : 2;
    const int last = A.t.n - 1;
    if (IN_LDS) {
        // several loads in flight per work-item before the LDS stores
        MGLevel<T> L0 = tail_level<T, true>(A, smem, 0);
        const T *gx = A.t.L[0].x, *gb = A.t.L[0].b;
        const int stride = blockDim.x;
        for (int e0 = threadIdx.x; e0 < L0.nn; e0 += SWEEP_BATCH * stride) {
            T xv[SWEEP_BATCH], bv[SWEEP_BATCH];
#pragma unroll
            for (int u = 0; u < SWEEP_BATCH; u++) {
                int e = e0 + u * stride;
                if (e < L0.nn) {
                    xv[u] = gx[e];
                    bv[u] = gb[e];
                }
            }
#pragma unroll
            for (int u = 0; u < SWEEP_BATCH; u++) {
                int e = e0 + u * stride;
                if (e < L0.nn) {
                    L0.x[e] = xv[u];
                    L0.b[e] = bv[u];
                }
            }
        }
        __syncthreads();
    }
    for (int l = 0; l < last; l++) {
        const MGLevel<T> L = tail_level<T, IN_LDS>(A, smem, l);
        const MGLevel<T> C = tail_level<T, IN_LDS>(A, smem, l + 1);
        const TailLevel &I = A.I[l];
        for (int sw = 0; sw < vdown; sw++)
            for (int c = 0; c < ncol; c++) wg_smooth(L, I, P, fem, c);
        if (fem) {
            for (int col = 0; col < 4; col++) {
                const FemColour &F = I.F[col];
                int n = F.NI + F.NA + F.NB;
                for (int e = threadIdx.x; e < n; e += blockDim.x) {
                    int i, j, k;
                    bool in = fem_colour_node(L, F, col & 1, e, i, j, k);
                    int c = i + L.d[0] * (j + L.d[1] * k);
                    L.r[c] = resid_value(L, P, 1, in, i, j, k, c);
                }
            }
        } else {
            for (int e = threadIdx.x; e < L.nn; e += blockDim.x) {
                int i, j, k;
                unflat(e, L.d[0], L.d[1], I.rd0, I.rd1, i, j, k);
                resid_at(L, P, 0, i, j, k);
            }
        }
        __syncthreads();
        const TailLevel &IC = A.I[l + 1];
        for (int e = threadIdx.x; e < C.nn; e += blockDim.x) {
            int I_, J, K;
            unflat(e, C.d[0], C.d[1], IC.rd0, IC.rd1, I_, J, K);
            restrict_at(L, C, P, fem, I_, J, K);
        }
        __syncthreads();
    }
    if (A.pinv) {
        // exact coarsest solve x = A^+ b: one work-item per row, the row
        // summed in column order (A^+ is symmetric, so column c of row r
        // is read at c n + r and a wavefront's loads are contiguous)
        const MGLevel<T> W = tail_level<T, IN_LDS>(A, smem, last);
        const int n = W.nn;
        const T *Pm = A.pinv;
        for (int r = threadIdx.x; r < n; r += blockDim.x) {
            T s = 0;
#pragma unroll 8
            for (int c = 0; c < n; c++) s += Pm[c * n + r] * W.b[c];
            W.x[r] = s;
        }
        __syncthreads();
    } else {
        const MGLevel<T> W = tail_level<T, IN_LDS>(A, smem, last);
        const TailLevel &I = A.I[last];
        for (int sw = 0; sw < csweeps; sw++)
            for (int c = 0; c < ncol; c++) {
                for (int e = threadIdx.x; e < W.nn; e += blockDim.x) {
                    int i, j, k;
                    unflat(e, W.d[0], W.d[1], I.rd0, I.rd1, i, j, k);
                    int my = fem ? ((j & 1) + 2 * (k & 1)) : ((i + j + k) & 1);
                    if (my != c) continue;
                    if (fem) fem_smooth_at(W, P, i, j, k);
                    else mac_smooth_at(W, P, i, j, k);
                }
                __syncthreads();
            }
    }
    for (int l = last - 1; l >= 0; l--) {
        const MGLevel<T> L = tail_level<T, IN_LDS>(A, smem, l);
        const MGLevel<T> C = tail_level<T, IN_LDS>(A, smem, l + 1);
        const TailLevel &I = A.I[l];
        for (int e = threadIdx.x; e < L.nn; e += blockDim.x) {
            int i, j, k;
            unflat(e, L.d[0], L.d[1], I.rd0, I.rd1, i, j, k);
            prolong_at(L, C, P, fem, i, j, k);
        }
        __syncthreads();
        for (int sw = 0; sw < vup; sw++)
            for (int c = 0; c < ncol; c++) wg_smooth(L, I, P, fem, c);
    }
    if (IN_LDS) {
        MGLevel<T> L0 = tail_level<T, true>(A, smem, 0);
        T *gx = A.t.L[0].x;
        for (int e = threadIdx.x; e < L0.nn; e += blockDim.x) gx[e] = L0.x[e];
    }
}

// mixed defect correction, first half: FP64 residual of the FP64 iterate
// straight into the FP32 right-hand side, FP32 correction zeroed
// FIRST: the FP32 correction starts from zero, so the first colour pass of
// the V-cycle's first smoothing sweep sees only zero neighbours and sets
// each node of colour 0 to a function of its own right-hand side alone.  That
// value is written here with the residual, so the V-cycle starts its first
// sweep at colour 1.  It equals the colour pass's value bit for bit: the pass
// sums +-0 terms to +0, so FEM gives (b - 0) / dg = b * rdg and MAC
// (1 - w) 0 + w (b - 0) / dg = w (b / dg).
template <bool FIRST>
__global__ void k_resid_to_f32(MGLevel<double> L, Prd P, NodeMap M, float *bf, float *xf) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= M.total) return;
    int i, j, k;
    bool in = map_node(L, M, t, i, j, k);
    int c = i + L.d[0] * (j + L.d[1] * k);
    float r = (float)resid_value(L, P, M.fem, in, i, j, k, c);
    bf[c] = r;
    float x0 = 0.f;
    if (FIRST) {
        if (M.fem) {
            if (((j | k) & 1) == 0) {
                float rdg = 0.03125f;
                if (!P.p[0] && (i == 0 || i == L.d[0] - 1)) rdg *= 2.f;
                if (!P.p[1] && (j == 0 || j == L.d[1] - 1)) rdg *= 2.f;
                if (!P.p[2] && (k == 0 || k == L.d[2] - 1)) rdg *= 2.f;
                x0 = r * rdg;
            }
        } else if (((i + j + k) & 1) == 0) {
            int miss = 0;
            miss += nbw(i, -1, L.d[0], P.p[0]) < 0;
            miss += nbw(i, 1, L.d[0], P.p[0]) < 0;
            miss += nbw(j, -1, L.d[1], P.p[1]) < 0;
            miss += nbw(j, 1, L.d[1], P.p[1]) < 0;
            miss += nbw(k, -1, L.d[2], P.p[2]) < 0;
            miss += nbw(k, 1, L.d[2], P.p[2]) < 0;
            float dg = float(-6.0 + miss);
            x0 = float(MAC_OMEGA) * (r / dg);
        }
    }
    if (xf) xf[c] = x0;
}

// k_resid_to_f32 for FEM levels with the interior marched along z
template <bool FIRST>
__global__ void k_resid_to_f32_march(MGLevel<double> L, Prd P, MarchMap A, float *bf, float *xf) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= A.items) return;
    resid_march(L, P, A, t, [&](int c, int j, int k, double rd) {
        float r = (float)rd;
        bf[c] = r;
        float x0 = 0.f;
        if (FIRST && ((j | k) & 1) == 0) {
            int i = c - L.d[0] * (j + L.d[1] * k);
            float rdg = 0.03125f;
            if (!P.p[0] && (i == 0 || i == L.d[0] - 1)) rdg *= 2.f;
            if (!P.p[1] && (j == 0 || j == L.d[1] - 1)) rdg *= 2.f;
            if (!P.p[2] && (k == 0 || k == L.d[2] - 1)) rdg *= 2.f;
            x0 = r * rdg;
        }
        if (xf) xf[c] = x0;
    });
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
void launch_mg_smooth(hipStream_t s, int fem, const MGLevel<T> &L, const int prd[3], int color, hipEvent_t start,
                      hipEvent_t stop, int bs, double *x64) {
    // default workgroup sizes, measured on the 128^3 / 129^3 levels: 1024 for
    // the red-black pass (7.8 us against 8.1 at 256, FP32; 8.8 against 10.0,
    // FP64); 64 for the FEM colour pass, whose wall-row work-items cost
    // several times an interior one and finish last, so small workgroups
    // spread them over more CUs (6.9 us against 8.9 at 256, FP32).  The
    // autotuner passes the size it measured fastest on the level at hand.
    Prd P = mkprd(prd);
    if (bs != 64 && bs != 128 && bs != 256 && bs != 512 && bs != 1024) bs = fem ? 64 : 1024;
    if (!fem) {
        int hx = (L.d[0] + 1) / 2, n = hx * L.d[1] * L.d[2];
        hipExtLaunchKernelGGL(k_mac_smooth<T>, dim3((n + bs - 1) / bs), dim3(bs), 0, s, start, stop, 0, L, P,
                              color, hx, 1.0f / hx, 1.0f / L.d[1], x64);
    } else {
        FemColour F = fem_colour(L.d, color);
        int n = F.NI + F.NA + F.NB;
        hipExtLaunchKernelGGL(k_fem_smooth<T>, dim3((n + bs - 1) / bs), dim3(bs), 0, s, start, stop, 0, L, P, color,
                              F, x64);
    }
}
static inline dim3 flat_grid(int n) { return dim3((n + 255) / 256); }

// FEM residuals marched along z (RMT3D_RES_MARCH=0: one work-item per node)
static bool res_march_on() {
    static const bool on = [] {
        const char *e = getenv("RMT3D_RES_MARCH");
        return !(e && e[0] == '0');
    }();
    return on;
}

// Tiles of at most 17, 9 or 5 nodes per dimension: the largest that gives
// 256 workgroups (one per CU) and whose region (tile plus halo, unknowns and
// right-hand sides) fits in LDS, else 5.  A region holds at most
// (5 + 2 np)^3 = 2197 unknowns at the smallest tile: 35 KB of LDS in FP64.
SweepPlan mg_sweep_plan(int fem, const int d[3], const int prd[3], int esize, int np) {
    SweepPlan S{};
    static const int sizes[3] = {17, 9, 5};
    for (int z = 0; z < 3; z++) {
        int tiles = 1, region = 1;
        for (int a = 0; a < 3; a++) {
            S.nt[a] = (d[a] + sizes[z] - 1) / sizes[z];
            S.ts[a] = (d[a] + S.nt[a] - 1) / S.nt[a];
            tiles *= S.nt[a];
            int r = S.ts[a] + 2 * np;
            if (!prd[a] && r > d[a]) r = d[a];
            region *= r;
        }
        S.region = region;
        if (tiles >= 256 && 2 * (size_t)region * esize <= MG_TAIL_LDS_MAX) break;
    }
    S.np = np;
    S.c0 = 0;
    S.lds = 2 * S.region * esize;
    // measured per sweep: FEM 65^3 (region 17^3) 17.4 us at 512 work-items
    // against 21.6 at 1024; MAC 64^3 (region 12^3) 8.1 against 8.7 at 256
    S.block = S.region >= 1024 ? 512 : 256;
    (void)fem;
    return S;
}

void mg_sweep_enable() {
    const int bytes = (int)MG_TAIL_LDS_MAX;
    (void)hipFuncSetAttribute((const void *)k_sweep<double, 0>, hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
    (void)hipFuncSetAttribute((const void *)k_sweep<double, 1>, hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
    (void)hipFuncSetAttribute((const void *)k_sweep<float, 0>, hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
    (void)hipFuncSetAttribute((const void *)k_sweep<float, 1>, hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
    (void)hipGetLastError();
}

template <typename T>
void launch_mg_sweep(hipStream_t s, int fem, const MGLevel<T> &L, const int prd[3], const SweepPlan &plan,
                     const T *xin, T *xout, hipEvent_t start, hipEvent_t stop) {
    dim3 grid(plan.nt[0] * plan.nt[1] * plan.nt[2]);
    if (fem)
        hipExtLaunchKernelGGL((k_sweep<T, 1>), grid, dim3(plan.block), plan.lds, s, start, stop, 0, L, mkprd(prd),
                              plan, xin, xout);
    else
        hipExtLaunchKernelGGL((k_sweep<T, 0>), grid, dim3(plan.block), plan.lds, s, start, stop, 0, L, mkprd(prd),
                              plan, xin, xout);
}
template <typename T>
void launch_mg_copy(hipStream_t s, const T *a, T *b, int n) {
    hipLaunchKernelGGL(k_copy<T>, dim3((n + 255) / 256), dim3(256), 0, s, a, b, n);
}

template <typename T>
void launch_mg_resid(hipStream_t s, int fem, const MGLevel<T> &L, const int prd[3]) {
    NodeMap M = node_map(L.d, fem);
    if (!M.plain && res_march_on()) {
        MarchMap A = march_map(L.d);
        hipLaunchKernelGGL(k_resid_march<T>, flat_grid(A.items), dim3(256), 0, s, L, mkprd(prd), A);
        return;
    }
    hipLaunchKernelGGL(k_resid<T>, flat_grid(M.total), dim3(256), 0, s, L, mkprd(prd), M);
}
void launch_mg_resid64_max(hipStream_t s, int fem, const MGLevel<double> &L, const int prd[3],
                           unsigned long long *maxbits) {
    NodeMap M = node_map(L.d, fem);
    if (!M.plain && res_march_on()) {
        MarchMap A = march_map(L.d);
        int nb = (A.items + 255) / 256;
        if (nb > 1024) nb = 1024;
        hipLaunchKernelGGL(k_resid64_max_march, dim3(nb), dim3(256), 0, s, L, mkprd(prd), A, maxbits);
        return;
    }
    int nb = (L.nn + 255) / 256;
    if (nb > 1024) nb = 1024;
    hipLaunchKernelGGL(k_resid64_max, dim3(nb), dim3(256), 0, s, L, mkprd(prd), M, maxbits);
}
// MAC residual and restriction in one pass: one work-item per coarse cell
// computes the residuals of its eight fine cells (every fine cell has one
// coarse parent, so nothing is computed twice) and sums them in
// restrict_at's order; the fine residual array is not written (the V-cycle
// reads only the coarse right-hand side).  A coarse cell away from walls and
// wraps loads the 32 distinct unknowns of its children and their face
// neighbours once and sums each child's terms in mac_off's order, so every
// value is the two-launch form's bit for bit.
template <typename T>
__global__ void k_mac_resid_restrict(MGLevel<T> F, MGLevel<T> C, Prd P, float r0, float r1) {
    FLAT_IJK(C.d[0], C.d[1], r0, r1)
    if (k >= C.d[2]) return;
    T s = 0;
    const bool inner = i >= 1 && i <= C.d[0] - 2 && j >= 1 && j <= C.d[1] - 2 && k >= 1 && k <= C.d[2] - 2;
    if (inner) {
        const int D0 = F.d[0], D01 = F.d[0] * F.d[1];
        const int c0 = 2 * i + D0 * (2 * j) + D01 * (2 * k);
        const T *x = F.x;
        T xc[2][2][2], bb[2][2][2], xm[2][2], xp[2][2], ym[2][2], yp[2][2], zm[2][2], zp[2][2];
#pragma unroll
        for (int c = 0; c < 2; c++)
#pragma unroll
            for (int b = 0; b < 2; b++) {
                const int row = c0 + b * D0 + c * D01;
                xm[c][b] = x[row - 1];
                xp[c][b] = x[row + 2];
#pragma unroll
                for (int a = 0; a < 2; a++) {
                    xc[c][b][a] = x[row + a];
                    bb[c][b][a] = F.b[row + a];
                }
            }
#pragma unroll
        for (int u = 0; u < 2; u++)
#pragma unroll
            for (int a = 0; a < 2; a++) {
                ym[u][a] = x[c0 + a + u * D01 - D0];        // (2i+a, 2j-1, 2k+u)
                yp[u][a] = x[c0 + a + u * D01 + 2 * D0];    // (2i+a, 2j+2, 2k+u)
                zm[u][a] = x[c0 + a + u * D0 - D01];        // (2i+a, 2j+u, 2k-1)
                zp[u][a] = x[c0 + a + u * D0 + 2 * D01];    // (2i+a, 2j+u, 2k+2)
            }
#pragma unroll
        for (int c = 0; c < 2; c++)
#pragma unroll
            for (int b = 0; b < 2; b++)
#pragma unroll
                for (int a = 0; a < 2; a++) {
                    T sm = 0;
                    sm += a ? xc[c][b][0] : xm[c][b];
                    sm += a ? xp[c][b] : xc[c][b][1];
                    sm += b ? xc[c][0][a] : ym[c][a];
                    sm += b ? yp[c][a] : xc[c][1][a];
                    sm += c ? xc[0][b][a] : zm[b][a];
                    sm += c ? zp[b][a] : xc[1][b][a];
                    s += bb[c][b][a] - (sm + T(-6.0) * xc[c][b][a]);
                }
    } else {
        for (int c = 0; c < 2; c++)
            for (int b = 0; b < 2; b++)
                for (int a = 0; a < 2; a++) {
                    const int fi = 2 * i + a, fj = 2 * j + b, fk = 2 * k + c;
                    const int fc = fi + F.d[0] * (fj + F.d[1] * fk);
                    s += resid_value(F, P, 0, false, fi, fj, fk, fc);
                }
    }
    s *= T(0.5);
    const int ci = i + C.d[0] * (j + C.d[1] * k);
    C.b[ci] = s;
    C.x[ci] = 0;
}

// r = b - A x on F, then C.b = R r and C.x = 0: one launch for MAC levels
// (F.r is left as it was), the residual and restriction launches for FEM
template <typename T>
void launch_mg_resid_restrict(hipStream_t s, int fem, const MGLevel<T> &F, const MGLevel<T> &C, const int prd[3]) {
    if (!fem) {
        hipLaunchKernelGGL(k_mac_resid_restrict<T>, flat_grid(C.nn), dim3(256), 0, s, F, C, mkprd(prd),
                           1.0f / C.d[0], 1.0f / C.d[1]);
        return;
    }
    launch_mg_resid(s, fem, F, prd);
    launch_mg_restrict(s, fem, F, C, prd);
}

template <typename T>
void launch_mg_restrict(hipStream_t s, int fem, const MGLevel<T> &F, const MGLevel<T> &C, const int prd[3]) {
    hipLaunchKernelGGL(k_restrict<T>, flat_grid(C.nn), dim3(256), 0, s, F, C, mkprd(prd), fem, 1.0f / C.d[0],
                       1.0f / C.d[1]);
}
template <typename T>
void launch_mg_prolong(hipStream_t s, int fem, const MGLevel<T> &F, const MGLevel<T> &C, const int prd[3],
                       int assign) {
    if (assign)
        hipLaunchKernelGGL((k_prolong<T, true>), flat_grid(F.nn), dim3(256), 0, s, F, C, mkprd(prd), fem,
                           1.0f / F.d[0], 1.0f / F.d[1]);
    else
        hipLaunchKernelGGL((k_prolong<T, false>), flat_grid(F.nn), dim3(256), 0, s, F, C, mkprd(prd), fem,
                           1.0f / F.d[0], 1.0f / F.d[1]);
}
template <typename T>
size_t mg_tail_lds_bytes(const MGLevel<T> *L, int n) {
    size_t b = 0;
    for (int l = 0; l < n; l++) b += 3 * sizeof(T) * (size_t)L[l].nn;
    return b;
}

template <typename T>
int mg_tail_enable_lds(size_t bytes) {
    int dev = 0, maxb = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 0;
    if (hipDeviceGetAttribute(&maxb, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess) return 0;
    if (bytes > (size_t)maxb) return 0;
    // raising the dynamic limit is required on some runtimes and a no-op on
    // others; it is raised to the most any shape uses, since every context
    // of the process shares the kernel's attribute
    (void)hipFuncSetAttribute((const void *)k_tail<T, true>, hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)std::min((size_t)maxb, MG_TAIL_LDS_MAX));
    (void)hipGetLastError();
    return 1;
}

template <typename T>
void launch_mg_tail(hipStream_t s, int fem, const MGLevel<T> *L, int n, const int prd[3], int vdown, int vup,
                    int csweeps, int lds, const T *pinv) {
    TailArgs<T> A{};
    A.t.n = n;
    A.pinv = pinv;
    int off = 0;
    for (int l = 0; l < n; l++) {
        A.t.L[l] = L[l];
        TailLevel &I = A.I[l];
        for (int c = 0; c < 4; c++) I.F[c] = fem_colour(L[l].d, c);
        I.rd0 = 1.0f / L[l].d[0];
        I.rd1 = 1.0f / L[l].d[1];
        I.hx = (L[l].d[0] + 1) / 2;
        I.rhx = 1.0f / I.hx;
        I.off = off;
        off += 3 * L[l].nn;
    }
    if (lds)
        hipLaunchKernelGGL((k_tail<T, true>), dim3(1), dim3(1024), mg_tail_lds_bytes(L, n), s, A, mkprd(prd), fem, vdown,
                           vup, csweeps);
    else
        hipLaunchKernelGGL((k_tail<T, false>), dim3(1), dim3(1024), 0, s, A, mkprd(prd), fem, vdown, vup, csweeps);
}
void launch_resid_to_f32(hipStream_t s, int fem, const MGLevel<double> &L, const int prd[3], float *bf, float *xf,
                         int first_pass) {
    NodeMap M = node_map(L.d, fem);
    if (!M.plain && res_march_on()) {
        MarchMap A = march_map(L.d);
        if (first_pass)
            hipLaunchKernelGGL(k_resid_to_f32_march<true>, flat_grid(A.items), dim3(256), 0, s, L, mkprd(prd), A, bf,
                               xf);
        else
            hipLaunchKernelGGL(k_resid_to_f32_march<false>, flat_grid(A.items), dim3(256), 0, s, L, mkprd(prd), A, bf,
                               xf);
        return;
    }
    if (first_pass)
        hipLaunchKernelGGL(k_resid_to_f32<true>, flat_grid(M.total), dim3(256), 0, s, L, mkprd(prd), M, bf, xf);
    else
        hipLaunchKernelGGL(k_resid_to_f32<false>, flat_grid(M.total), dim3(256), 0, s, L, mkprd(prd), M, bf, xf);
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

// ---------------------------------------------------------------------------
// Fast-diagonalisation solve of the coarsest level (Lynch, Rice & Thomas,
// Numer. Math. 6, 1964).  Both operators are sums of Kronecker products of
// 1D matrices: MAC T.I.I + I.T.I + I.I.T (T the 1D second difference with
// Neumann ends folded in), FEM 12 (K.M.M + M.K.M + M.M.K) (K, M the 1D Q1
// stiffness and mass matrices of unit elements, Neumann ends halved).  With
// per-dimension bases S (S^T M S = I, S^T K S = diag(lam); M = I for MAC),
// x = (Sx.Sy.Sz) D^+ (Sx.Sy.Sz)^T b, D the sum of the three lam (times 12
// for FEM): three contractions forward (the last scaled by D^+) and three
// back.  Each output sums its d terms in ascending order, as the oracle's
// restatement does, so FP64 results agree bit for bit.
// ---------------------------------------------------------------------------

// The closed-form bases, angles reduced exactly in integers:
//   FEM wall (d nodes, n = d-1 elements): cos(pi p q / n),
//     lam = 6 (1 - cos(pi p / n)) / (2 + cos(pi p / n));
//   MAC wall (d cells): cos(pi p (2q+1) / 2d), lam = -2 (1 - cos(pi p / d));
//   periodic: p = 0 constant, p = 2k-1 cos(2 pi k q / d), p = 2k
//     sin(2 pi k q / d), with the lam of the angle 2 pi k / d.
// Each mode is scaled to unit M-norm (FEM M = tridiag(1, 4, 1) / 6, wall
// ends 2/6; MAC M = I), summed in point order.
void mg_fdm_basis(int fem, int d, int prd, double *S, double *lam) {
    const double pi = 3.14159265358979323846;
    std::vector<double> v(d);
    for (int p = 0; p < d; p++) {
        double c;
        if (!prd && fem) {
            int n = d - 1;
            for (int q = 0; q < d; q++) v[q] = det_cos(pi * (double)(((long)p * q) % (2L * n)) / n);
            c = det_cos(pi * p / n);
        } else if (!prd) {
            for (int q = 0; q < d; q++) v[q] = det_cos(pi * (double)(((long)p * (2 * q + 1)) % (4L * d)) / (2.0 * d));
            c = det_cos(pi * p / d);
        } else {
            int k = (p + 1) / 2;
            for (int q = 0; q < d; q++) {
                double ang = 2.0 * pi * (double)(((long)k * q) % d) / d;
                v[q] = (p == 0 || (p & 1)) ? det_cos(ang) : det_sin(ang);
            }
            c = det_cos(2.0 * pi * k / d);
        }
        lam[p] = fem ? 6.0 * (1.0 - c) / (2.0 + c) : -2.0 * (1.0 - c);
        double nrm = 0;
        for (int q = 0; q < d; q++) {
            if (!fem) {
                nrm += v[q] * v[q];
                continue;
            }
            int lo = q - 1, hi = q + 1;
            int end = 0;
            if (prd) {
                lo = (lo + d) % d;
                hi = hi % d;
            } else {
                end = (q == 0 || q == d - 1);
            }
            double mv = (end ? 2.0 : 4.0) * v[q];
            if (lo >= 0) mv += v[lo];
            if (hi < d) mv += v[hi];
            nrm += v[q] * (mv / 6.0);
        }
        double sq = sqrt(nrm);
        for (int q = 0; q < d; q++) S[(size_t)q * d + p] = v[q] / sq;
    }
}

// out[.., p, ..] = sum over q of W[q*d + p] in[.., q, ..] along one
// dimension, each output summed over ascending q from +0.  A wavefront
// computes PB consecutive outputs p for 64 pencils (the same p for all
// lanes, so its W entries are LDS broadcasts).  The kernels are bound by
// memory latency, not bandwidth or arithmetic: each workgroup issues all the
// loads it stages into LDS (W's fdm_rows(da) x PB slices, and for x the
// row tile) as one batch; along y and z it streams pencil values from memory
// FDM_QU at a time, double-buffered.
constexpr int FDM_QU = 32;

// W slice element e of output block p0: W[q*da + p0 + min(r, last)] with
// q = e / PB, r = e % PB (columns past the end repeat column da-1,
// unused); rows q >= da are zero.  The terms run in branch-free chunks of
// FDM_QU, then of 4: the slice has fdm_rows(da) rows (da rounded up to a
// multiple of 4) and pencil values past da are zero too.  The extra terms
// are +0 x w, which leave a sum that started at +0 unchanged (such a sum is
// never -0).
__host__ __device__ constexpr int fdm_rows(int da) { return (da + 3) / 4 * 4; }

template <int PB, typename T>
__device__ __forceinline__ T fdm_w_elem(const T *__restrict__ W, int da, int p0, int e) {
    int q = e / PB, r = e - q * PB;
    int last = da - 1 - p0;
    T w = W[min(q, da - 1) * da + p0 + (r < last ? r : last)];
    return q < da ? w : T(0);
}

template <int U, int PB, typename T>
__device__ __forceinline__ void fdm_terms(T (&acc)[PB], const T *v, const T *wsl, int q0) {
#pragma unroll
    for (int u = 0; u < U; u++) {
        const T *w = wsl + (q0 + u) * PB;
#pragma unroll
        for (int r = 0; r < PB; r++) acc[r] += w[r] * v[u];
    }
}

// y or z: one wavefront per workgroup, lanes on consecutive x positions
// (coalesced loads and stores), blockIdx.y the output block.  Pencil c
// starts at (c % inner) + (c / inner) * ostride and steps by stride; SCALE
// multiplies each output by dinv.
template <typename T, bool SCALE, int PB>
__global__ void __launch_bounds__(64) k_fdm_cols(const T *__restrict__ in, T *__restrict__ out,
                                                 const T *__restrict__ W, const T *__restrict__ dinv, int ncols,
                                                 int inner, int ostride, int da, int stride) {
    extern __shared__ unsigned char fdm_lds[];
    T *wsl = reinterpret_cast<T *>(fdm_lds);
    const int p0 = blockIdx.y * PB;
    const int nw = fdm_rows(da) * PB;
    const int c = min((int)(blockIdx.x * blockDim.x + threadIdx.x), ncols - 1);
    const int co = c / inner;
    const int base = (c - co * inner) + co * ostride;
    const int dfull = da / FDM_QU * FDM_QU, dpad = fdm_rows(da);
    T v[FDM_QU], vn[FDM_QU];
    // the first pencil values and the W slice, one batch of loads
#pragma unroll
    for (int u = 0; u < FDM_QU; u++) {
        v[u] = in[base + min(u, da - 1) * stride];
        vn[u] = T(0);
    }
    for (int e0 = 0; e0 < nw; e0 += 64 * FDM_QU) {
        T wv[FDM_QU];
#pragma unroll
        for (int u = 0; u < FDM_QU; u++) wv[u] = fdm_w_elem<PB>(W, da, p0, min(e0 + u * 64 + (int)threadIdx.x, nw - 1));
#pragma unroll
        for (int u = 0; u < FDM_QU; u++) {
            int e = e0 + u * 64 + threadIdx.x;
            if (e < nw) wsl[e] = wv[u];
        }
    }
    __syncthreads();
    T acc[PB];
#pragma unroll
    for (int r = 0; r < PB; r++) acc[r] = T(0);
    for (int q0 = 0; q0 < dfull; q0 += FDM_QU) {
        if (q0 + FDM_QU < dfull)
#pragma unroll
            for (int u = 0; u < FDM_QU; u++) vn[u] = in[base + (q0 + FDM_QU + u) * stride];
        fdm_terms<FDM_QU, PB>(acc, v, wsl, q0);
#pragma unroll
        for (int u = 0; u < FDM_QU; u++) v[u] = vn[u];
    }
    for (int q0 = dfull; q0 < dpad; q0 += 4) {
        T t4[4];
#pragma unroll
        for (int u = 0; u < 4; u++) {
            T x = in[base + min(q0 + u, da - 1) * stride];
            t4[u] = q0 + u < da ? x : T(0);
        }
        fdm_terms<4, PB>(acc, t4, wsl, q0);
    }
    if (blockIdx.x * blockDim.x + threadIdx.x >= (unsigned)ncols) return;
#pragma unroll
    for (int r = 0; r < PB; r++)
        if (p0 + r < da) {
            int e = base + (p0 + r) * stride;
            T o = acc[r];
            if (SCALE) o *= dinv[e];
            out[e] = o;
        }
}

// x: a workgroup of nwv wavefronts stages rb (<= 64) consecutive rows in LDS
// (pitch d0 | 1: an odd number of words or doublewords, so the 64 lanes'
// rows start in different LDS banks) and the W slices of nwv output
// blocks, in one batch of
// loads; wavefront w computes block blockIdx.y * nwv + w for all rows (lane
// = row) and stores it straight to the rows.
template <typename T, int PB>
__global__ void __launch_bounds__(256) k_fdm_rows(const T *__restrict__ in, T *__restrict__ out,
                                                  const T *__restrict__ W, int d0, int nrows, int rb) {
    extern __shared__ unsigned char fdm_lds[];
    const int nwv = blockDim.x >> 6, npb = (d0 + PB - 1) / PB;
    const int nw = fdm_rows(d0) * PB;         // one block's W slice
    T *wsl = reinterpret_cast<T *>(fdm_lds);      // nwv slices
    T *tile = wsl + nwv * nw;
    const int ld = d0 | 1;
    const int r0 = blockIdx.x * rb;
    const int nr = min(rb, nrows - r0);
    const int pb0 = blockIdx.y * nwv;
    const int nbw = min(nwv, npb - pb0);          // blocks of this workgroup
    const T *src = in + (size_t)r0 * d0;
    const int nt = nr * d0, ntot = nbw * nw + nt;
    for (int e0 = 0; e0 < ntot; e0 += blockDim.x * FDM_QU) {
        T v[FDM_QU];
#pragma unroll
        for (int u = 0; u < FDM_QU; u++) {
            int e = min(e0 + u * (int)blockDim.x + (int)threadIdx.x, ntot - 1);
            if (e < nbw * nw) {
                int b = e / nw;
                v[u] = fdm_w_elem<PB>(W, d0, (pb0 + b) * PB, e - b * nw);
            } else {
                v[u] = src[e - nbw * nw];
            }
        }
#pragma unroll
        for (int u = 0; u < FDM_QU; u++) {
            int e = e0 + u * blockDim.x + threadIdx.x;
            if (e < nbw * nw) {
                wsl[e] = v[u];
            } else if (e < ntot) {
                int t = e - nbw * nw, r = t / d0;
                tile[r * ld + (t - r * d0)] = v[u];
            }
        }
    }
    __syncthreads();
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (wv >= nbw) return;
    const int p0 = (pb0 + wv) * PB;
    const T *ws = wsl + wv * nw;
    const int rl = lane < nr ? lane : nr - 1;     // idle lanes repeat the last row, unstored
    const T *row = tile + rl * ld;
    T acc[PB];
#pragma unroll
    for (int r = 0; r < PB; r++) acc[r] = T(0);
    // row values come from LDS; here a guarded chunk of FDM_QU terms
    // measured faster than the zero-padded chunks of the y and z passes
    for (int q0 = 0; q0 < d0; q0 += FDM_QU) {
        T v[FDM_QU];
#pragma unroll
        for (int u = 0; u < FDM_QU; u++) v[u] = row[min(q0 + u, d0 - 1)];
#pragma unroll
        for (int u = 0; u < FDM_QU; u++) {
            if (q0 + u < d0) {
                const T *w = ws + (q0 + u) * PB;
#pragma unroll
                for (int r = 0; r < PB; r++) acc[r] += w[r] * v[u];
            }
        }
    }
    if (lane < nr) {
        T *dst = out + (size_t)(r0 + lane) * d0;
#pragma unroll
        for (int r = 0; r < PB; r++)
            if (p0 + r < d0) dst[p0 + r] = acc[r];
    }
}

template <typename T, int PB>
static void fdm_pass_pb(hipStream_t s, const MGLevel<T> &L, int a, const T *W, const T *dinv, const T *in, T *out) {
    int d0 = L.d[0], d1 = L.d[1], d2 = L.d[2];
    int da = L.d[a];
    size_t wbytes = sizeof(T) * (size_t)fdm_rows(da) * PB;
    if (a == 0) {
        // up to 4 output blocks and 64 rows per workgroup, within 64 KiB of LDS
        int npb = (d0 + PB - 1) / PB;
        int nwv = npb < 4 ? npb : 4;
        while (nwv > 1 && nwv * wbytes > 32768) nwv--;
        int rb = 64;
        while (rb > 1 && nwv * wbytes + sizeof(T) * (size_t)rb * (d0 | 1) > 65536) rb /= 2;
        int nrows = d1 * d2;
        size_t lds = nwv * wbytes + sizeof(T) * (size_t)rb * (d0 | 1);
        hipLaunchKernelGGL((k_fdm_rows<T, PB>), dim3((nrows + rb - 1) / rb, (npb + nwv - 1) / nwv), dim3(64 * nwv), lds,
                           s, in, out, W, d0, nrows, rb);
        return;
    }
    int ncols = L.nn / da;
    int inner = a == 1 ? d0 : d0 * d1, ostride = a == 1 ? d0 * d1 : 0, stride = a == 1 ? d0 : d0 * d1;
    dim3 g((ncols + 63) / 64, (da + PB - 1) / PB);
    if (dinv)
        hipLaunchKernelGGL((k_fdm_cols<T, true, PB>), g, dim3(64), wbytes, s, in, out, W, dinv, ncols, inner, ostride,
                           da, stride);
    else
        hipLaunchKernelGGL((k_fdm_cols<T, false, PB>), g, dim3(64), wbytes, s, in, out, W, (const T *)nullptr, ncols,
                           inner, ostride, da, stride);
}

// pb: outputs per work-item (2, 4 or 8); every choice sums each output's
// terms in the same order, so they give identical results
template <typename T>
static void fdm_pass(hipStream_t s, const MGLevel<T> &L, int a, int pb, const T *W, const T *dinv, const T *in, T *out) {
    if (pb == 2) fdm_pass_pb<T, 2>(s, L, a, W, dinv, in, out);
    else if (pb == 4) fdm_pass_pb<T, 4>(s, L, a, W, dinv, in, out);
    else fdm_pass_pb<T, 8>(s, L, a, W, dinv, in, out);
}

template <typename T>
void launch_mg_fdm_pass(hipStream_t s, const MGLevel<T> &L, int a, int pb, const T *W, const T *dinv, const T *in,
                        T *out) {
    fdm_pass<T>(s, L, a, pb, W, dinv, in, out);
}

// ---- FP32 (the mixed path's preconditioner): three launches on MFMA ----
// k_fdm_plane_mx: one workgroup per z-plane, forward: the x then the y
// contraction, inverse: y then x.  k_fdm_pencils_mx: 64 consecutive
// z-pencils per workgroup: the forward z contraction, the D^+ scaling and
// the inverse z contraction, written back in place.  Each contraction is a
// small GEMM on v_mfma_f32_16x16x4f32 with every extent padded to a multiple
// of 16 with zeros in LDS.  The sums run in the matrix cores' order, not
// the six-launch form's: the preconditioner's values differ in the last
// FP32 bits, and the FP64 defect correction around it converges to the
// same tolerance.
typedef float fdm_f4 __attribute__((ext_vector_type(4)));
constexpr int FDM_MX_BLOCK = 1024;

__host__ __device__ constexpr int fdm_x16(int d) { return (d + 15) & ~15; }
__host__ __device__ constexpr int fdm_xp(int d) { return fdm_x16(d) + 4; }   // LDS row pitch (floats)

// C[m][n] = sum_k A[m][k] B[k][n] for m < M, n < N, k < K (multiples of 16):
// A row-major (pitch pa), B given as its transpose BT[n][k] (pitch pb), both
// in LDS.  A wavefront takes 16 x 16 tiles.  In an MFMA step lane l gives
// the A entry of row l % 16 and the B entry of column l % 16 at k-slot
// l / 16; here slot t of the four steps of a 16-wide k chunk is k = kb + 4 t
// + s (s = step), so each lane reads its four A and four B entries with one
// 16-byte load each.  The result D[m][n] of a tile lands in lane
// (n % 16) + 16 (m % 16 / 4), element m % 4.
template <typename F>
__device__ __forceinline__ void fdm_mx_gemm(const float *A, int pa, const float *BT, int pb, int M, int N, int K,
                                            F store) {
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, nw = blockDim.x >> 6;
    const int r = lane & 15, slot = lane >> 4;
    const int tn = N >> 4, nt = (M >> 4) * tn;
    for (int t = wave; t < nt; t += nw) {
        const int m0 = (t / tn) << 4, n0 = (t - (t / tn) * tn) << 4;
        const float *ap = A + (m0 + r) * pa + 4 * slot;
        const float *bp = BT + (n0 + r) * pb + 4 * slot;
        fdm_f4 acc = {0.f, 0.f, 0.f, 0.f};
        for (int kb = 0; kb < K; kb += 16) {
            const fdm_f4 av = *reinterpret_cast<const fdm_f4 *>(ap + kb);
            const fdm_f4 bv = *reinterpret_cast<const fdm_f4 *>(bp + kb);
            acc = __builtin_amdgcn_mfma_f32_16x16x4f32(av[0], bv[0], acc, 0, 0, 0);
            acc = __builtin_amdgcn_mfma_f32_16x16x4f32(av[1], bv[1], acc, 0, 0, 0);
            acc = __builtin_amdgcn_mfma_f32_16x16x4f32(av[2], bv[2], acc, 0, 0, 0);
            acc = __builtin_amdgcn_mfma_f32_16x16x4f32(av[3], bv[3], acc, 0, 0, 0);
        }
#pragma unroll
        for (int v = 0; v < 4; v++) store(m0 + 4 * slot + v, n0 + r, acc[v]);
    }
}

// Staging of a kernel's LDS operands: element e < total is load(e, o),
// stored at lds[o]; each work-item issues FDM_MX_U loads before its first
// store, so the loads of all the operands are in flight together.
constexpr int FDM_MX_U = 16;
template <typename L>
__device__ __forceinline__ void fdm_mx_stage(float *lds, int total, L load) {
    for (int e0 = threadIdx.x; e0 < total; e0 += FDM_MX_U * blockDim.x) {
        float v[FDM_MX_U];
        int o[FDM_MX_U];
#pragma unroll
        for (int u = 0; u < FDM_MX_U; u++) {
            int e = e0 + u * blockDim.x;
            o[u] = -1;
            v[u] = e < total ? load(e, o[u]) : 0.f;
        }
#pragma unroll
        for (int u = 0; u < FDM_MX_U; u++)
            if (o[u] >= 0) lds[o[u]] = v[u];
    }
}

// element e of the rows x cols matrix src (row-major, pitch sp) padded to a
// rows16 x cols16 block at LDS offset off with pitch dp (zero outside), or
// transposed (stored at [j][i]); cols16 = C, rC = 1 / C
struct FdmMxPad {
    const float *src;
    int off, sp, dp, rows, cols, C, tr;
    float rC;
    __device__ __forceinline__ float load(int e, int &o) const {
        int i = qdiv(e, C, rC), j = e - i * C;
        o = off + (tr ? j * dp + i : i * dp + j);
        return (i < rows && j < cols) ? src[i * sp + j] : 0.f;
    }
    __device__ __forceinline__ int count() const { return fdm_x16(rows) * C; }
};
__device__ __forceinline__ FdmMxPad fdm_mx_pad(int off, int dp, const float *src, int sp, int rows, int cols,
                                               int tr = 0) {
    const int C = fdm_x16(cols);
    return FdmMxPad{src, off, sp, dp, rows, cols, C, tr, 1.0f / C};
}

// three padded copies staged together
__device__ __forceinline__ void fdm_mx_stage3(float *lds, const FdmMxPad &a, const FdmMxPad &b, const FdmMxPad &c) {
    const int na = a.count(), nb = b.count(), n = na + nb + c.count();
    fdm_mx_stage(lds, n, [&](int e, int &o) {
        return e < na ? a.load(e, o) : e < na + nb ? b.load(e - na, o) : c.load(e - na - nb, o);
    });
}

// forward: out = (S1^T (S0^T planes)) ; inverse: out = S0 (S1 planes) per
// plane, with S = the bases (S[q][p] = basis q of mode p) and ST their
// transposes, all d x d row-major
template <bool INV>
__global__ void __launch_bounds__(FDM_MX_BLOCK) k_fdm_plane_mx(const float *__restrict__ in, float *__restrict__ out,
                                                               const float *__restrict__ S0,
                                                               const float *__restrict__ ST0,
                                                               const float *__restrict__ S1,
                                                               const float *__restrict__ ST1, int d0, int d1) {
    extern __shared__ unsigned char fdm_lds[];
    const int X0 = fdm_x16(d0), X1 = fdm_x16(d1), p0 = fdm_xp(d0), p1 = fdm_xp(d1), np = d0 * d1;
    float *buf = reinterpret_cast<float *>(fdm_lds);
    const float *src = in + (size_t)blockIdx.x * np;
    float *dst = out + (size_t)blockIdx.x * np;
    if (!INV) {
        float *P = buf;                 // [X1][p0]  plane[j][q]
        float *B1 = P + X1 * p0;        // [X0][p0]  S0^T[p][q]
        float *A2 = B1 + X0 * p0;       // [X1][p1]  S1^T[p1][q]
        float *TT = A2 + X1 * p1;       // [X0][p1]  T[j][p] stored at [p][j]
        fdm_mx_stage3(buf, fdm_mx_pad(0, p0, src, d0, d1, d0), fdm_mx_pad(B1 - buf, p0, ST0, d0, d0, d0),
                      fdm_mx_pad(A2 - buf, p1, ST1, d1, d1, d1));
        __syncthreads();
        // T[j][p] = sum_q plane[j][q] S0[q][p]
        fdm_mx_gemm(P, p0, B1, p0, X1, X0, X0, [&](int m, int n, float v) { TT[n * p1 + m] = v; });
        __syncthreads();
        // out[p1][i] = sum_q S1[q][p1] T[q][i]
        fdm_mx_gemm(A2, p1, TT, p1, X1, X0, X1, [&](int m, int n, float v) {
            if (m < d1 && n < d0) dst[m * d0 + n] = v;
        });
    } else {
        float *XT = buf;                // [X0][p1]  plane[q][i] stored at [i][q]
        float *A1 = XT + X0 * p1;       // [X1][p1]  S1[p1][q]
        float *O1 = A1 + X1 * p1;       // [X1][p0]
        float *B2 = O1 + X1 * p0;       // [X0][p0]  S0[p][q]
        // the plane is read along i (coalesced) and stored transposed
        fdm_mx_stage3(buf, fdm_mx_pad(0, p1, src, d0, d1, d0, 1), fdm_mx_pad(A1 - buf, p1, S1, d1, d1, d1),
                      fdm_mx_pad(B2 - buf, p0, S0, d0, d0, d0));
        __syncthreads();
        // O1[p1][i] = sum_q ST1[q][p1] plane[q][i]
        fdm_mx_gemm(A1, p1, XT, p1, X1, X0, X1, [&](int m, int n, float v) { O1[m * p0 + n] = v; });
        __syncthreads();
        // out[j][p] = sum_q O1[j][q] ST0[q][p]
        fdm_mx_gemm(O1, p0, B2, p0, X1, X0, X0, [&](int m, int n, float v) {
            if (m < d1 && n < d0) dst[m * d0 + n] = v;
        });
    }
}

__global__ void __launch_bounds__(FDM_MX_BLOCK) k_fdm_pencils_mx(float *__restrict__ io, const float *__restrict__ S2,
                                                                 const float *__restrict__ ST2,
                                                                 const float *__restrict__ dinv, int np, int d2) {
    extern __shared__ unsigned char fdm_lds[];
    const int X2 = fdm_x16(d2), pz = fdm_xp(d2);
    float *VT = reinterpret_cast<float *>(fdm_lds);   // [64][pz]  values[q][c] at [c][q]
    float *UT = VT + 64 * pz;                           // [64][pz]
    float *Af = UT + 64 * pz;                           // [X2][pz]  S2^T[p][q]
    float *Ai = Af + X2 * pz;                           // [X2][pz]  S2[p][q]
    const int c0 = blockIdx.x * 64, nc = min(64, np - c0);
    // the pencils are read along c (coalesced) and stored transposed
    FdmMxPad v = fdm_mx_pad(0, pz, io + c0, np, d2, nc, 1);
    v.C = 64;
    v.rC = 1.0f / 64;
    fdm_mx_stage3(VT, v, fdm_mx_pad(Af - VT, pz, ST2, d2, d2, d2), fdm_mx_pad(Ai - VT, pz, S2, d2, d2, d2));
    __syncthreads();
    // U[p][c] = D^+ sum_q S2[q][p] values[q][c]
    fdm_mx_gemm(Af, pz, VT, pz, X2, 64, X2, [&](int m, int n, float v) {
        if (m < d2 && n < nc) v *= dinv[(size_t)m * np + c0 + n];
        UT[n * pz + m] = v;
    });
    __syncthreads();
    // values[p][c] = sum_q ST2[q][p] U[q][c]
    fdm_mx_gemm(Ai, pz, UT, pz, X2, 64, X2, [&](int m, int n, float v) {
        if (m < d2 && n < nc) io[(size_t)m * np + c0 + n] = v;
    });
}

// ---- the same three launches with direct operands (every extent <= 128) --
// A staged workgroup holds a whole plane (or 64 pencils) and its bases in
// LDS, so one CU's matrix cores run all of a plane's tiles and the level
// occupies only d2 CUs.  Here a workgroup computes one 16-wide block of a
// plane's output (or 16 pencils), one wavefront per 16-row tile: the
// operands are read straight from memory (the bases stay in L1/L2) into
// registers, every chunk of a tile at once, and only the 16-wide
// intermediate product passes through LDS.  Operands, k-slot order and
// roundings are those of the staged kernels, so the results are identical.
constexpr int FDM_DX_MAXD = 128;
constexpr int FDM_DX_PITCH = FDM_DX_MAXD + 4;

// One 16 x 16 tile over nc <= NC 16-wide chunks of the contraction:
// ld_a(c) and ld_b(c) give this lane's A[m0 + r][16 c + 4 slot + s] and
// B[16 c + 4 slot + s][n0 + r] for s = 0..3 (r = lane % 16, slot = lane / 16).
// All NC chunks are loaded unconditionally (the loaders return zeros past
// the extents), so the loads form one batch; the products stop at nc.
template <int NC, typename LA, typename LB>
__device__ __forceinline__ fdm_f4 fdm_dx_tile(int nc, LA ld_a, LB ld_b) {
    fdm_f4 a[NC], b[NC];
#pragma unroll
    for (int c = 0; c < NC; c++) {
        a[c] = ld_a(c);
        b[c] = ld_b(c);
    }
    fdm_f4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int c = 0; c < NC; c++)
        if (c < nc) {
            acc = __builtin_amdgcn_mfma_f32_16x16x4f32(a[c][0], b[c][0], acc, 0, 0, 0);
            acc = __builtin_amdgcn_mfma_f32_16x16x4f32(a[c][1], b[c][1], acc, 0, 0, 0);
            acc = __builtin_amdgcn_mfma_f32_16x16x4f32(a[c][2], b[c][2], acc, 0, 0, 0);
            acc = __builtin_amdgcn_mfma_f32_16x16x4f32(a[c][3], b[c][3], acc, 0, 0, 0);
        }
    return acc;
}

// four consecutive entries M[row][k0 .. k0+3] of a rows x cols row-major
// matrix (pitch ld), zero outside it.  Entries outside are loaded from M[0]
// and dropped, so that every load is unconditional and a tile's loads are
// all in flight together.
__device__ __forceinline__ fdm_f4 fdm_dx_row4(const float *__restrict__ M, int ld, int row, int rows, int k0,
                                              int cols) {
    fdm_f4 v;
    const bool in = row < rows;
#pragma unroll
    for (int s = 0; s < 4; s++) {
        const bool ok = in && k0 + s < cols;
        const float x = M[ok ? (size_t)row * ld + k0 + s : 0];
        v[s] = ok ? x : 0.f;
    }
    return v;
}
// four consecutive entries of row `row` of a zero-padded X x X basis
// (X = 16 nc): chunk c < nc, the chunks past nc read chunk nc - 1 (their
// products are not taken)
__device__ __forceinline__ fdm_f4 fdm_dx_pad4(const float *__restrict__ M, int X, int row, int c, int nc, int slot) {
    const int cc = c < nc ? c : nc - 1;
    return *reinterpret_cast<const fdm_f4 *>(M + row * X + 16 * cc + 4 * slot);
}
// M[k0 + s][col], s = 0..3, zero outside
__device__ __forceinline__ fdm_f4 fdm_dx_col4(const float *__restrict__ M, int ld, int col, int cols, int k0,
                                              int rows) {
    fdm_f4 v;
    const bool in = col < cols;
#pragma unroll
    for (int s = 0; s < 4; s++) {
        const bool ok = in && k0 + s < rows;
        const float x = M[ok ? (size_t)(k0 + s) * ld + col : 0];
        v[s] = ok ? x : 0.f;
    }
    return v;
}

// forward, plane blockIdx.y, output columns p in [16 blockIdx.x, +16):
// T[j][p] = sum_q plane[j][q] S0[q][p] (wave w: rows j of tile w), then
// out[p1][p] = sum_j S1[j][p1] T[j][p] (wave w: rows p1 of tile w);
// 64 ceil(d1 / 16) work-items
template <int NC>
__global__ void __launch_bounds__(64 * NC) k_fdm_plane_dx_fwd(const float *__restrict__ in,
                                                                     float *__restrict__ out,
                                                                     const float *__restrict__ ST0,
                                                                     const float *__restrict__ ST1, int d0, int d1) {
    // ST0, ST1: padded (pitch 16 nc0, 16 nc1)
    __shared__ __align__(16) float TT[16 * FDM_DX_PITCH];   // T[j][n0 + r] at [r][j]
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63, r = lane & 15, slot = lane >> 4;
    const int n0 = blockIdx.x * 16, m = 16 * w + r;
    const size_t np = (size_t)d0 * d1;
    const float *src = in + blockIdx.y * np;
    float *dst = out + blockIdx.y * np;
    const int nc0 = fdm_x16(d0) >> 4, nc1 = fdm_x16(d1) >> 4;
    fdm_f4 t = fdm_dx_tile<NC>(
        nc0, [&](int c) { return fdm_dx_row4(src, d0, m, d1, 16 * c + 4 * slot, d0); },
        [&](int c) { return fdm_dx_pad4(ST0, 16 * nc0, n0 + r, c, nc0, slot); });
    *reinterpret_cast<fdm_f4 *>(TT + r * FDM_DX_PITCH + 16 * w + 4 * slot) = t;
    __syncthreads();
    fdm_f4 o = fdm_dx_tile<NC>(
        nc1, [&](int c) { return fdm_dx_pad4(ST1, 16 * nc1, m, c, nc1, slot); },
        [&](int c) { return *reinterpret_cast<const fdm_f4 *>(TT + r * FDM_DX_PITCH + 16 * c + 4 * slot); });
    const int p = n0 + r;
#pragma unroll
    for (int v = 0; v < 4; v++) {
        const int p1 = 16 * w + 4 * slot + v;
        if (p1 < d1 && p < d0) dst[p1 * d0 + p] = o[v];
    }
}

// inverse, plane blockIdx.y, output rows j in [16 blockIdx.x, +16):
// O1[j][i] = sum_q S1[j][q] plane[q][i] (wave w: columns i of tile w), then
// out[j][p] = sum_i O1[j][i] S0[p][i] (wave w: columns p of tile w);
// 64 ceil(d0 / 16) work-items
template <int NC>
__global__ void __launch_bounds__(64 * NC) k_fdm_plane_dx_inv(const float *__restrict__ in,
                                                                     float *__restrict__ out,
                                                                     const float *__restrict__ S0,
                                                                     const float *__restrict__ S1, int d0, int d1) {
    // S0, S1: padded (pitch 16 nc0, 16 nc1)
    __shared__ __align__(16) float O1[16 * FDM_DX_PITCH];   // O1[m0 + a][i] at [a][i]
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63, r = lane & 15, slot = lane >> 4;
    const int m0 = blockIdx.x * 16, n = 16 * w + r;
    const size_t np = (size_t)d0 * d1;
    const float *src = in + blockIdx.y * np;
    float *dst = out + blockIdx.y * np;
    const int nc0 = fdm_x16(d0) >> 4, nc1 = fdm_x16(d1) >> 4;
    fdm_f4 o = fdm_dx_tile<NC>(
        nc1, [&](int c) { return fdm_dx_pad4(S1, 16 * nc1, m0 + r, c, nc1, slot); },
        [&](int c) { return fdm_dx_col4(src, d0, n, d0, 16 * c + 4 * slot, d1); });
#pragma unroll
    for (int v = 0; v < 4; v++) O1[(4 * slot + v) * FDM_DX_PITCH + n] = o[v];
    __syncthreads();
    fdm_f4 u = fdm_dx_tile<NC>(
        nc0, [&](int c) { return *reinterpret_cast<const fdm_f4 *>(O1 + r * FDM_DX_PITCH + 16 * c + 4 * slot); },
        [&](int c) { return fdm_dx_pad4(S0, 16 * nc0, n, c, nc0, slot); });
#pragma unroll
    for (int v = 0; v < 4; v++) {
        const int j = m0 + 4 * slot + v;
        if (j < d1 && n < d0) dst[j * d0 + n] = u[v];
    }
}

// pencils c in [16 blockIdx.x, +16) of the np = d0 d1 z-pencils, in place:
// U[p][c] = D^+ sum_q S2[q][p] values[q][c], then values[p][c] =
// sum_q S2[p][q] U[q][c] (wave w: rows p of tile w); 64 ceil(d2 / 16)
// work-items
template <int NC>
__global__ void __launch_bounds__(64 * NC) k_fdm_pencils_dx(float *__restrict__ io,
                                                                   const float *__restrict__ S2,
                                                                   const float *__restrict__ ST2,
                                                                   const float *__restrict__ dinv, int np, int d2) {
    // S2, ST2: padded (pitch 16 nc2)
    __shared__ __align__(16) float UT[16 * FDM_DX_PITCH];   // U[p][c0 + r] at [r][p]
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63, r = lane & 15, slot = lane >> 4;
    const int c = blockIdx.x * 16 + r, m = 16 * w + r;
    const int nc2 = fdm_x16(d2) >> 4;
    fdm_f4 u = fdm_dx_tile<NC>(
        nc2, [&](int k) { return fdm_dx_pad4(ST2, 16 * nc2, m, k, nc2, slot); },
        [&](int k) { return fdm_dx_col4(io, np, c, np, 16 * k + 4 * slot, d2); });
#pragma unroll
    for (int v = 0; v < 4; v++) {
        const int p = 16 * w + 4 * slot + v;
        if (p < d2 && c < np) u[v] *= dinv[(size_t)p * np + c];
    }
    *reinterpret_cast<fdm_f4 *>(UT + r * FDM_DX_PITCH + 16 * w + 4 * slot) = u;
    __syncthreads();
    fdm_f4 o = fdm_dx_tile<NC>(
        nc2, [&](int k) { return fdm_dx_pad4(S2, 16 * nc2, m, k, nc2, slot); },
        [&](int k) { return *reinterpret_cast<const fdm_f4 *>(UT + r * FDM_DX_PITCH + 16 * k + 4 * slot); });
#pragma unroll
    for (int v = 0; v < 4; v++) {
        const int p = 16 * w + 4 * slot + v;
        if (p < d2 && c < np) io[(size_t)p * np + c] = o[v];
    }
}

static bool fdm_dx_fits(const int d[3]) {
    return d[0] <= FDM_DX_MAXD && d[1] <= FDM_DX_MAXD && d[2] <= FDM_DX_MAXD;
}

// NC: chunks loaded per tile, at least the chunks of either contraction of
// the launch (a plane kernel's work-items: 64 per 16-row tile, fewer than
// 64 NC when the extents differ)
template <int NC>
static void fdm_direct_nc(hipStream_t s, const MGLevel<float> &L, const FdmDev<float> &F, int plane) {
    const int d0 = L.d[0], d1 = L.d[1], d2 = L.d[2], np = d0 * d1;
    const int nc0 = fdm_x16(d0) >> 4, nc1 = fdm_x16(d1) >> 4, nc2 = fdm_x16(d2) >> 4;
    if (plane == 0)
        hipLaunchKernelGGL(k_fdm_plane_dx_fwd<NC>, dim3(nc0, d2), dim3(64 * nc1), 0, s, (const float *)L.b, L.r,
                           F.STP[0], F.STP[1], d0, d1);
    else if (plane == 1)
        hipLaunchKernelGGL(k_fdm_pencils_dx<NC>, dim3((np + 15) / 16), dim3(64 * nc2), 0, s, L.r, F.SP[2], F.STP[2],
                           F.dinv, np, d2);
    else
        hipLaunchKernelGGL(k_fdm_plane_dx_inv<NC>, dim3(nc1, d2), dim3(64 * nc0), 0, s, (const float *)L.r, L.x,
                           F.SP[0], F.SP[1], d0, d1);
}

static void fdm_direct(hipStream_t s, const MGLevel<float> &L, const FdmDev<float> &F) {
    const int nc01 = fdm_x16(std::max(L.d[0], L.d[1])) >> 4, nc2 = fdm_x16(L.d[2]) >> 4;
    for (int k = 0; k < 3; k++) {
        switch (k == 1 ? nc2 : nc01) {
            case 1: fdm_direct_nc<1>(s, L, F, k); break;
            case 2: fdm_direct_nc<2>(s, L, F, k); break;
            case 3: fdm_direct_nc<3>(s, L, F, k); break;
            case 4: fdm_direct_nc<4>(s, L, F, k); break;
            case 5: fdm_direct_nc<5>(s, L, F, k); break;
            case 6: fdm_direct_nc<6>(s, L, F, k); break;
            case 7: fdm_direct_nc<7>(s, L, F, k); break;
            default: fdm_direct_nc<8>(s, L, F, k); break;
        }
    }
}

static size_t fdm_mx_plane_bytes(const int d[3]) {
    size_t X0 = fdm_x16(d[0]), X1 = fdm_x16(d[1]), p0 = fdm_xp(d[0]), p1 = fdm_xp(d[1]);
    return sizeof(float) * (X1 * p0 + X0 * p0 + X1 * p1 + X0 * p1);
}
static size_t fdm_mx_pencil_bytes(const int d[3]) {
    size_t X2 = fdm_x16(d[2]), pz = fdm_xp(d[2]);
    return sizeof(float) * (2 * 64 * pz + 2 * X2 * pz);
}

size_t mg_fdm_fused_bytes(const int d[3]) { return std::max(fdm_mx_plane_bytes(d), fdm_mx_pencil_bytes(d)); }

int mg_fdm_fused_enable(const int d[3]) {
    const char *dx = getenv("RMT3D_FDM_DIRECT");
    const int direct = (fdm_dx_fits(d) && !(dx && dx[0] == '0')) ? 2 : 0;
    int dev = 0, maxb = 0;
    if (hipGetDevice(&dev) != hipSuccess) return direct;
    if (hipDeviceGetAttribute(&maxb, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess) return direct;
    size_t bytes = mg_fdm_fused_bytes(d);
    if (bytes > (size_t)maxb || bytes > MG_TAIL_LDS_MAX) return direct;
    // every context of the process shares the kernels' attribute: raise it
    // to the most any shape may use
    const int lim = (int)std::min((size_t)maxb, MG_TAIL_LDS_MAX);
    (void)hipFuncSetAttribute((const void *)k_fdm_plane_mx<false>, hipFuncAttributeMaxDynamicSharedMemorySize, lim);
    (void)hipFuncSetAttribute((const void *)k_fdm_plane_mx<true>, hipFuncAttributeMaxDynamicSharedMemorySize, lim);
    (void)hipFuncSetAttribute((const void *)k_fdm_pencils_mx, hipFuncAttributeMaxDynamicSharedMemorySize, lim);
    (void)hipGetLastError();
    return 1 | direct;
}

static void fdm_fused(hipStream_t s, const MGLevel<float> &L, const FdmDev<float> &F) {
    const int d0 = L.d[0], d1 = L.d[1], d2 = L.d[2], np = d0 * d1;
    const size_t bp = fdm_mx_plane_bytes(L.d), bz = fdm_mx_pencil_bytes(L.d);
    const dim3 blk(FDM_MX_BLOCK);
    hipLaunchKernelGGL(k_fdm_plane_mx<false>, dim3(d2), blk, bp, s, (const float *)L.b, L.r, F.S[0], F.ST[0], F.S[1],
                       F.ST[1], d0, d1);
    hipLaunchKernelGGL(k_fdm_pencils_mx, dim3((np + 63) / 64), blk, bz, s, L.r, F.S[2], F.ST[2], F.dinv, np, d2);
    hipLaunchKernelGGL(k_fdm_plane_mx<true>, dim3(d2), blk, bp, s, (const float *)L.r, L.x, F.S[0], F.ST[0], F.S[1],
                       F.ST[1], d0, d1);
}
// forward along x, y, z (the last scaled by D^+), back along z, y, x,
// through the level's r and x arrays (b is left as it is): one launch per
// contraction, or (FP32 with F.fused) the three MFMA launches
template <typename T>
void launch_mg_fdm(hipStream_t s, const MGLevel<T> &L, const FdmDev<T> &F) {
    if constexpr (std::is_same<T, float>::value) {
        if (F.fused == 2) {
            fdm_direct(s, L, F);
            return;
        }
        if (F.fused) {
            fdm_fused(s, L, F);
            return;
        }
    }
    fdm_pass<T>(s, L, 0, F.pb[0], F.S[0], nullptr, L.b, L.r);
    fdm_pass<T>(s, L, 1, F.pb[1], F.S[1], nullptr, L.r, L.x);
    fdm_pass<T>(s, L, 2, F.pb[2], F.S[2], F.dinv, L.x, L.r);
    fdm_pass<T>(s, L, 2, F.pb[2], F.ST[2], nullptr, L.r, L.x);
    fdm_pass<T>(s, L, 1, F.pb[1], F.ST[1], nullptr, L.x, L.r);
    fdm_pass<T>(s, L, 0, F.pb[0], F.ST[0], nullptr, L.r, L.x);
}

#define INST(T)                                                                                              \
    template void launch_mg_fdm<T>(hipStream_t, const MGLevel<T> &, const FdmDev<T> &);                     \
    template void launch_mg_smooth<T>(hipStream_t, int, const MGLevel<T> &, const int *, int, hipEvent_t,    \
                                      hipEvent_t, int, double *);                                            \
    template void launch_mg_fdm_pass<T>(hipStream_t, const MGLevel<T> &, int, int, const T *, const T *,    \
                                        const T *, T *);                                                     \
    template void launch_mg_sweep<T>(hipStream_t, int, const MGLevel<T> &, const int *, const SweepPlan &,   \
                                     const T *, T *, hipEvent_t, hipEvent_t);                                \
    template void launch_mg_copy<T>(hipStream_t, const T *, T *, int);                                      \
    template void launch_mg_resid<T>(hipStream_t, int, const MGLevel<T> &, const int *);                    \
    template void launch_mg_resid_restrict<T>(hipStream_t, int, const MGLevel<T> &, const MGLevel<T> &,      \
                                              const int *);                                                 \
    template void launch_mg_restrict<T>(hipStream_t, int, const MGLevel<T> &, const MGLevel<T> &, const int *); \
    template void launch_mg_prolong<T>(hipStream_t, int, const MGLevel<T> &, const MGLevel<T> &, const int *,   \
                                       int);                                                                  \
    template void launch_mg_tail<T>(hipStream_t, int, const MGLevel<T> *, int, const int *, int, int, int, int,  \
                                    const T *);                                                              \
    template size_t mg_tail_lds_bytes<T>(const MGLevel<T> *, int);                                          \
    template int mg_tail_enable_lds<T>(size_t);
INST(double)
INST(float)

}  // namespace rmt
