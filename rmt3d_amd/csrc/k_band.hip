// Band ("extraps") kernels: layered least-squares extrapolation of the
// reference map, the cell-sorted band build, the predictor/corrector updates
// and redraw_boundary.
//
// Extrapolation follows fluid_3d::extrapolate (fluid_3d_ex.cc:31-328) with
// extrap_helper's masks (Extrapolation/extrap_helper.{hh,cc}):
//  - one object at a time; the object's maps live in a dense per-cell
//    scratch (lid: -1 none, 0 primary, L layer; val = xpred), which is exactly
//    what get_refmap(ind, oid) sees during that object's extrapolation;
//  - the BFS keeps the reference's u32 mask array and nmask counter; the
//    immediate mask writes of add_to_next_list become a candidate flag that is
//    applied after the pass (same set: the write only de-duplicates);
//  - the layer list lives on the device; kernels loop over its device count,
//    so no host round trip is needed per layer;
//  - every list cell fits its weighted affine model in FP64 with two passes
//    over the search cube (means, then centred moments) in the reference's
//    summation order.
#include <limits.h>

#include <algorithm>
#include "dev.h"
#include "launch.h"

namespace rmt {

__device__ __forceinline__ int IXb(const Geo &g, int i, int j, int k) { return i + g.m4 * (j + g.n4 * k); }
__device__ __forceinline__ void unpack(const Geo &g, int ind, int &i, int &j, int &k) {
    k = ind / g.mn4; j = (ind % g.mn4) / g.m4; i = ind % g.m4;
}

// periodic source of a ghost cell (cell table)
__device__ __forceinline__ int psrc(const Geo &g, int i, int j, int k) {
    int c[3] = {i, j, k};
    const int s[3] = {g.m, g.n, g.o};
    bool any = false;
    for (int d = 0; d < 3; d++) {
        int t = (c[d] < 2) ? 0 : ((c[d] >= s[d] + 2) ? 2 : 1);
        if (t != 1) {
            if (!g.prd[d]) return -1;
            any = true;
            c[d] += (t == 0) ? s[d] : -s[d];
        }
    }
    if (!any) return -1;
    return IXb(g, c[0], c[1], c[2]);
}

// extrap_helper::within_extrapolation_domain on one rank (extrap_helper.hh:295-404)
__device__ __forceinline__ bool within_extrap(const Geo &g, int ind) {
    int c[3];
    unpack(g, ind, c[0], c[1], c[2]);
    const int s[3] = {g.m, g.n, g.o};
    int t[3];
    bool any = false, has = true;
    for (int d = 0; d < 3; d++) {
        t[d] = (c[d] < 2) ? 0 : ((c[d] > s[d] + 1) ? 2 : 1);
        if (t[d] != 1) { any = true; if (!g.prd[d]) has = false; }
    }
    if (!any) return true;
    if (t[0] == 2 && t[1] == 2 && t[2] == 2) return true;
    return !has;
}
__device__ __forceinline__ bool within_inner_ghost(const Geo &g, int ind) {
    int i, j, k;
    unpack(g, ind, i, j, k);
    return i >= 1 && i <= g.m4 - 2 && j >= 1 && j <= g.n4 - 2 && k >= 1 && k <= g.o4 - 2;
}

// ---- object scratch and masks ----
__global__ void k_obj_init(Geo g, RMap R, Scratch W, uint32_t oid, uint32_t nmask) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= g.Ng) return;
    if (R.c[t] == oid) {
        W.lid[t] = 0;
        for (int a = 0; a < 3; a++) W.val[a][t] = R.xpred[a][t];
        W.mask[t] = nmask;
    } else {
        W.lid[t] = -1;
    }
    W.flag[t] = 0;
}

__global__ void k_mask_reset(Geo g, Scratch W) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < g.Ng) W.mask[t] = 0;
}

__device__ __forceinline__ void push6(const Geo &g, Scratch &W, int ind, uint32_t cur_mask) {
    const int st[6] = {-1, 1, -g.m4, g.m4, -g.mn4, g.mn4};
    for (int a = 0; a < 6; a++) {
        int cand = ind + st[a];
        if (W.mask[cand] < cur_mask) W.flag[cand] = 1;
    }
}

// add_first_layer (extrap_helper.cc:207-219): primary cells in the inner box
__global__ void k_first_push(Geo g, Scratch W, uint32_t cur_mask) {
    int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
    int j = blockIdx.y + 1, k = blockIdx.z + 1;
    if (i > g.m4 - 2) return;
    int ind = IXb(g, i, j, k);
    if (W.mask[ind] == cur_mask) push6(g, W, ind, cur_mask);
}

// next-layer candidates from the list (fluid_3d_ex.cc:277-284)
__global__ void k_list_push(Geo g, Scratch W, const int *list, const int *cnt, uint32_t cur_mask) {
    int n = *cnt;
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < n; t += gridDim.x * blockDim.x) {
        int ind = list[t];
        if (within_inner_ghost(g, ind)) push6(g, W, ind, cur_mask);
    }
}

__global__ void k_apply_flags(Geo g, Scratch W, uint32_t nmask) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= g.Ng) return;
    if (W.flag[t]) W.mask[t] = nmask;
}

// extrap_helper::comm_mask on one rank (periodic copy) followed by
// scan_ghost_region (extrap_helper.cc:261-327)
__global__ void k_comm_mask(Geo g, Scratch W) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= g.Ng) return;
    int i, j, k;
    unpack(g, t, i, j, k);
    int s = psrc(g, i, j, k);
    if (s >= 0) W.mask[t] = W.mask[s];
}
__global__ void k_scan_ghost(Geo g, Scratch W, uint32_t nmask) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= g.Ng) return;
    int i, j, k;
    unpack(g, t, i, j, k);
    bool ghost = (i < 2 || i > g.m + 1 || j < 2 || j > g.n + 1 || k < 2 || k > g.o + 1);
    if (ghost && W.mask[t] == nmask) W.flag[t] = 1;
}

// update_list: the flagged cells become the next list (set semantics)
__global__ __launch_bounds__(1024) void k_compact(Geo g, Scratch W, int *list, int *cnt) {
    __shared__ int sh[17];
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    bool f = t < g.Ng && W.flag[t] != 0;
    int p = block_append(f, cnt, sh);
    if (f) {
        list[p] = t;
        W.flag[t] = 0;
    }
}

// ---- per-cell weighted least squares (fluid_3d_ex.cc:94-271, extrap_helper.cc:34-169) ----
// The average unit normal of the object's previous-step face normals around
// cell ind (fluid_3d_ex.cc:101-140): the cell's three lower faces and the
// lower face of its +x, +y and +z neighbours, each from that cell's map of
// the object in the previous band (get_refmap_prev, fluid_3d_utils.cc:44-55).
// Called by a whole wavefront for one cell: lanes 0-3 look up the four cells'
// maps and load their face normals at once (the lookups are chains of
// dependent loads), then every lane adds them up in the reference's order,
// cell by cell and face by face, from values broadcast by shuffles, so the
// sums are the serial loop's.
__device__ void prev_normal(const Geo &g, const RMap &R, const Band &T, int ind, unsigned oid, double gp[3],
                            double &gnorm) {
    int ii, jj, kk;
    unpack(g, ind, ii, jj, kk);
    gp[0] = gp[1] = gp[2] = 0;
    gnorm = 0;
    if (!(ii >= 2 && ii <= g.m4 - 3 && jj >= 2 && jj <= g.n4 - 3 && kk >= 2 && kk <= g.o4 - 3)) return;
    const int lane = __lane_id();
    const int pp = lane < 4 ? lane : 0;
    const int c = pp == 0 ? ind : pp == 1 ? ind + 1 : pp == 2 ? ind + g.m4 : ind + g.mn4;
    int found = 0;
    // face normals of the lane's cell, in scalars (no private arrays): cell 0
    // contributes its three lower faces (a*, b*, c* = faces 0, 1, 2), cells
    // 1-3 the face pp - 1, their lower face towards cell 0 (a*)
    double ax = 0, ay = 0, az = 0, bx = 0, by = 0, bz = 0, cx = 0, cy = 0, cz = 0;
    if (lane < 4) {
        int at = -1;
        if (R.c[c] == oid) {
            found = 1;
            at = c;
        } else {
            int s0 = T.start[c], e0 = s0 + T.cnt[c];
            for (int a = s0; a < e0; a++)
                if ((T.c[a] & OMASK) == oid) { found = 1; at = -2 - a; break; }
        }
        // component f of the cell's previous-step face normals
        auto ld = [&](int f) { return at >= 0 ? R.gphi[f][at] : T.gphi[f][-2 - at]; };
        if (found) {
            const int d0 = pp == 0 ? 0 : pp - 1;
            ax = ld(3 * d0);
            ay = ld(3 * d0 + 1);
            az = ld(3 * d0 + 2);
            if (pp == 0) {
                bx = ld(3); by = ld(4); bz = ld(5);
                cx = ld(6); cy = ld(7); cz = ld(8);
            }
        }
    }
    int nf = 0;
    double sx = 0, sy = 0, sz = 0;
    // one face normal (x, y, z) of face dd: skipped when its dd component
    // carries the unset marker, counted when any component is nonzero
    auto face = [&](double x, double y, double z, int dd) {
        const double vd = dd == 0 ? x : dd == 1 ? y : z;
        if (vd == -1.e200) return;
        sx += x;
        sy += y;
        sz += z;
        if (x != 0.0 || y != 0.0 || z != 0.0) nf++;
    };
    if (__shfl(found, 0, 64)) {
        face(__shfl(ax, 0, 64), __shfl(ay, 0, 64), __shfl(az, 0, 64), 0);
        face(__shfl(bx, 0, 64), __shfl(by, 0, 64), __shfl(bz, 0, 64), 1);
        face(__shfl(cx, 0, 64), __shfl(cy, 0, 64), __shfl(cz, 0, 64), 2);
    }
#pragma unroll
    for (int p = 1; p < 4; p++)
        if (__shfl(found, p, 64)) face(__shfl(ax, p, 64), __shfl(ay, p, 64), __shfl(az, p, 64), p - 1);
    gp[0] = sx;
    gp[1] = sy;
    gp[2] = sz;
    if (nf == 6) {
        double nrm = 0;
        for (int a = 0; a < 3; a++) { gp[a] /= nf; nrm += gp[a] * gp[a]; }
        if (nrm != 0) {
            nrm = sqrt(nrm);
            for (int a = 0; a < 3; a++) gp[a] /= nrm;
            gnorm = nrm;
        } else {
            gp[0] = gp[1] = gp[2] = 0;
        }
    } else {
        gp[0] = gp[1] = gp[2] = 0;
    }
}

struct LsqArgs {
    int L;
    int max_rs;
    const double *wpow;
};

// weight of neighbour (p,q,r) or a negative value if it is not a data point
__device__ __forceinline__ double nb_weight(const Geo &g, const Scratch &W, int ind, int ii, int jj, int kk, int p,
                                            int q, int r, int L, bool cosw, const double gp[3], const double *wpow) {
    if (ii + p > g.m4 - 1 || ii + p < 0) return -1;
    if (jj + q > g.n4 - 1 || jj + q < 0) return -1;
    if (kk + r > g.o4 - 1 || kk + r < 0) return -1;
    int nb = ind + p + g.m4 * q + g.mn4 * r;
    int lid = W.lid[nb];
    if (lid < 0 || lid >= L) return -1;
    int dist = abs(p) + abs(q) + abs(r);
    if (cosw) {
        double ev0 = g.lx[ii] - g.lx[ii + p], ev1 = g.ly[jj] - g.ly[jj + q], ev2 = g.lz[kk] - g.lz[kk + r];
        double dw = 0., en = 0.;
        dw += ev0 * gp[0]; en += ev0 * ev0;
        dw += ev1 * gp[1]; en += ev1 * ev1;
        dw += ev2 * gp[2]; en += ev2 * ev2;
        en = sqrt(en);
        dw /= en;
        dw *= wpow[dist];
        if (!(dw > 0)) return -1;
        return dw;
    }
    return wpow[dist];
}

// One wavefront per cell of the layer.  The lanes evaluate the candidate
// stencil points in parallel, 64 at a time in the (r, q, p) order of the
// reference's loops, and park each point's terms in LDS; then one lane per
// sum accumulates its sum over the points in that order.  Every sum is thus
// the same sequence of additions as the serial loop of extrap_helper
// (bit-identical), while the stencil gathers of a cell run 64-wide.  A point
// that is not a data point parks +0 terms: the running sums start at +0 and
// are never -0, so adding them leaves the sums bit-identical and the summing
// loops are branch-free.  Terms are stored quantity-major ([quantity][point],
// rows padded to 65 doubles): a store puts the 64 lanes' terms in 64
// adjacent words, and the summing lanes (one per quantity) read rows whose
// starts fall in different banks, so neither side conflicts on LDS banks.
constexpr int LSQ_T1 = 7;    // ravg[3], davg[3], wsum
constexpr int LSQ_T2 = 15;   // am[6], rhs[9]

// cnt_next: the next layer's list counter, cleared here (nothing reads it
// during this launch)
__global__ __launch_bounds__(64) void k_gather_lsq(Geo g, RMap R, Band T, Scratch W, const int *list, const int *cnt,
                                                   LsqArgs A, const BoxBatch *__restrict__ B, int *status,
                                                   int *cnt_next) {
    if (blockIdx.x == 0 && threadIdx.x == 0) *cnt_next = 0;
    __shared__ double term[LSQ_T2][65];
    __shared__ double bc[LSQ_T2];
    const int lane = threadIdx.x;
    int n = *cnt;
    for (int t = blockIdx.x; t < n; t += gridDim.x) {
        int ind = list[t];
        if (!within_extrap(g, ind)) continue;
        int ii, jj, kk;
        unpack(g, ind, ii, jj, kk);
        // the cell's object: the one batch box that holds it (lane b tests box b)
        bool inb = false;
        if (lane < B->n) {
            const int *lo = B->lo[lane], *ex = B->ext[lane];
            inb = ii >= lo[0] && ii < lo[0] + ex[0] && jj >= lo[1] && jj < lo[1] + ex[1] && kk >= lo[2] &&
                  kk < lo[2] + ex[2];
        }
        unsigned long long hit = __ballot(inb);
        unsigned oid = hit ? (unsigned)B->oid[__ffsll((long long)hit) - 1] : 0u;
        // the previous normal only weights the first two layers
        // (fluid_3d_ex.cc:196: grad_phi_norm != 0 && n_layers <= 2)
        double gp[3] = {0, 0, 0}, gnorm = 0;
        if (A.L <= 2) prev_normal(g, R, T, ind, oid, gp, gnorm);
        bool cosw = (gnorm != 0. && A.L <= 2);
        bool ok = false;
        double davg[3], ravg[3], coeff[9];
        for (int sr = 2; !ok && sr <= A.max_rs; sr++) {
            const int D = 2 * sr + 1, nc = D * D * D;
            // the 5^3 stencil is two chunks: each lane keeps its two points'
            // weights and data in registers for the second pass
            const bool keep = nc <= 2 * 64;
            double kw[2] = {-1, -1}, krv[2][3], kdp[2][3];
            // pass 1: weighted sums (extrap_helper.cc:49-62)
            double acc = 0;
            int npt = 0;
            if (keep) {
                // both chunks' weights and data first, so that their
                // gathers are in flight together, then the two summations
#pragma unroll
                for (int ch = 0; ch < 2; ch++) {
                    int k = ch * 64 + lane;
                    double w = -1;
                    int p = 0, q = 0, r = 0;
                    if (k < nc) {
                        p = k % D - sr; q = (k / D) % D - sr; r = k / (D * D) - sr;
                        w = nb_weight(g, W, ind, ii, jj, kk, p, q, r, A.L, cosw, gp, A.wpow);
                    }
                    kw[ch] = w;
                    for (int a = 0; a < 3; a++) krv[ch][a] = kdp[ch][a] = 0;
                    if (w >= 0) {
                        int nb = ind + p + g.m4 * q + g.mn4 * r;
                        krv[ch][0] = W.val[0][nb]; krv[ch][1] = W.val[1][nb]; krv[ch][2] = W.val[2][nb];
                        kdp[ch][0] = g.lx[ii + p]; kdp[ch][1] = g.ly[jj + q]; kdp[ch][2] = g.lz[kk + r];
                    }
                }
#pragma unroll
                for (int ch = 0; ch < 2; ch++) {
                    const int k0 = ch * 64;
                    if (k0 >= nc) break;
                    const double w = kw[ch];
                    bool v = w >= 0;
                    if (v) {
                        term[0][lane] = krv[ch][0] * w;
                        term[1][lane] = krv[ch][1] * w;
                        term[2][lane] = krv[ch][2] * w;
                        term[3][lane] = kdp[ch][0] * w;
                        term[4][lane] = kdp[ch][1] * w;
                        term[5][lane] = kdp[ch][2] * w;
                        term[6][lane] = w;
                    } else {
                        for (int a = 0; a < LSQ_T1; a++) term[a][lane] = 0.0;
                    }
                    npt += __popcll(__ballot(v));
                    __syncthreads();
                    if (lane < LSQ_T1) {
                        int ne = min(64, nc - k0);
#pragma unroll 8
                        for (int e = 0; e < ne; e++) acc += term[lane][e];
                    }
                    __syncthreads();
                }
            } else {
                for (int k0 = 0; k0 < nc; k0 += 64) {
                    int k = k0 + lane;
                    double w = -1;
                    int p = 0, q = 0, r = 0;
                    if (k < nc) {
                        p = k % D - sr; q = (k / D) % D - sr; r = k / (D * D) - sr;
                        w = nb_weight(g, W, ind, ii, jj, kk, p, q, r, A.L, cosw, gp, A.wpow);
                    }
                    bool v = w >= 0;
                    if (v) {
                        int nb = ind + p + g.m4 * q + g.mn4 * r;
                        double rv[3] = {W.val[0][nb], W.val[1][nb], W.val[2][nb]};
                        double dp[3] = {g.lx[ii + p], g.ly[jj + q], g.lz[kk + r]};
                        term[0][lane] = rv[0] * w;
                        term[1][lane] = rv[1] * w;
                        term[2][lane] = rv[2] * w;
                        term[3][lane] = dp[0] * w;
                        term[4][lane] = dp[1] * w;
                        term[5][lane] = dp[2] * w;
                        term[6][lane] = w;
                    } else {
                        for (int a = 0; a < LSQ_T1; a++) term[a][lane] = 0.0;
                    }
                    npt += __popcll(__ballot(v));
                    __syncthreads();
                    if (lane < LSQ_T1) {
                        int ne = min(64, nc - k0);
#pragma unroll 8
                        for (int e = 0; e < ne; e++) acc += term[lane][e];
                    }
                    __syncthreads();
                }
            }
            if (npt == 0) continue;
            if (lane < LSQ_T1) bc[lane] = acc;
            __syncthreads();
            double wsum = bc[6];
            for (int a = 0; a < 3; a++) { ravg[a] = bc[a]; davg[a] = bc[3 + a]; }
            __syncthreads();
            for (int a = 0; a < 3; a++) ravg[a] /= wsum;
            for (int a = 0; a < 3; a++) davg[a] /= wsum;
            // pass 2: centred moments (extrap_helper.cc:68-75, 129-146)
            acc = 0;
            for (int k0 = 0, ch = 0; k0 < nc; k0 += 64, ch++) {
                int k = k0 + lane;
                double Q = -1;
                double rv[3] = {0, 0, 0}, dpv[3] = {0, 0, 0};
                if (keep) {
                    Q = ch ? kw[1] : kw[0];
                    for (int a = 0; a < 3; a++) {
                        rv[a] = ch ? krv[1][a] : krv[0][a];
                        dpv[a] = ch ? kdp[1][a] : kdp[0][a];
                    }
                } else if (k < nc) {
                    int p = k % D - sr, q = (k / D) % D - sr, r = k / (D * D) - sr;
                    Q = nb_weight(g, W, ind, ii, jj, kk, p, q, r, A.L, cosw, gp, A.wpow);
                    if (Q >= 0) {
                        int nb = ind + p + g.m4 * q + g.mn4 * r;
                        for (int a = 0; a < 3; a++) rv[a] = W.val[a][nb];
                        dpv[0] = g.lx[ii + p]; dpv[1] = g.ly[jj + q]; dpv[2] = g.lz[kk + r];
                    }
                }
                bool v = Q >= 0;
                if (v) {
                    double dp0 = dpv[0], dp1 = dpv[1], dp2 = dpv[2];
                    term[0][lane] = Q * (dp0 - davg[0]) * (dp0 - davg[0]);
                    term[1][lane] = Q * (dp1 - davg[1]) * (dp1 - davg[1]);
                    term[2][lane] = Q * (dp2 - davg[2]) * (dp2 - davg[2]);
                    term[3][lane] = Q * (dp0 - davg[0]) * (dp1 - davg[1]);
                    term[4][lane] = Q * (dp0 - davg[0]) * (dp2 - davg[2]);
                    term[5][lane] = Q * (dp1 - davg[1]) * (dp2 - davg[2]);
                    double dd[3] = {dp0 - davg[0], dp1 - davg[1], dp2 - davg[2]};
                    for (int i = 0; i < 3; i++) {
                        double lr = Q * (rv[i] - ravg[i]);
                        for (int j = 0; j < 3; j++) term[6 + 3 * i + j][lane] = dd[j] * lr;
                    }
                } else {
                    for (int a = 0; a < LSQ_T2; a++) term[a][lane] = 0.0;
                }
                __syncthreads();
                if (lane < LSQ_T2) {
                    int ne = min(64, nc - k0);
#pragma unroll 8
                    for (int e = 0; e < ne; e++) acc += term[lane][e];
                }
                __syncthreads();
            }
            if (lane < LSQ_T2) bc[lane] = acc;
            __syncthreads();
            double am[6], rhs[9];
            for (int a = 0; a < 6; a++) am[a] = bc[a];
            for (int a = 0; a < 9; a++) rhs[a] = bc[6 + a];
            __syncthreads();
            // mat3::inverse (mat3.hh:61-67) and extrap_helper::dot
            double a11 = am[0], a12 = am[3], a13 = am[4], a21 = am[3], a22 = am[1], a23 = am[5], a31 = am[4],
                   a32 = am[5], a33 = am[2];
            double det = a11 * (a22 * a33 - a23 * a32) + a21 * (a13 * a32 - a12 * a33) + a31 * (a12 * a23 - a13 * a22);
            double id = 1 / det;
            double Ai[9] = {(a22 * a33 - a23 * a32) * id, (a13 * a32 - a12 * a33) * id, (a12 * a23 - a13 * a22) * id,
                            (a23 * a31 - a21 * a33) * id, (a11 * a33 - a13 * a31) * id, (a13 * a21 - a11 * a23) * id,
                            (a21 * a32 - a22 * a31) * id, (a12 * a31 - a11 * a32) * id, (a11 * a22 - a12 * a21) * id};
            if (fabs(det) < 1e-17) continue;
            for (int i = 0; i < 3; i++) {
                int x = 3 * i, y = x + 1, z = x + 2;
                coeff[x] = Ai[0] * rhs[x] + Ai[1] * rhs[y] + Ai[2] * rhs[z];
                coeff[y] = Ai[3] * rhs[x] + Ai[4] * rhs[y] + Ai[5] * rhs[z];
                coeff[z] = Ai[6] * rhs[x] + Ai[7] * rhs[y] + Ai[8] * rhs[z];
            }
            ok = true;
        }
        if (lane != 0) continue;
        if (!ok) { atomicOr(status, ST_FATAL_EXTRAP); continue; }
        // fluid_3d::transfer (fluid_3d_ex.cc:293-313)
        double ps[3] = {g.lx[ii] - davg[0], g.ly[jj] - davg[1], g.lz[kk] - davg[2]};
        double out[3];
        for (int a = 0; a < 3; a++) {
            out[a] = ravg[a];
            for (int b = 0; b < 3; b++) out[a] += coeff[3 * a + b] * ps[b];
        }
        if (isnan(out[0]) || isnan(out[1]) || isnan(out[2])) continue;
        for (int a = 0; a < 3; a++) W.val[a][ind] = out[a];
        W.lid[ind] = (signed char)A.L;
    }
}

// communicate_a_layer on one rank: a periodic ghost takes its source's map of
// this layer (fluid_3d_comm.cc:203-280)
__global__ void k_wrap_layer(Geo g, Scratch W, int L) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= g.Ng) return;
    int i, j, k;
    unpack(g, t, i, j, k);
    int s = psrc(g, i, j, k);
    if (s < 0 || W.lid[s] != L) return;
    W.lid[t] = (signed char)L;
    for (int a = 0; a < 3; a++) W.val[a][t] = W.val[a][s];
}

// the object's layers become staged band entries
__global__ __launch_bounds__(1024) void k_stage_obj(Geo g, Scratch W, uint32_t oid, int *status) {
    __shared__ int sh[17];
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    int L = t < g.Ng ? W.lid[t] : 0;
    int p = block_append(L >= 1, W.st_n, sh);
    if (L < 1) return;
    if (p >= W.st_cap) { atomicOr(status, ST_BAND_OVERFLOW); return; }
    W.st_cell[p] = t;
    W.st_c[p] = oid | ((uint32_t)L << LDIG);
    for (int a = 0; a < 3; a++) { W.st_x[a][p] = W.val[a][t]; W.st_xpred[a][p] = W.val[a][t]; }
}

// ---- per-object boxes ----
// Every per-object pass of the extrapolation touches only the object's
// primary cells dilated by the layers and the search radius: the BFS marks
// cells within nlayers of the primaries, a list cell reads masks one cell
// further and least-squares data max_rs cells further, and prev_normal one
// more.  So each object's passes run over the bounding box of its primary
// cells (padded index space) grown by nlayers + max_rs + 2 and clamped to
// the padded grid, instead of over the whole grid.  Outside the box the
// scratch keeps whatever earlier objects left, which no pass of this object
// reads: masks there are below this object's mask values, and flags are
// zero (every flag a pass sets inside the box is cleared by the compaction
// in the box).  In a periodic dimension where the grown box reaches the ghost
// frame or the strip that feeds the opposite ghost frame, the box takes the
// whole padded extent, so that the periodic copies find their sources and
// targets inside it.  bx[8 o .. 8 o + 7] = lo[3], ext[3], volume.
constexpr int BOX_INTS = 8;

__global__ void k_box_init(int *bb, int n_obj) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n_obj) return;
    for (int d = 0; d < 3; d++) { bb[BOX_INTS * t + d] = INT_MAX; bb[BOX_INTS * t + 3 + d] = -1; }
}

// index extremes of each object's primary cells: a workgroup reduces its
// cells in LDS, then adds one global atomic per extreme and object it saw
__global__ __launch_bounds__(256) void k_box_reduce(Geo g, RMap R, int *bb, int n_obj) {
    __shared__ int sm[6 * MAXOBJ];
    for (int e = threadIdx.x; e < 6 * n_obj; e += blockDim.x) sm[e] = (e % 6) < 3 ? INT_MAX : -1;
    __syncthreads();
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < g.Ng; t += gridDim.x * blockDim.x) {
        uint32_t c = R.c[t];
        if (c >= 1 && c <= (uint32_t)n_obj) {
            int ijk[3];
            unpack(g, t, ijk[0], ijk[1], ijk[2]);
            int *o = sm + 6 * (c - 1);
            for (int d = 0; d < 3; d++) {
                atomicMin(o + d, ijk[d]);
                atomicMax(o + 3 + d, ijk[d]);
            }
        }
    }
    __syncthreads();
    for (int e = threadIdx.x; e < 6 * n_obj; e += blockDim.x) {
        int ob = e / 6, d = e % 6;
        if (sm[6 * ob + 3] < 0) continue;
        int *dst = bb + BOX_INTS * ob + d;
        if (d < 3) atomicMin(dst, sm[e]);
        else atomicMax(dst, sm[e]);
    }
}

__global__ void k_box_finish(Geo g, int *bb, int n_obj, int grow) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n_obj) return;
    int *b = bb + BOX_INTS * t;
    if (b[3] < 0) {   // no primary cell: empty box
        for (int e = 0; e < BOX_INTS; e++) b[e] = 0;
        return;
    }
    const int s[3] = {g.m, g.n, g.o}, d4[3] = {g.m4, g.n4, g.o4};
    int lo[3], ext[3];
    for (int d = 0; d < 3; d++) {
        int l = b[d] - grow, h = b[3 + d] + grow;
        if (g.prd[d] && (l <= 3 || h >= s[d])) { l = 0; h = d4[d] - 1; }
        l = l < 0 ? 0 : l;
        h = h > d4[d] - 1 ? d4[d] - 1 : h;
        lo[d] = l;
        ext[d] = h - l + 1;
    }
    for (int d = 0; d < 3; d++) { b[d] = lo[d]; b[3 + d] = ext[d]; }
    b[6] = ext[0] * ext[1] * ext[2];
    b[7] = 0;
}

// A batch is a set of objects whose boxes are pairwise disjoint.  Their
// extrapolations touch disjoint cells of the scratch (each stays in its own
// box) and read nothing the others write, so one sequence of launches runs
// all of them, each cell knowing its object from the box it lies in.  The
// mask tags of a pass are shared by the batch's objects; a tag only has to
// differ from every earlier tag within the object's own box.
// A cell of the batch: work-item t of [0, total) -> (box b, padded index)
__device__ __forceinline__ int batch_cell(const Geo &g, const BoxBatch *__restrict__ B, int t, int &b) {
    if (B->n == 1) {   // one box: every index below is uniform
        b = 0;
        int e0 = B->ext[0][0], e1 = B->ext[0][1];
        int q = t / e0;
        int i = t - q * e0;
        int k = q / e1;
        int j = q - k * e1;
        return IXb(g, B->lo[0][0] + i, B->lo[0][1] + j, B->lo[0][2] + k);
    }
    int lo = 0, hi = B->n - 1;
    while (lo < hi) {   // the last box whose start is <= t
        int mid = (lo + hi + 1) >> 1;
        if (B->start[mid] <= t) lo = mid;
        else hi = mid - 1;
    }
    b = lo;
    int r = t - B->start[lo];
    int e0 = B->ext[lo][0], e1 = B->ext[lo][1];
    int q = r / e0;
    int i = r - q * e0;
    int k = q / e1;
    int j = q - k * e1;
    return IXb(g, B->lo[lo][0] + i, B->lo[lo][1] + j, B->lo[lo][2] + k);
}

#define BATCH_LOOP(B)                                                                      \
    const int tot_ = (B)->total;                                                           \
    for (int t_ = blockIdx.x * blockDim.x + threadIdx.x; t_ < tot_; t_ += gridDim.x * blockDim.x)

// also clears the counter of the first layer's cell list
__global__ void k_batch_init(Geo g, RMap R, Scratch W, const BoxBatch *__restrict__ B, uint32_t nmask, int *cnt0) {
    if (blockIdx.x == 0 && threadIdx.x == 0) *cnt0 = 0;
    BATCH_LOOP(B) {
        int b;
        int t = batch_cell(g, B, t_, b);
        if (R.c[t] == (uint32_t)B->oid[b]) {
            W.lid[t] = 0;
            for (int a = 0; a < 3; a++) W.val[a][t] = R.xpred[a][t];
            W.mask[t] = nmask;
        } else {
            W.lid[t] = -1;
        }
        W.flag[t] = 0;
    }
}
__global__ void k_batch_first_push(Geo g, Scratch W, const BoxBatch *__restrict__ B, uint32_t cur_mask) {
    BATCH_LOOP(B) {
        int b;
        int ind = batch_cell(g, B, t_, b);
        if (within_inner_ghost(g, ind) && W.mask[ind] == cur_mask) push6(g, W, ind, cur_mask);
    }
}
__global__ void k_batch_apply_flags(Geo g, Scratch W, const BoxBatch *__restrict__ B, uint32_t nmask) {
    BATCH_LOOP(B) {
        int b;
        int t = batch_cell(g, B, t_, b);
        if (W.flag[t]) W.mask[t] = nmask;
    }
}
__global__ void k_batch_comm_mask(Geo g, Scratch W, const BoxBatch *__restrict__ B) {
    BATCH_LOOP(B) {
        int b;
        int t = batch_cell(g, B, t_, b);
        int i, j, k;
        unpack(g, t, i, j, k);
        int s = psrc(g, i, j, k);
        if (s >= 0) W.mask[t] = W.mask[s];
    }
}
__global__ void k_batch_wrap_layer(Geo g, Scratch W, const BoxBatch *__restrict__ B, int L) {
    BATCH_LOOP(B) {
        int b;
        int t = batch_cell(g, B, t_, b);
        int i, j, k;
        unpack(g, t, i, j, k);
        int s = psrc(g, i, j, k);
        if (s < 0 || W.lid[s] != L) continue;
        W.lid[t] = (signed char)L;
        for (int a = 0; a < 3; a++) W.val[a][t] = W.val[a][s];
    }
}
// The next layer's cell list: the flagged cells, appended and unflagged.
// The workgroup-uniform loops keep every thread at each block_append.  Each
// cell's decision reads only its own state, so the passes before it fold
// in: APPLY (no periodic dimension) first gives each flagged cell the new
// mask tag (as k_batch_apply_flags does; without periodic images a ghost
// cell carries the new tag only if it was flagged), and SCAN (periodic)
// first flags the ghost cells that carry the tag after the periodic mask
// copy.
template <bool APPLY, bool SCAN>
__global__ __launch_bounds__(1024) void k_batch_compact(Geo g, Scratch W, const BoxBatch *__restrict__ B, int *list,
                                                        int *cnt, uint32_t nmask) {
    __shared__ int sh[17];
    const int tot = B->total;
    for (int b0 = blockIdx.x * blockDim.x; b0 < tot; b0 += gridDim.x * blockDim.x) {
        int tt = b0 + threadIdx.x, b;
        int t = tt < tot ? batch_cell(g, B, tt, b) : 0;
        bool f = tt < tot && W.flag[t] != 0;
        if (SCAN && tt < tot && !f) {
            int i, j, k;
            unpack(g, t, i, j, k);
            bool ghost = (i < 2 || i > g.m + 1 || j < 2 || j > g.n + 1 || k < 2 || k > g.o + 1);
            f = ghost && W.mask[t] == nmask;
        }
        int p = block_append(f, cnt, sh);
        if (f) {
            if (APPLY) W.mask[t] = nmask;
            list[p] = t;
            W.flag[t] = 0;
        }
    }
}
__global__ __launch_bounds__(1024) void k_batch_stage(Geo g, Scratch W, const BoxBatch *__restrict__ B, int *status) {
    __shared__ int sh[17];
    const int tot = B->total;
    for (int b0 = blockIdx.x * blockDim.x; b0 < tot; b0 += gridDim.x * blockDim.x) {
        int tt = b0 + threadIdx.x, b = 0;
        int t = tt < tot ? batch_cell(g, B, tt, b) : 0;
        int L = tt < tot ? W.lid[t] : 0;
        int p = block_append(L >= 1, W.st_n, sh);
        if (L < 1) continue;
        if (p >= W.st_cap) { atomicOr(status, ST_BAND_OVERFLOW); continue; }
        W.st_cell[p] = t;
        W.st_c[p] = (uint32_t)B->oid[b] | ((uint32_t)L << LDIG);
        for (int a = 0; a < 3; a++) { W.st_x[a][p] = W.val[a][t]; W.st_xpred[a][p] = W.val[a][t]; }
    }
}

// ---- band build: count, scan, place, per-cell order, gather ----
__global__ void k_count_cells(Geo g, Band B, Scratch W) {
    int n = *W.st_n;
    if (n > W.st_cap) n = W.st_cap;
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < n; t += gridDim.x * blockDim.x)
        atomicAdd(&B.cnt[W.st_cell[t]], 1);
}

// exclusive scan of B.cnt into B.start, total into *B.n (three phases)
constexpr int SCAN_BS = 1024;
// exclusive scan of SCAN_BS counts per workgroup: each wavefront scans its
// 64 with lane shuffles, one wavefront scans the wavefront totals, and two
// barriers join them (integer sums: the same values as any order)
__global__ void k_scan_blocks(const int *in, int *out, int *bsum, int n) {
    constexpr int NW = SCAN_BS / 64;
    __shared__ int ws[NW];
    const int t = blockIdx.x * SCAN_BS + threadIdx.x;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int v = (t < n) ? in[t] : 0;
    int incl = v;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const int u = __shfl_up(incl, off, 64);
        if (lane >= off) incl += u;
    }
    if (lane == 63) ws[w] = incl;
    __syncthreads();
    if (w == 0) {
        const int s0 = lane < NW ? ws[lane] : 0;
        int si = s0;
#pragma unroll
        for (int off = 1; off < NW; off <<= 1) {
            const int u = __shfl_up(si, off, 64);
            if (lane >= off) si += u;
        }
        if (lane < NW) ws[lane] = si - s0;
    }
    __syncthreads();
    const int excl = incl - v + ws[w];
    if (t < n) out[t] = excl;
    if (threadIdx.x == SCAN_BS - 1) bsum[blockIdx.x] = excl + v;
}
// exclusive scan of the block sums in one workgroup of SCAN_BS threads, each
// thread owning a contiguous run of ceil(nb / SCAN_BS) sums
// cap_status (when not null): flags ST_BAND_OVERFLOW there if the total
// exceeds cap
__global__ void k_scan_sums(int *bsum, int nb, int *total, int cap = 0, int *cap_status = nullptr) {
    __shared__ int sh[SCAN_BS];
    int per = (nb + SCAN_BS - 1) / SCAN_BS;
    int b0 = threadIdx.x * per, b1 = min(b0 + per, nb);
    int loc = 0;
    for (int b = b0; b < b1; b++) loc += bsum[b];
    sh[threadIdx.x] = loc;
    __syncthreads();
    for (int off = 1; off < SCAN_BS; off <<= 1) {
        int add = (threadIdx.x >= (unsigned)off) ? sh[threadIdx.x - off] : 0;
        __syncthreads();
        sh[threadIdx.x] += add;
        __syncthreads();
    }
    int run = sh[threadIdx.x] - loc;
    for (int b = b0; b < b1; b++) {
        int v = bsum[b];
        bsum[b] = run;
        run += v;
    }
    if (threadIdx.x == SCAN_BS - 1) {
        *total = sh[SCAN_BS - 1];
        if (cap_status && sh[SCAN_BS - 1] > cap) atomicOr(cap_status, ST_BAND_OVERFLOW);
    }
}
__global__ void k_scan_add(int *out, const int *bsum, int n) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) out[t] += bsum[t / SCAN_BS];
}

// exclusive scan of n ints (n <= SCAN_BS^2 * 64); bsum holds n/SCAN_BS + 1
// ints of scratch, *total the sum
void launch_exclusive_scan(hipStream_t s, const int *in, int *out, int *bsum, int n, int *total) {
    int nb = (n + SCAN_BS - 1) / SCAN_BS;
    hipLaunchKernelGGL(k_scan_blocks, dim3(nb), dim3(SCAN_BS), 0, s, in, out, bsum, n);
    hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(SCAN_BS), 0, s, bsum, nb, total);
    hipLaunchKernelGGL(k_scan_add, dim3((n + 255) / 256), dim3(256), 0, s, out, (const int *)bsum, n);
}

__global__ void k_place(Geo g, Band B, Scratch W) {
    int n = *W.st_n;
    if (n > W.st_cap) n = W.st_cap;
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < n; t += gridDim.x * blockDim.x) {
        int c = W.st_cell[t];
        // the cell's start: its in-block prefix plus its block's offset
        // (k_sort_cells writes the sum back into B.start)
        int slot = B.start[c] + W.scan_tmp[c / SCAN_BS] + atomicAdd(&W.fill[c], 1);
        if (slot < B.cap) W.perm[slot] = t;
    }
}

// multimap order inside a cell: object order, a redrawn primary last.  Also
// completes the scan: B.start gets its block's offset added.
__global__ void k_sort_cells(Geo g, Band B, Scratch W, int nlayers) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= g.Ng) return;
    const int s = B.start[t] + W.scan_tmp[t / SCAN_BS];
    B.start[t] = s;
    int n = B.cnt[t];
    if (n < 2) return;
    if (s + n > B.cap) return;
    uint32_t moved = (uint32_t)(nlayers + 1);
    for (int a = 1; a < n; a++) {
        int v = W.perm[s + a];
        uint32_t cv = W.st_c[v];
        uint32_t kv = ((cv >> LDIG) == moved ? 4096u : 0u) + (cv & OMASK);
        int b = a - 1;
        while (b >= 0) {
            uint32_t cb = W.st_c[W.perm[s + b]];
            uint32_t kb = ((cb >> LDIG) == moved ? 4096u : 0u) + (cb & OMASK);
            if (kb <= kv) break;
            W.perm[s + b + 1] = W.perm[s + b];
            b--;
        }
        W.perm[s + b + 1] = v;
    }
}

__global__ void k_gather_entries(Band B, Scratch W, int carry) {
    int n = *B.n;
    if (n > B.cap) n = B.cap;
    for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < n; e += gridDim.x * blockDim.x) {
        int v = W.perm[e];
        B.cell[e] = W.st_cell[v];
        B.c[e] = W.st_c[v];
        for (int a = 0; a < 3; a++) { B.x[a][e] = W.st_x[a][v]; B.xpred[a][e] = W.st_xpred[a][v]; }
        for (int a = 0; a < 9; a++) B.gphi[a][e] = -1.e200;
        if (carry)
            for (int a = 0; a < 3; a++) { B.detF[a][e] = W.st_detF[a][v]; B.en[a][e] = W.st_en[a][v]; }
        else
            for (int a = 0; a < 3; a++) { B.detF[a][e] = -1.e200; B.en[a][e] = -1.e200; }
        for (int a = 0; a < 18; a++) B.fx[a][e] = 0;
    }
}

__global__ void k_zero_i32(int *p, int n) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) p[t] = 0;
}
__global__ void k_zero_u64(unsigned long long *p, int n) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) p[t] = 0;
}
// two int arrays of n zeroed in one launch
__global__ void k_zero2_i32(int *p, int *q, int n) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) {
        p[t] = 0;
        q[t] = 0;
    }
}

void launch_zero_i32(hipStream_t s, int *p, int n) {
    hipLaunchKernelGGL(k_zero_i32, dim3((n + 255) / 256), dim3(256), 0, s, p, n);
}
void launch_zero_u64(hipStream_t s, unsigned long long *p, int n) {
    hipLaunchKernelGGL(k_zero_u64, dim3((n + 255) / 256), dim3(256), 0, s, p, n);
}

static inline dim3 g1(int n) { return dim3((n + 255) / 256); }
static const int GRID_STRIDE_BLOCKS = 2048;

void band_build_from_staging(hipStream_t s, const Geo &g, Band B, Scratch &W, int nlayers, int *status, int carry) {
    hipLaunchKernelGGL(k_zero2_i32, g1(g.Ng), dim3(256), 0, s, B.cnt, W.fill, g.Ng);
    hipLaunchKernelGGL(k_count_cells, dim3(GRID_STRIDE_BLOCKS), dim3(256), 0, s, g, B, W);
    int nb = (g.Ng + SCAN_BS - 1) / SCAN_BS;
    hipLaunchKernelGGL(k_scan_blocks, dim3(nb), dim3(SCAN_BS), 0, s, B.cnt, B.start, W.scan_tmp, g.Ng);
    hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(SCAN_BS), 0, s, W.scan_tmp, nb, B.n, B.cap, status);
    hipLaunchKernelGGL(k_place, dim3(GRID_STRIDE_BLOCKS), dim3(256), 0, s, g, B, W);
    hipLaunchKernelGGL(k_sort_cells, g1(g.Ng), dim3(256), 0, s, g, B, W, nlayers);
    hipLaunchKernelGGL(k_gather_entries, dim3(GRID_STRIDE_BLOCKS), dim3(256), 0, s, B, W, carry);
}

// fluid_3d::extrapolate (fluid_3d_ex.cc:31-47); the caller has swapped the
// previous band into the temp slot.  nmask is the host copy of the mask
// counter (extrap_helper::nmask), advanced exactly as the reference does.
// one object: its batch is built on the device, with no host read-back
__global__ void k_batch_single(const int *bx, BoxBatch *B) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    B->n = 1;
    B->oid[0] = 1;
    for (int d = 0; d < 3; d++) { B->lo[0][d] = bx[d]; B->ext[0][d] = bx[3 + d]; }
    B->start[0] = 0;
    B->start[1] = bx[6];
    B->total = bx[6];
}

// Objects in index order go to the first batch whose boxes they do not
// overlap (a box spans [lo, lo + ext) in each dimension); empty boxes (no
// primary cell) take no part.
static int make_batches(const int *bx, int n_obj, BoxBatch *out) {
    int nb = 0;
    for (int o = 0; o < n_obj; o++) {
        const int *b = bx + BOX_INTS * o;
        if (b[6] <= 0) continue;
        int dst = -1;
        for (int q = 0; q < nb && dst < 0; q++) {
            bool clash = false;
            for (int e = 0; e < out[q].n && !clash; e++) {
                bool ov = true;
                for (int d = 0; d < 3; d++)
                    ov = ov && b[d] < out[q].lo[e][d] + out[q].ext[e][d] && out[q].lo[e][d] < b[d] + b[3 + d];
                clash = ov;
            }
            if (!clash) dst = q;
        }
        if (dst < 0) {
            dst = nb++;
            out[dst].n = 0;
            out[dst].start[0] = 0;
        }
        BoxBatch &q = out[dst];
        int e = q.n++;
        q.oid[e] = o + 1;
        for (int d = 0; d < 3; d++) { q.lo[e][d] = b[d]; q.ext[e][d] = b[3 + d]; }
        q.start[e + 1] = q.start[e] + b[6];
        q.total = q.start[e + 1];
    }
    return nb;
}

int band_extrapolate(hipStream_t s, const Geo &g, const Phys &P, const ObjC *O, RMap R, Band B, Scratch &W,
                     int nlayers, int max_rs, const double *wpow_dev, uint32_t &nmask, int *status, int n_obj) {
    (void)P; (void)O;
    launch_zero_i32(s, W.st_n, 1);
    bool per = g.prd[0] || g.prd[1] || g.prd[2];
    // temp band needed by prev_normal is passed in B's place by the caller
    // through W; see fluid3d.hip
    (void)B;
    if (n_obj < 1) return 0;
    // the objects' boxes, from their primary cells, then (on the host) the
    // batches of objects with disjoint boxes
    hipLaunchKernelGGL(k_box_init, dim3((n_obj + 63) / 64), dim3(64), 0, s, W.box, n_obj);
    hipLaunchKernelGGL(k_box_reduce, dim3(std::min(1024, (g.Ng + 255) / 256)), dim3(256), 0, s, g, R, W.box, n_obj);
    hipLaunchKernelGGL(k_box_finish, dim3((n_obj + 63) / 64), dim3(64), 0, s, g, W.box, n_obj, nlayers + max_rs + 2);
    int nb = 1;
    if (n_obj == 1) {
        hipLaunchKernelGGL(k_batch_single, dim3(1), dim3(64), 0, s, (const int *)W.box, W.batch);
    } else {
        if (hipMemcpyAsync(W.h_box, W.box, sizeof(int) * BOX_INTS * n_obj, hipMemcpyDeviceToHost, s) != hipSuccess ||
            hipStreamSynchronize(s) != hipSuccess)
            return -1;
        nb = make_batches(W.h_box, n_obj, W.h_batch);
        if (nb == 0) return 0;
        if (hipMemcpyAsync(W.batch, W.h_batch, sizeof(BoxBatch) * nb, hipMemcpyHostToDevice, s) != hipSuccess)
            return -1;
    }
    const dim3 ge(std::min(1024, (g.Ng + 255) / 256)), gc(std::min(256, (g.Ng + 1023) / 1024));
    for (int q = 0; q < nb; q++) {
        const BoxBatch *bb = W.batch + q;
        if (nmask > UINT_MAX - 50) {
            nmask = 1;
            hipLaunchKernelGGL(k_mask_reset, g1(g.Ng), dim3(256), 0, s, g, W);
        }
        uint32_t cur_mask = nmask;
        int cur = 0;
        hipLaunchKernelGGL(k_batch_init, ge, dim3(256), 0, s, g, R, W, bb, nmask, W.cnt + cur);
        nmask++;
        hipLaunchKernelGGL(k_batch_first_push, ge, dim3(256), 0, s, g, W, bb, cur_mask);
        // flags -> mask tag, periodic mask copy, ghost scan, compaction
        auto next_list = [&](int *list, int *cnt) {
            if (per) {
                hipLaunchKernelGGL(k_batch_apply_flags, ge, dim3(256), 0, s, g, W, bb, nmask);
                hipLaunchKernelGGL(k_batch_comm_mask, ge, dim3(256), 0, s, g, W, bb);
                hipLaunchKernelGGL((k_batch_compact<false, true>), gc, dim3(1024), 0, s, g, W, bb, list, cnt, nmask);
            } else {
                hipLaunchKernelGGL((k_batch_compact<true, false>), gc, dim3(1024), 0, s, g, W, bb, list, cnt, nmask);
            }
        };
        next_list(W.list[cur], W.cnt + cur);
        for (int L = 1; L <= nlayers; L++) {
            nmask++;
            LsqArgs A;
            A.L = L; A.max_rs = max_rs; A.wpow = wpow_dev;
            int nxt = cur ^ 1;
            hipLaunchKernelGGL(k_gather_lsq, dim3(8192), dim3(64), 0, s, g, R, W.tband, W,
                               W.list[cur], W.cnt + cur, A, bb, status, W.cnt + nxt);
            if (L < nlayers) {
                hipLaunchKernelGGL(k_list_push, dim3(GRID_STRIDE_BLOCKS), dim3(256), 0, s, g, W, W.list[cur],
                                   W.cnt + cur, cur_mask);
                next_list(W.list[nxt], W.cnt + nxt);
            }
            cur = nxt;
            if (per) hipLaunchKernelGGL(k_batch_wrap_layer, ge, dim3(256), 0, s, g, W, bb, L);
        }
        hipLaunchKernelGGL(k_batch_stage, gc, dim3(1024), 0, s, g, W, bb, status);
    }
    return nb;
}

// ---- predictor, corrector, full update ----
// update_reference_map predictor (fluid_3d.cc:1153-1192)
// flat grid over the m x n x o interior, x fastest
__global__ void k_xpred(Geo g, Phys P, Fluid U, RMap R) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= g.m * g.n * g.o) return;
    const int row = t / g.m, i = t - row * g.m;
    const int k = row / g.n, j = row - k * g.n;
    int c = IXb(g, i + 2, j + 2, k + 2);
    if ((R.c[c] & OMASK) == 0) return;
    double cdt = P.dt;
    double av[3];
    av[0] = -0.5 * cdt * (U.fvel[0][c] + U.fvel[3][c]) * g.dxsp;
    av[1] = -0.5 * cdt * (U.fvel[7][c] + U.fvel[10][c]) * g.dysp;
    av[2] = -0.5 * cdt * (U.fvel[14][c] + U.fvel[17][c]) * g.dzsp;
    for (int d = 0; d < 3; d++) {
        double dxt = av[0] * (R.fx[3 + d][c] - R.fx[d][c]) + av[1] * (R.fx[9 + d][c] - R.fx[6 + d][c]) +
                     av[2] * (R.fx[15 + d][c] - R.fx[12 + d][c]);
        R.xpred[d][c] = R.x[d][c] + dxt;
    }
}
void launch_xpred(hipStream_t s, const Geo &g, const Phys &P, const Fluid &U, RMap R) {
    hipLaunchKernelGGL(k_xpred, dim3((g.m * g.n * g.o + 255) / 256), dim3(256), 0, s, g, P, U, R);
}

// corrector (fluid_3d.cc:1205-1248) over every padded cell
__global__ void k_corrector(Geo g, Phys P, const ObjC *O, RMap R, Band B, Band T, int *err) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= g.Ng) return;
    if (R.c[t] & OMASK)
        for (int a = 0; a < 3; a++) R.x[a][t] = 0.5 * (R.x[a][t] + R.xpred[a][t]);
    int s = B.start[t], n = B.cnt[t];
    for (int e = s; e < s + n; e++) {
        uint32_t oid = B.c[e] & OMASK;
        int old = -1;
        int ts = T.start[t], tn = T.cnt[t];
        for (int b = ts; b < ts + tn; b++)
            if ((T.c[b] & OMASK) == oid) { old = b; break; }
        if (old >= 0) {
            for (int q = 0; q < 3; q++) B.x[q][e] = 0.5 * (T.x[q][old] + B.xpred[q][e]);
        } else {
            double dist = obj_phi(O, P.n_obj, (int)oid - 1, B.xpred[0][e], B.xpred[1][e], B.xpred[2][e]);
            if (dist < P.eps && (B.c[e] >> LDIG) <= 3) atomicAdd(err, 1);
            else
                for (int q = 0; q < 3; q++) B.x[q][e] = B.xpred[q][e];
        }
    }
}
void launch_corrector(hipStream_t s, const Geo &g, const Phys &P, const ObjC *O, RMap R, Band B, Band T, int *err) {
    hipLaunchKernelGGL(k_corrector, g1(g.Ng), dim3(256), 0, s, g, P, O, R, B, T, err);
}

// revert_reference_map_update (fluid_3d.cc:1257-1282)
__global__ void k_revert(Geo g, RMap R, Band B, Band T) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= g.Ng) return;
    if (R.c[t] & OMASK)
        for (int a = 0; a < 3; a++) R.x[a][t] = 2 * R.x[a][t] - R.xpred[a][t];
    int s = B.start[t], n = B.cnt[t];
    for (int e = s; e < s + n; e++) {
        uint32_t oid = B.c[e] & OMASK;
        int ts = T.start[t], tn = T.cnt[t];
        for (int b = ts; b < ts + tn; b++)
            if ((T.c[b] & OMASK) == oid) {
                for (int q = 0; q < 3; q++) B.x[q][e] = T.x[q][b];
                break;
            }
    }
}
void launch_revert(hipStream_t s, const Geo &g, RMap R, Band B, Band T) {
    hipLaunchKernelGGL(k_revert, g1(g.Ng), dim3(256), 0, s, g, R, B, T);
}

// update_reference_map_full (fluid_3d.cc:1285-1306)
__global__ void k_full_prim(Geo g, RMap R) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= g.Ng) return;
    if ((R.c[t] & OMASK) > 0)
        for (int a = 0; a < 3; a++) R.x[a][t] = R.xpred[a][t];
}
__global__ void k_band_x_eq_xpred(Band B) {
    int n = *B.n;
    if (n > B.cap) n = B.cap;
    for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < n; e += gridDim.x * blockDim.x)
        for (int a = 0; a < 3; a++) B.x[a][e] = B.xpred[a][e];
}
void launch_full_update(hipStream_t s, const Geo &g, RMap R, Band B) {
    hipLaunchKernelGGL(k_full_prim, g1(g.Ng), dim3(256), 0, s, g, R);
    hipLaunchKernelGGL(k_band_x_eq_xpred, dim3(GRID_STRIDE_BLOCKS), dim3(256), 0, s, B);
}
void launch_band_x_eq_xpred(hipStream_t s, Band B) {
    hipLaunchKernelGGL(k_band_x_eq_xpred, dim3(GRID_STRIDE_BLOCKS), dim3(256), 0, s, B);
}

// ---- redraw_boundary (fluid_3d_ex.cc:336-394) ----
constexpr int RD_MAX = 32;
constexpr int RD_LIGHT = 4;    // list entries kept in registers
// One cell's redraw.  Its list (the band maps, then the primary if it
// leaves) holds up to CAP entries; with UNR = CAP every loop over it is
// unrolled to CAP with guards on the list length, so the list lives in
// registers (a private array indexed at run time lives in scratch memory,
// whose per-wavefront allocation caps how many wavefronts run).  k_redraw
// takes the cells with at most RD_LIGHT - 1 band maps that way and lists
// the others, which k_redraw_heavy takes with CAP = RD_MAX.  Every work-item
// of k_redraw reaches the staging append at the end, where a wavefront
// reserves its cells' slots with one atomic (the counts scanned across the
// wavefront).  The staging order does not matter: the band build sorts
// each cell's entries.
template <int CAP, int UNR, bool HEAVY>
__device__ __forceinline__ void redraw_cell(const Geo &g, const Phys &P, const ObjC *O, const RMap &R, const Band &B,
                                            const Scratch &W, int nlayers, int *status, int t, bool live) {
    bool fatal = false;
    uint32_t lc[CAP];
    double lx[CAP][3], lp[CAP][3];
    // where each listed map came from: its band entry, or -1 for this
    // cell's primary; the whole record moves in the reference
    // (fluid_3d_ex.cc:345-389), so face energies and det F go with it
    int lsrc[CAP];
    int nl = 0;
    int s = B.start[t], n = live ? B.cnt[t] : 0;
    if (n > CAP - 1) n = CAP - 1;
#pragma unroll UNR
    for (int q = 0; q < CAP - 1; q++) {
        if (q >= n) break;
        const int e = s + q;
        lc[q] = B.c[e];
        lsrc[q] = e;
        for (int a = 0; a < 3; a++) { lx[q][a] = B.x[a][e]; lp[q][a] = B.xpred[a][e]; }
        nl = q + 1;
    }
    double pen[3], pdet[3];
    uint32_t lflag = (uint32_t)(nlayers + 1) << LDIG;
    uint32_t pc = R.c[t];
    bool occupied = false, changed = false, cleared = false;
    int take = -1;   // band entry that became the primary
    int oidp = live ? (int)(pc & OMASK) : 0;
    // the primary map's values matter only where the cell has one (else they
    // are overwritten before any use), so a cell without one, most of the
    // grid, reads three words instead of nine
    double px[3] = {0, 0, 0}, pp[3] = {0, 0, 0};
    if (oidp > 0)
        for (int a = 0; a < 3; a++) { px[a] = R.x[a][t]; pp[a] = R.xpred[a][t]; }
    if (oidp > 0) {
        double pphi = obj_phi(O, P.n_obj, oidp - 1, px[0], px[1], px[2]);
        if (pphi > 0) {
#pragma unroll UNR
            for (int q = 0; q < CAP; q++)
                if (q == nl) {
                    lc[q] = lflag | pc;
                    lsrc[q] = -1;
                    for (int a = 0; a < 3; a++) { lx[q][a] = px[a]; lp[q][a] = pp[a]; }
                }
            for (int a = 0; a < 3; a++) { pen[a] = R.en[a][t]; pdet[a] = R.detF[a][t]; }
            nl++;
            pc = 0;
            for (int a = 0; a < 3; a++) { px[a] = 0; pp[a] = 0; }
            changed = true;
            cleared = true;
            atomicAdd(&P.counters[1], 1);
        } else {
            occupied = true;
        }
    }
    const int N = nl;
#pragma unroll UNR
    for (int l = 0; l < CAP; l++) {
        if (l >= N || fatal) break;
        if (l >= nl) continue;
        uint32_t oid = lc[l] & OMASK;
        double sphi = obj_phi(O, P.n_obj, (int)oid - 1, lx[l][0], lx[l][1], lx[l][2]);
        bool inside = (sphi < 0 && (lc[l] >> LDIG) == 1);
        if (!inside) continue;
        int po = (int)(pc & OMASK);
        double pphi = po > 0 ? obj_phi(O, P.n_obj, po - 1, px[0], px[1], px[2]) : __builtin_nan("");
        if ((occupied && sphi < pphi) || !occupied) {
            if (pphi < -g.dx) {
                atomicOr(status, ST_FATAL_REDRAW);
                fatal = true;
                break;
            }
            pc = oid;
            for (int a = 0; a < 3; a++) { px[a] = lx[l][a]; pp[a] = lp[l][a]; }
            take = lsrc[l];
            occupied = true;
            atomicAdd(&P.counters[2], 1);
            changed = true;
            cleared = false;
            // remove the first entry of that object, shifting the rest down
            int j0 = -1;
#pragma unroll UNR
            for (int b = 0; b < CAP; b++)
                if (b < nl && (lc[b] & OMASK) == oid && j0 == -1) j0 = b;
            if (j0 >= 0) {
#pragma unroll UNR
                for (int b = 0; b < CAP - 1; b++) {
                    if (b >= j0 && b < nl - 1) {
                        lc[b] = lc[b + 1];
                        lsrc[b] = lsrc[b + 1];
                        for (int a = 0; a < 3; a++) { lx[b][a] = lx[b + 1][a]; lp[b][a] = lp[b + 1][a]; }
                    }
                }
                nl--;
            }
        }
    }
    if (fatal) nl = 0;
    if (changed && !fatal) {
        R.c[t] = pc;
        for (int a = 0; a < 3; a++) { R.x[a][t] = px[a]; R.xpred[a][t] = pp[a]; }
        if (cleared) {
            for (int a = 0; a < 18; a++) R.fx[a][t] = 0;
            for (int a = 0; a < 9; a++) R.gphi[a][t] = 0;
            for (int a = 0; a < 3; a++) { R.detF[a][t] = 0; R.en[a][t] = 0; }
        } else if (take >= 0) {
            for (int a = 0; a < 3; a++) { R.detF[a][t] = B.detF[a][take]; R.en[a][t] = B.en[a][take]; }
        }
    }
    int base;
    if (HEAVY) {
        base = nl > 0 ? atomicAdd(W.st_n, nl) : 0;
    } else {
        // wavefront scan of the counts, one reservation per wavefront
        const int lane = __lane_id();
        int incl = nl;
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            int v = __shfl_up(incl, off, 64);
            if (lane >= off) incl += v;
        }
        const int total = __shfl(incl, 63, 64);
        int b0 = 0;
        if (lane == 63 && total > 0) b0 = atomicAdd(W.st_n, total);
        base = __shfl(b0, 63, 64) + incl - nl;
    }
#pragma unroll UNR
    for (int e = 0; e < CAP; e++) {
        if (e >= nl) break;
        int p = base + e;
        if (p >= W.st_cap) { atomicOr(status, ST_BAND_OVERFLOW); break; }
        W.st_cell[p] = t;
        W.st_c[p] = lc[e];
        for (int a = 0; a < 3; a++) { W.st_x[a][p] = lx[e][a]; W.st_xpred[a][p] = lp[e][a]; }
        int src = lsrc[e];
        for (int a = 0; a < 3; a++) {
            W.st_en[a][p] = src >= 0 ? B.en[a][src] : pen[a];
            W.st_detF[a][p] = src >= 0 ? B.detF[a][src] : pdet[a];
        }
    }
}

// The cells that hold a map (a primary one or band entries): redraw_cell
// does nothing on the others, most of the grid, so the register-heavy
// redraw runs over this list only.  Block 0 also clears the staging and
// heavy-list counters that the redraw launches use next.  A workgroup
// gathers the listed cells of RD_SCAN_CHUNKS x 256 cells in LDS and
// reserves their place in the list with one global atomic: one per 256
// cells serialised on the counter (24 us at 128^3).
constexpr int RD_SCAN_CHUNKS = 8;
__global__ __launch_bounds__(256) void k_redraw_scan(Geo g, RMap R, Band B, int *list, int *count, int *st_n,
                                                     int *hcount) {
    __shared__ int buf[256 * RD_SCAN_CHUNKS];
    __shared__ int nbuf, base;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        *st_n = 0;
        *hcount = 0;
    }
    if (threadIdx.x == 0) nbuf = 0;
    __syncthreads();
    bool f[RD_SCAN_CHUNKS];
#pragma unroll
    for (int r = 0; r < RD_SCAN_CHUNKS; r++) {
        const int t = (blockIdx.x * RD_SCAN_CHUNKS + r) * 256 + threadIdx.x;
        f[r] = t < g.Ng && ((R.c[t] & OMASK) != 0 || B.cnt[t] > 0);
    }
#pragma unroll
    for (int r = 0; r < RD_SCAN_CHUNKS; r++)
        if (f[r]) buf[atomicAdd(&nbuf, 1)] = (blockIdx.x * RD_SCAN_CHUNKS + r) * 256 + threadIdx.x;
    __syncthreads();
    if (threadIdx.x == 0) base = nbuf ? atomicAdd(count, nbuf) : 0;
    __syncthreads();
    for (int e = threadIdx.x; e < nbuf; e += 256) list[base + e] = buf[e];
}
// the listed cells with at most RD_LIGHT - 1 band maps; the others go to
// hlist.  The loop bound is uniform per workgroup, so every lane of a
// wavefront reaches redraw_cell's reservation scan.
__global__ void k_redraw(Geo g, Phys P, const ObjC *O, RMap R, Band B, Scratch W, int nlayers, int *status,
                         const int *list, const int *count, int *hlist, int *hcount) {
    const int n = *count;
    for (int q0 = blockIdx.x * blockDim.x; q0 < n; q0 += gridDim.x * blockDim.x) {
        const int q = q0 + threadIdx.x;
        bool live = q < n;
        const int t = live ? list[q] : 0;
        if (live && B.cnt[t] > RD_LIGHT - 1) {
            hlist[atomicAdd(hcount, 1)] = t;
            live = false;
        }
        redraw_cell<RD_LIGHT, RD_LIGHT, false>(g, P, O, R, B, W, nlayers, status, t, live);
    }
}
__global__ void k_redraw_heavy(Geo g, Phys P, const ObjC *O, RMap R, Band B, Scratch W, int nlayers, int *status,
                               const int *hlist, const int *hcount) {
    int n = *hcount;
    for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < n; q += gridDim.x * blockDim.x)
        redraw_cell<RD_MAX, 1, true>(g, P, O, R, B, W, nlayers, status, hlist[q], true);
}
void launch_redraw(hipStream_t s, const Geo &g, const Phys &P, const ObjC *O, RMap R, Band B, Scratch &W, int nlayers,
                   int *status) {
    launch_zero_i32(s, W.cnt, 1);
    hipLaunchKernelGGL(k_redraw_scan, dim3((g.Ng + 256 * RD_SCAN_CHUNKS - 1) / (256 * RD_SCAN_CHUNKS)), dim3(256), 0,
                       s, g, R, B, W.list[0], W.cnt, W.st_n, W.cnt + 1);
    hipLaunchKernelGGL(k_redraw, dim3(1024), dim3(256), 0, s, g, P, O, R, B, W, nlayers, status,
                       (const int *)W.list[0], (const int *)W.cnt, W.list[1], W.cnt + 1);
    hipLaunchKernelGGL(k_redraw_heavy, dim3(64), dim3(64), 0, s, g, P, O, R, B, W, nlayers, status,
                       (const int *)W.list[1], (const int *)(W.cnt + 1));
}

// ---- reductions ----
// fixed blocked order (256 chunks, 256 strided partial sums, binary tree,
// chunk sums added in order): deterministic and reproduced by the oracle
__global__ void k_sum_blocks(const double *a, int n, double scale, double *partial) {
    __shared__ double sh[256];
    double v = 0;
    int per = (n + gridDim.x - 1) / gridDim.x;
    int s = blockIdx.x * per, e = s + per < n ? s + per : n;
#pragma unroll 8
    for (int t = s + threadIdx.x; t < e; t += blockDim.x) v += a[t] * scale;
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if (threadIdx.x < (unsigned)off) sh[threadIdx.x] += sh[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = sh[0];
}
// the partials are loaded by all work-items at once, then added in order by
// one (nb <= 256)
__global__ void k_finish_sum(const double *partial, int nb, double *out) {
    __shared__ double sh[256];
    for (int b = threadIdx.x; b < nb; b += blockDim.x) sh[b] = partial[b];
    __syncthreads();
    if (threadIdx.x != 0) return;
    double v = 0;
#pragma unroll 16
    for (int b = 0; b < nb; b++) v += sh[b];
    *out = v;
}
void launch_sum_blocks(hipStream_t s, const double *a, int n, double scale, double *partial) {
    hipLaunchKernelGGL(k_sum_blocks, dim3(256), dim3(256), 0, s, a, n, scale, partial);
}
void launch_finish_sum(hipStream_t s, const double *partial, double *out) {
    hipLaunchKernelGGL(k_finish_sum, dim3(1), dim3(256), 0, s, partial, 256, out);
}

}  // namespace rmt
