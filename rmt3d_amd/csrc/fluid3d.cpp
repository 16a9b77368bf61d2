// Host orchestration of the MI355X FSI step (see fluid3d.h).
#include <map>
#include <mutex>
#include "fluid3d.h"
#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <algorithm>
#include <vector>
#include <cstdio>

namespace rmt {

#define HCHK(x)                                                                              \
    do {                                                                                     \
        hipError_t e_ = (x);                                                                 \
        if (e_ != hipSuccess) {                                                              \
            err = std::string(#x) + ": " + hipGetErrorString(e_);                            \
            return RMT3D_ERR_HIP;                                                            \
        }                                                                                    \
    } while (0)

static const int ORC_VEL = 1, ORC_PRES = 2, ORC_FVEL = 4, ORC_C = 8, ORC_X = 16, ORC_XPRED = 32, ORC_FX = 64,
                 ORC_DVEL = 2048;

// ---------------------------------------------------------------------------
// multigrid
// ---------------------------------------------------------------------------
// a periodic dimension is halved only while the half stays even, keeping the
// smoothers' parity colourings valid across the wrap
static bool can_coarsen(int fem, int d, int prd) {
    if (prd) return (d % 4 == 0) && d / 2 >= 2;
    if (!fem) return (d % 2 == 0) && d / 2 >= 2;
    return (d % 2 == 1) && d >= 3 && (d - 1) / 2 + 1 >= 2;
}
static int coarse_dim(int fem, int d, int prd) { return (!fem || prd) ? d / 2 : (d - 1) / 2 + 1; }

// The coarsest operator as a dense matrix (row-major), entry by entry as
// mac_off / fem_off weight them: MAC +1 per face neighbour present and
// -6 + missing on the diagonal; FEM the 20 edge and corner weights (-2, -1)
// halved per wall dimension whose offset is zero, 32 halved likewise on the
// diagonal.  A neighbour reached twice (a periodic extent of 2) adds twice.
// Every entry is a small multiple of 1/8, so the sums are exact.
static void coarse_operator(int fem, const int d[3], const int prd[3], double *A) {
    int n = d[0] * d[1] * d[2];
    for (int e = 0; e < n * n; e++) A[e] = 0;
    auto nb = [&](int i, int di, int a) {
        int j = i + di;
        if (j < 0) return prd[a] ? j + d[a] : -1;
        if (j >= d[a]) return prd[a] ? j - d[a] : -1;
        return j;
    };
    for (int k = 0; k < d[2]; k++)
        for (int j = 0; j < d[1]; j++)
            for (int i = 0; i < d[0]; i++) {
                int r = i + d[0] * (j + d[1] * k);
                double *row = A + (size_t)r * n;
                if (!fem) {
                    int g[3] = {i, j, k}, miss = 0;
                    for (int a = 0; a < 3; a++)
                        for (int sg = -1; sg <= 1; sg += 2) {
                            int q = nb(g[a], sg, a);
                            if (q < 0) { miss++; continue; }
                            int h[3] = {i, j, k};
                            h[a] = q;
                            row[h[0] + d[0] * (h[1] + d[1] * h[2])] += 1;
                        }
                    row[r] += -6.0 + miss;
                    continue;
                }
                bool w0 = !prd[0] && (i == 0 || i == d[0] - 1);
                bool w1 = !prd[1] && (j == 0 || j == d[1] - 1);
                bool w2 = !prd[2] && (k == 0 || k == d[2] - 1);
                double sx = w0 ? 0.5 : 1, sy = w1 ? 0.5 : 1, sz = w2 ? 0.5 : 1;
                for (int ok = -1; ok <= 1; ok++)
                    for (int oj = -1; oj <= 1; oj++)
                        for (int oi = -1; oi <= 1; oi++) {
                            int nz = (oi != 0) + (oj != 0) + (ok != 0);
                            if (nz < 2) continue;
                            int ii = nb(i, oi, 0), jj = nb(j, oj, 1), kk = nb(k, ok, 2);
                            if (ii < 0 || jj < 0 || kk < 0) continue;
                            double w = (nz == 2) ? -2 : -1;
                            if (oi == 0) w *= sx;
                            if (oj == 0) w *= sy;
                            if (ok == 0) w *= sz;
                            row[ii + d[0] * (jj + d[1] * kk)] += w;
                        }
                row[r] += 32 * sx * sy * sz;
            }
}

// Moore-Penrose pseudo-inverse of a symmetric matrix from its eigen-
// decomposition by cyclic Jacobi rotations (Golub & Van Loan, Matrix
// Computations, 4th ed., sec. 8.5): rotations in (p, q) row order until the
// off-diagonal mass falls below 1e-28 of the diagonal's, then
// P = sum over eigenvalues |l_i| > 1e-10 max|l| of v_i v_i^T / l_i, summed in
// eigenvalue order, computed on and above the diagonal and mirrored.  The
// operation order is fixed (no contraction, no reassociation), so the result
// is the same bits on every IEEE-754 host.
static void sym_pinv(int n, const double *A0, double *P) {
    size_t nn = (size_t)n * n;
    std::vector<double> A(A0, A0 + nn), V(nn, 0.0);
    for (int i = 0; i < n; i++) V[(size_t)i * n + i] = 1;
    for (int sweep = 0; sweep < 100; sweep++) {
        double off = 0, dia = 0;
        for (int p = 0; p < n; p++) {
            dia += A[(size_t)p * n + p] * A[(size_t)p * n + p];
            for (int q = p + 1; q < n; q++) off += A[(size_t)p * n + q] * A[(size_t)p * n + q];
        }
        if (off <= 1e-28 * dia) break;
        for (int p = 0; p < n - 1; p++)
            for (int q = p + 1; q < n; q++) {
                double apq = A[(size_t)p * n + q];
                if (apq == 0) continue;
                double theta = (A[(size_t)q * n + q] - A[(size_t)p * n + p]) / (2 * apq);
                double t = 1 / (fabs(theta) + sqrt(theta * theta + 1));
                if (theta < 0) t = -t;
                double c = 1 / sqrt(t * t + 1), s = t * c;
                for (int k = 0; k < n; k++) {
                    double akp = A[(size_t)k * n + p], akq = A[(size_t)k * n + q];
                    A[(size_t)k * n + p] = c * akp - s * akq;
                    A[(size_t)k * n + q] = s * akp + c * akq;
                }
                for (int k = 0; k < n; k++) {
                    double apk = A[(size_t)p * n + k], aqk = A[(size_t)q * n + k];
                    A[(size_t)p * n + k] = c * apk - s * aqk;
                    A[(size_t)q * n + k] = s * apk + c * aqk;
                }
                for (int k = 0; k < n; k++) {
                    double vkp = V[(size_t)k * n + p], vkq = V[(size_t)k * n + q];
                    V[(size_t)k * n + p] = c * vkp - s * vkq;
                    V[(size_t)k * n + q] = s * vkp + c * vkq;
                }
            }
    }
    double lmax = 0;
    for (int i = 0; i < n; i++) lmax = std::max(lmax, fabs(A[(size_t)i * n + i]));
    double thr = 1e-10 * lmax;
    for (int r = 0; r < n; r++)
        for (int c = r; c < n; c++) {
            double s = 0;
            for (int i = 0; i < n; i++) {
                double l = A[(size_t)i * n + i];
                if (fabs(l) > thr) s += V[(size_t)r * n + i] * V[(size_t)c * n + i] / l;
            }
            P[(size_t)r * n + c] = s;
            P[(size_t)c * n + r] = s;
        }
}

// words of fast-diagonalisation operands the hierarchy of a d[0] x d[1] x
// d[2] problem allocates: the coarsening rule of MGSolver::setup
static size_t mg_fdm_words(int fem, const int d[3], const int prd[3]) {
    int cur[3] = {d[0], d[1], d[2]};
    for (int lev = 0; lev < 16; lev++) {
        size_t nn = (size_t)cur[0] * cur[1] * cur[2];
        if (nn <= (size_t)MG_DIRECT_NN) return 0;
        if (lev > 0 && nn > (size_t)MG_TAIL_NN && nn <= (size_t)MG_FDM_NN && cur[0] <= MG_FDM_DMAX &&
            cur[1] <= MG_FDM_DMAX && cur[2] <= MG_FDM_DMAX)
            return nn + 2 * ((size_t)cur[0] * cur[0] + (size_t)cur[1] * cur[1] + (size_t)cur[2] * cur[2]);
        for (int a = 0; a < 3; a++)
            if (!can_coarsen(fem, cur[a], prd[a])) return 0;
        for (int a = 0; a < 3; a++) cur[a] = coarse_dim(fem, cur[a], prd[a]);
    }
    return 0;
}

void MGSolver::setup(int fem_, const int d[3], const int prd_[3], bool mixed_) {
    fem = fem_;
    mixed = mixed_;
    for (int a = 0; a < 3; a++) prd[a] = prd_[a];
    int cur[3] = {d[0], d[1], d[2]};
    nlev = 0;
    for (;;) {
        MGLevel<double> &l = L[nlev];
        for (int a = 0; a < 3; a++) l.d[a] = cur[a];
        l.nn = cur[0] * cur[1] * cur[2];
        l.x = l.b = l.r = nullptr;
        bool ok_alloc = hipMalloc(&l.x, sizeof(double) * l.nn) == hipSuccess &&
                        hipMalloc(&l.b, sizeof(double) * l.nn) == hipSuccess &&
                        hipMalloc(&l.r, sizeof(double) * l.nn) == hipSuccess;
        bytes += 3 * sizeof(double) * (size_t)l.nn;
        if (ok_alloc) {
            hipMemset(l.x, 0, sizeof(double) * l.nn);
            hipMemset(l.b, 0, sizeof(double) * l.nn);
            hipMemset(l.r, 0, sizeof(double) * l.nn);
        }
        if (mixed) {
            MGLevel<float> &f = Lf[nlev];
            for (int a = 0; a < 3; a++) f.d[a] = cur[a];
            f.nn = l.nn;
            f.x = f.b = f.r = nullptr;
            ok_alloc = ok_alloc && hipMalloc(&f.x, sizeof(float) * f.nn) == hipSuccess &&
                       hipMalloc(&f.b, sizeof(float) * f.nn) == hipSuccess &&
                       hipMalloc(&f.r, sizeof(float) * f.nn) == hipSuccess;
            bytes += 3 * sizeof(float) * (size_t)f.nn;
            if (ok_alloc) {
                hipMemset(f.x, 0, sizeof(float) * f.nn);
                hipMemset(f.b, 0, sizeof(float) * f.nn);
                hipMemset(f.r, 0, sizeof(float) * f.nn);
            }
        }
        nlev++;
        if (!ok_alloc) { alloc_failed = true; break; }
        if (l.nn <= MG_DIRECT_NN) break;
        if (nlev > 1 && l.nn > MG_TAIL_NN && l.nn <= MG_FDM_NN && cur[0] <= MG_FDM_DMAX && cur[1] <= MG_FDM_DMAX &&
            cur[2] <= MG_FDM_DMAX) {
            fdm = true;
            break;
        }
        bool ok = nlev < 16;
        for (int a = 0; a < 3; a++) ok = ok && can_coarsen(fem, cur[a], prd[a]);
        if (!ok) break;
        for (int a = 0; a < 3; a++) cur[a] = coarse_dim(fem, cur[a], prd[a]);
    }
    if (!alloc_failed && fdm) {
        // bases S[a] and transposes ST[a] of the three dimensions, then
        // 1 / eigenvalue per mode, packed in one array
        const MGLevel<double> &c = L[nlev - 1];
        size_t off[7], tot = 0;
        for (int a = 0; a < 3; a++) {
            off[a] = tot; tot += (size_t)c.d[a] * c.d[a];
            off[3 + a] = tot; tot += (size_t)c.d[a] * c.d[a];
        }
        off[6] = tot; tot += c.nn;
        std::vector<double> h(tot);
        std::vector<double> lam[3];
        for (int a = 0; a < 3; a++) {
            int d = c.d[a];
            lam[a].resize(d);
            double *S = h.data() + off[a], *ST = h.data() + off[3 + a];
            mg_fdm_basis(fem, d, prd[a], S, lam[a].data());
            for (int q = 0; q < d; q++)
                for (int p = 0; p < d; p++) ST[(size_t)p * d + q] = S[(size_t)q * d + p];
        }
        double scale = fem ? 12.0 : 1.0;
        double *dinv = h.data() + off[6];
        for (int k = 0; k < c.d[2]; k++)
            for (int j = 0; j < c.d[1]; j++)
                for (int i = 0; i < c.d[0]; i++) {
                    double D = scale * ((lam[0][i] + lam[1][j]) + lam[2][k]);
                    dinv[i + c.d[0] * (j + c.d[1] * k)] = (D == 0) ? 0.0 : 1.0 / D;
                }
        bool ok = hipMalloc(&d_fdm, sizeof(double) * tot) == hipSuccess &&
                  hipMemcpy(d_fdm, h.data(), sizeof(double) * tot, hipMemcpyHostToDevice) == hipSuccess;
        bytes += sizeof(double) * tot;
        if (ok) {
            for (int a = 0; a < 3; a++) { fdm64.S[a] = d_fdm + off[a]; fdm64.ST[a] = d_fdm + off[3 + a]; }
            fdm64.dinv = d_fdm + off[6];
        }
        if (ok && mixed) {
            // FP32: the same arrays, then zero-padded copies of the bases
            // (rows of 16-float multiples, each copy 64-float aligned) for
            // the direct-operand MFMA form
            std::vector<float> hf(h.begin(), h.end());
            size_t offp[6];
            for (int a = 0; a < 3; a++) {
                const size_t d = c.d[a], X = (d + 15) & ~(size_t)15;
                for (int t = 0; t < 2; t++) {
                    hf.resize((hf.size() + 63) & ~(size_t)63, 0.f);
                    offp[2 * a + t] = hf.size();
                    hf.resize(hf.size() + X * X, 0.f);
                    const double *src = h.data() + off[t ? 3 + a : a];
                    float *dst = hf.data() + offp[2 * a + t];
                    for (size_t i = 0; i < d; i++)
                        for (size_t j = 0; j < d; j++) dst[i * X + j] = (float)src[i * d + j];
                }
            }
            ok = hipMalloc(&f_fdm, sizeof(float) * hf.size()) == hipSuccess &&
                 hipMemcpy(f_fdm, hf.data(), sizeof(float) * hf.size(), hipMemcpyHostToDevice) == hipSuccess;
            bytes += sizeof(float) * hf.size();
            if (ok) {
                for (int a = 0; a < 3; a++) {
                    fdm32.S[a] = f_fdm + off[a];
                    fdm32.ST[a] = f_fdm + off[3 + a];
                    fdm32.SP[a] = f_fdm + offp[2 * a];
                    fdm32.STP[a] = f_fdm + offp[2 * a + 1];
                }
                fdm32.dinv = f_fdm + off[6];
            }
        }
        if (!ok) alloc_failed = true;
    } else if (!alloc_failed && L[nlev - 1].nn <= MG_DIRECT_NN) {
        const MGLevel<double> &c = L[nlev - 1];
        int n = c.nn;
        std::vector<double> A((size_t)n * n), Pm((size_t)n * n);
        coarse_operator(fem, c.d, prd, A.data());
        sym_pinv(n, A.data(), Pm.data());
        bool ok = hipMalloc(&d_pinv, sizeof(double) * n * n) == hipSuccess &&
                  hipMemcpy(d_pinv, Pm.data(), sizeof(double) * n * n, hipMemcpyHostToDevice) == hipSuccess;
        bytes += sizeof(double) * (size_t)n * n;
        if (ok && mixed) {
            std::vector<float> Pf(Pm.begin(), Pm.end());
            ok = hipMalloc(&f_pinv, sizeof(float) * n * n) == hipSuccess &&
                 hipMemcpy(f_pinv, Pf.data(), sizeof(float) * n * n, hipMemcpyHostToDevice) == hipSuccess;
            bytes += sizeof(float) * (size_t)n * n;
        }
        if (!ok) alloc_failed = true;
    }
    // the tail starts at the first level small enough for one workgroup,
    // with all its levels in LDS when they fit
    tail_lev = fdm ? nlev : nlev - 1;
    while (!fdm && tail_lev > 0 && L[tail_lev - 1].nn <= MG_TAIL_NN && nlev - (tail_lev - 1) <= MG_TAIL_MAXLEV &&
           mg_tail_lds_bytes(L + tail_lev - 1, nlev - tail_lev + 1) <= MG_TAIL_LDS_MAX)
        tail_lev--;
    // RMT3D_NO_GRAPH=1 launches the V-cycles one kernel at a time (for
    // tracing tools that do not follow graph launches)
    const char *ng = getenv("RMT3D_NO_GRAPH");
    no_graph = ng && ng[0] == '1';
    const char *fo = getenv("RMT3D_MIXED_FUSE");
    fuse_off = fo && fo[0] == '0';
    // colour passes per launch above the tail: one per launch on the finest
    // level (bandwidth-bound, where a halo's redundant loads cost more than
    // launches save), a whole sweep per launch below it (latency-bound);
    // RMT3D_SWEEP="n0,n1,..." sets them per level (1: one launch per colour)
    for (int l = 0; l < 16; l++) npass[l] = (l == 0) ? 1 : (fem ? 4 : 2);
    if (const char *sw = getenv("RMT3D_SWEEP")) {
        int l = 0;
        for (const char *p = sw; *p && l < 16; l++) {
            int v = atoi(p);
            if (v >= 1 && v <= (fem ? 4 : 2)) npass[l] = v;
            while (*p && *p != ',') p++;
            if (*p == ',') p++;
        }
    }
    mg_sweep_enable();
    if (fdm) {
        // the FP32 fast-diagonalisation solve (mixed path) runs in three
        // MFMA launches when the level's planes, pencil blocks and bases fit
        // a workgroup's LDS; RMT3D_FDM_FUSED=0 keeps one launch per
        // contraction.  The FP64 solve always takes the six launches.
        const char *ff = getenv("RMT3D_FDM_FUSED");
        bool allow = !(ff && ff[0] == '0');
        fdm64.fused = 0;
        fdm32.forms = (allow && mixed) ? mg_fdm_fused_enable(L[nlev - 1].d) : 0;
        fdm32.fused = (fdm32.forms & 2) ? 2 : (fdm32.forms & 1);
        fdm_fused_ok = fdm32.forms ? 2 : 0;
    } else {
        tail_lds64 = mg_tail_enable_lds<double>(mg_tail_lds_bytes(L + tail_lev, nlev - tail_lev));
        tail_lds32 = mixed ? mg_tail_enable_lds<float>(mg_tail_lds_bytes(Lf + tail_lev, nlev - tail_lev)) : 0;
    }
    double eps = DBL_EPSILON, c = fem ? 32.0 : 6.0;
    tol = c * c * eps * eps * 1e4;
    max_norm_tol = c * eps * 1e2;
    log_tol = log10(tol);
    hipMalloc(&d_max, sizeof(unsigned long long));
    hipHostMalloc((void **)&h_max, sizeof(unsigned long long), hipHostMallocDefault);
    hipEventCreateWithFlags(&ev_check, hipEventDisableTiming);
}

void MGSolver::release() {
    if (gexec) hipGraphExecDestroy(gexec);
    gexec = nullptr;
    for (int l = 0; l < nlev; l++) {
        if (L[l].x) hipFree(L[l].x);
        if (L[l].b) hipFree(L[l].b);
        if (L[l].r) hipFree(L[l].r);
        if (mixed) {
            if (Lf[l].x) hipFree(Lf[l].x);
            if (Lf[l].b) hipFree(Lf[l].b);
            if (Lf[l].r) hipFree(Lf[l].r);
        }
    }
    nlev = 0;
    if (d_pinv) hipFree(d_pinv);
    if (f_pinv) hipFree(f_pinv);
    d_pinv = nullptr; f_pinv = nullptr;
    if (d_fdm) hipFree(d_fdm);
    if (f_fdm) hipFree(f_fdm);
    d_fdm = nullptr; f_fdm = nullptr;
    fdm = false;
    if (d_max) hipFree(d_max);
    if (h_max) hipHostFree(h_max);
    if (ev_check) hipEventDestroy(ev_check);
    d_max = nullptr; h_max = nullptr; ev_check = nullptr;
}

void MGSolver::check_queue(hipStream_t s) {
    launch_zero_u64(s, d_max, 1);
    launch_mg_resid64_max(s, fem, L[0], prd, d_max);
    hipMemcpyAsync(h_max, d_max, sizeof(unsigned long long), hipMemcpyDeviceToHost, s);
    hipEventRecord(ev_check, s);
}

double MGSolver::check_wait() {
    hipEventSynchronize(ev_check);
    double v;
    memcpy(&v, h_max, sizeof v);
    return v;
}

double MGSolver::resid_max(hipStream_t s) {
    launch_zero_u64(s, d_max, 1);
    launch_mg_resid64_max(s, fem, L[0], prd, d_max);
    hipMemcpyAsync(h_max, d_max, sizeof(unsigned long long), hipMemcpyDeviceToHost, s);
    hipStreamSynchronize(s);
    double v;
    memcpy(&v, h_max, sizeof v);
    return v;
}

// nsweeps smoothing sweeps: each sweep is the level's colour passes in
// order, np of them per launch (np = 1: one in-place single-colour launch per
// pass).  A multi-pass launch reads the unknowns from one array and writes
// them to another, so the launches alternate between x and the level's
// residual array (free while smoothing); an odd count ends with a copy back
// into x.  Per-kernel profiling stamps each launch's own start and end.
// first_col / x64 (single-colour launches only, the mixed path's finest
// level): the first sweep starts at colour first_col (the colours before it
// were applied by the residual kernel), and every pass of the last sweep adds
// its nodes' new values to the FP64 solution x64.
template <typename T>
static void smooth(hipStream_t s, int fem, const MGLevel<T> &l, const int prd[3], int np, int nsweeps, KProf *kp,
                   int bs = 0, int first_col = 0, double *x64 = nullptr) {
    int ncol = fem ? 4 : 2;
    // a halo no wider than a periodic extent wraps at most once
    for (int a = 0; a < 3; a++)
        if (prd[a] && l.d[a] < np) np = l.d[a];
    if (np < 1) np = 1;
    int m = 0;
    for (int sw = 0; sw < nsweeps; sw++)
        for (int c0 = (sw == 0 && np == 1) ? first_col : 0; c0 < ncol; c0 += np) {
            hipEvent_t e0 = nullptr, e1 = nullptr;
            if (kp) kp->pair(e0, e1);
            if (np == 1) {
                launch_mg_smooth(s, fem, l, prd, c0, e0, e1, bs, sw == nsweeps - 1 ? x64 : nullptr);
                continue;
            }
            SweepPlan P = mg_sweep_plan(fem, l.d, prd, (int)sizeof(T), std::min(np, ncol - c0));
            P.c0 = c0;
            const T *src = (m & 1) ? l.r : l.x;
            T *dst = (m & 1) ? l.x : l.r;
            launch_mg_sweep(s, fem, l, prd, P, src, dst, e0, e1);
            m++;
        }
    if (m & 1) launch_mg_copy(s, (const T *)l.r, l.x, l.nn);
}

// ---------------------------------------------------------------------------
// Autotuner of the multigrid launch shapes.  The reference picks its
// V-cycle schedule at run time by measurement (mg_opt searches (v_up,
// v_down), mg_probs.hh:80-208, off by default).  Here the smoothing counts
// stay the reference's 2 + 2 and what is searched is how the same arithmetic
// is laid out on the GPU, level by level, in the precision the V-cycles run
// in:
//   - per smoothing level above the tail: colour passes per launch (1, 2 or
//     4 FEM; 1 or 2 MAC) and, for single-pass launches, the workgroup size
//     (64 ... 1024) of the colour-pass kernel;
//   - per dimension of a fast-diagonalisation level: outputs per work-item
//     (2, 4, 8) of its contractions, and then that six-launch form against
//     the three-launch one.
// Every candidate performs the same operations in the same order on every
// unknown (the fused sweeps are bit-identical to single-colour passes), so a
// choice changes time only, never results.  Each candidate is timed over a
// few whole iterations of the solve on the solver's own arrays (zero at
// set-up, zeroed again afterwards), replayed from a graph as the solve runs
// them, with events on the context stream; the minimum over three
// interleaved rounds decides.  Choices are cached per operator, level shape and
// precision for the life of the process, so contexts of one shape (a serving
// mix) tune once.  RMT3D_TUNE=0 keeps the built-in defaults; RMT3D_SWEEP, when
// set, fixes the passes per launch.
// ---------------------------------------------------------------------------
namespace {
struct TuneChoice {
    int npass[16];
    int bs[16];
    int pb[3];
    int fused;
    int vd, vu;
};
std::mutex g_tune_mu;
std::mutex g_tune_run_mu;
std::map<std::vector<int>, TuneChoice> g_tune_cache;

// minimum over rounds of the time of `reps` runs of op(c), per candidate c
template <typename Op>
std::vector<double> time_candidates(hipStream_t s, int ncand, Op op) {
    const int reps = 3, rounds = 3;
    std::vector<double> best(ncand, 1e30);
    hipEvent_t e0, e1;
    if (hipEventCreate(&e0) != hipSuccess) return best;
    if (hipEventCreate(&e1) != hipSuccess) { hipEventDestroy(e0); return best; }
    for (int c = 0; c < ncand; c++) op(c);   // warm-up (code objects, caches)
    for (int r = 0; r < rounds; r++)
        for (int c = 0; c < ncand; c++) {
            hipEventRecord(e0, s);
            for (int k = 0; k < reps; k++) op(c);
            hipEventRecord(e1, s);
            float ms = 0;
            if (hipEventSynchronize(e1) == hipSuccess && hipEventElapsedTime(&ms, e0, e1) == hipSuccess)
                best[c] = std::min(best[c], (double)ms);
        }
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    return best;
}

int argmin(const std::vector<double> &v) { return (int)(std::min_element(v.begin(), v.end()) - v.begin()); }
}  // namespace

// Times one iteration of the solve per candidate as the solve will run it:
// each candidate's iteration (after configure(c)) is captured into a graph
// of its own and the graphs are replayed, or, where the solve runs eagerly
// (RMT3D_NO_GRAPH, per-kernel profiling, a failed capture), the launches
// are issued eagerly.  Eager timing adds each launch's host cost, which
// favours the candidates with fewer launches.  Runs inside Sim::create,
// which holds hip_setup_mutex, so no other context captures meanwhile.
template <typename Cfg>
std::vector<double> MGSolver::time_iterations(hipStream_t s, int ncand, Cfg configure) {
    bool eager = graph_failed || no_graph || (prof && prof->enabled);
    std::vector<hipGraphExec_t> ex(ncand, nullptr);
    for (int c = 0; c < ncand && !eager; c++) {
        configure(c);
        hipGraph_t gr = nullptr;
        if (hipStreamBeginCapture(s, hipStreamCaptureModeRelaxed) != hipSuccess) {
            (void)hipGetLastError();
            eager = true;
            break;
        }
        cycle_body(s);
        hipError_t e1 = hipStreamEndCapture(s, &gr);
        hipError_t e2 = (e1 == hipSuccess) ? hipGraphInstantiate(&ex[c], gr, nullptr, nullptr, 0) : e1;
        if (gr) hipGraphDestroy(gr);
        if (e2 != hipSuccess) {
            hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
            if (hipStreamIsCapturing(s, &cs) == hipSuccess && cs != hipStreamCaptureStatusNone) {
                hipGraph_t g2 = nullptr;
                (void)hipStreamEndCapture(s, &g2);
                if (g2) hipGraphDestroy(g2);
            }
            (void)hipGetLastError();
            ex[c] = nullptr;
            eager = true;
        }
    }
    std::vector<double> t;
    if (eager)
        t = time_candidates(s, ncand, [&](int c) {
            configure(c);
            cycle_body(s);
        });
    else
        t = time_candidates(s, ncand, [&](int c) { hipGraphLaunch(ex[c], s); });
    hipStreamSynchronize(s);
    for (hipGraphExec_t e : ex)
        if (e) hipGraphExecDestroy(e);
    return t;
}

template <typename T>
void MGSolver::tune_levels(hipStream_t s, MGLevel<T> *LV, FdmDev<T> &F, bool sweep_fixed) {
    const int ncol = fem ? 4 : 2;
    int last = fdm ? nlev - 1 : tail_lev;   // levels [0, last) are smoothed by launches
    // Each candidate is timed inside whole iterations of the solve (the
    // V-cycle, or the defect-correction step around it), so that it meets
    // the caches the other kernels of the cycle leave it: a colour pass timed
    // alone on data that stays in L2 prefers other workgroup sizes than it
    // does inside the cycle.  Choices are made from the coarsest level up,
    // every other level keeping its current (default or chosen) shape.
    if (fdm) {
        // the six-launch form's outputs per work-item, then that form
        // against the three-launch ones that fit
        const int fused_ok = F.fused;
        F.fused = 0;
        for (int a = 0; a < 3; a++) {
            static const int pbs[3] = {2, 4, 8};
            int keep = F.pb[a];
            std::vector<double> t = time_iterations(s, 3, [&](int k) { F.pb[a] = pbs[k]; });
            F.pb[a] = t[argmin(t)] < 1e29 ? pbs[argmin(t)] : keep;
        }
        if (fused_ok) {
            std::vector<int> forms = {0};
            for (int f = 1; f <= 2; f++)
                if (F.forms & f) forms.push_back(f);
            std::vector<double> t = time_iterations(s, (int)forms.size(), [&](int k) { F.fused = forms[k]; });
            F.fused = t[argmin(t)] < 1e29 ? forms[argmin(t)] : fused_ok;
        }
    }
    for (int l = last - 1; l >= 0; l--) {
        // candidates: (np = 1, bs) for each workgroup size, then np = 2 (, 4)
        std::vector<std::pair<int, int>> cand;
        static const int sizes[5] = {64, 128, 256, 512, 1024};
        for (int bsz : sizes) cand.push_back({1, bsz});
        if (!sweep_fixed)
            for (int np = 2; np <= ncol; np *= 2) cand.push_back({np, 0});
        else if (npass[l] > 1)
            cand.assign(1, {npass[l], 0});
        int keep_np = npass[l], keep_bs = smooth_bs[l];
        std::vector<double> t = time_iterations(s, (int)cand.size(), [&](int c) {
            npass[l] = cand[c].first;
            smooth_bs[l] = cand[c].second;
        });
        int w = argmin(t);
        if (t[w] < 1e29) {
            npass[l] = cand[w].first;
            smooth_bs[l] = cand[w].second;
        } else {
            npass[l] = keep_np;
            smooth_bs[l] = keep_bs;
        }
    }
    // the iterations ran on zero arrays; leave every level as set-up left it
    for (int l = 0; l < nlev; l++) {
        hipMemsetAsync(L[l].x, 0, sizeof(double) * L[l].nn, s);
        hipMemsetAsync(L[l].b, 0, sizeof(double) * L[l].nn, s);
        hipMemsetAsync(L[l].r, 0, sizeof(double) * L[l].nn, s);
        if (mixed) {
            hipMemsetAsync(Lf[l].x, 0, sizeof(float) * Lf[l].nn, s);
            hipMemsetAsync(Lf[l].b, 0, sizeof(float) * Lf[l].nn, s);
            hipMemsetAsync(Lf[l].r, 0, sizeof(float) * Lf[l].nn, s);
        }
    }
}

// the schedule RMT3D_MIXED_VCYCLE fixes, as 10 down + up, or 0
static int mixed_vcycle_env() {
    const char *vc = getenv("RMT3D_MIXED_VCYCLE");
    int d = 0, u = 0;
    if (vc && sscanf(vc, "%d,%d", &d, &u) == 2 && d >= 0 && u >= 1 && d <= 4 && u <= 4) return 10 * d + u;
    return 0;
}

void MGSolver::tune(hipStream_t s) {
    const char *tv = getenv("RMT3D_TUNE");
    if (tv && tv[0] == '0') {
        if (mixed)
            if (int f = mixed_vcycle_env()) {
                v_down = f / 10;
                v_up = f % 10;
            }
        return;
    }
    bool sweep_fixed = getenv("RMT3D_SWEEP") != nullptr;
    std::vector<int> key = {fem, mixed ? 4 : 8, sweep_fixed ? 1 : 0, fuse_off ? 1 : 0, fdm_fused_ok, fdm32.forms,
                            mixed ? mixed_vcycle_env() : 0, nlev, fdm ? 1 : 0,
                            prd[0], prd[1], prd[2]};
    for (int l = 0; l < nlev; l++)
        for (int a = 0; a < 3; a++) key.push_back(L[l].d[a]);
    // one tuning at a time in the process: contexts of one shape created
    // together (a serving mix) wait for the first one's choice and share it
    std::lock_guard<std::mutex> run_lk(g_tune_run_mu);
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        auto it = g_tune_cache.find(key);
        if (it != g_tune_cache.end()) {
            for (int l = 0; l < 16; l++) { npass[l] = it->second.npass[l]; smooth_bs[l] = it->second.bs[l]; }
            for (int a = 0; a < 3; a++) fdm64.pb[a] = fdm32.pb[a] = it->second.pb[a];
            fdm64.fused = (fdm_fused_ok & 1) ? it->second.fused : 0;
            fdm32.fused = (fdm_fused_ok & 2) ? it->second.fused : 0;
            v_down = it->second.vd;
            v_up = it->second.vu;
            tuned = 2;
            return;
        }
    }
    if (mixed) {
        tune_levels<float>(s, Lf, fdm32, sweep_fixed);
        for (int a = 0; a < 3; a++) fdm64.pb[a] = fdm32.pb[a];
        fdm64.fused = (fdm_fused_ok & 1) ? fdm32.fused : 0;
        tune_schedule(s);
    } else {
        tune_levels<double>(s, L, fdm64, sweep_fixed);
        for (int a = 0; a < 3; a++) fdm32.pb[a] = fdm64.pb[a];
        fdm32.fused = (fdm_fused_ok & 2) ? fdm64.fused : 0;
    }
    hipStreamSynchronize(s);
    (void)hipGetLastError();
    TuneChoice ch;
    for (int l = 0; l < 16; l++) { ch.npass[l] = npass[l]; ch.bs[l] = smooth_bs[l]; }
    for (int a = 0; a < 3; a++) ch.pb[a] = fdm64.pb[a];
    ch.fused = mixed ? fdm32.fused : fdm64.fused;
    ch.vd = v_down;
    ch.vu = v_up;
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tune_cache[key] = ch;
    tuned = 1;
}

// Smoothing sweeps of the mixed path's FP32 V-cycle.  The reference picks
// (v_up, v_down) by measurement when its mg_opt search is on (mg_probs.hh:
// 80-208: digits of residual reduction per second over neighbouring
// schedules).  The FP64 path keeps the reference's default 2 + 2, so that
// its iterates are the oracle's.  The mixed path's FP32 V-cycle is a
// preconditioner inside the FP64 defect correction, whose stopping rule and
// converged solution do not depend on it, so there the schedule is picked
// here: for 2 + 2, 2 + 1, 1 + 2 and (FEM) 0 + 3 sweeps (with an exact or nearly
// exact coarse solve the sweeps after it do the work of the ones before,
// and without sweeps before it the finest level's own residual pass is not
// needed), the iterations that take a fixed
// zero-mean right-hand side (a 64-bit LCG) from zero below the solver's
// FP64 tolerance, plus the two insurance iterations every solve runs after
// that (mg_probs.cc:43-47), times the event-timed cost of one iteration;
// the least product wins, 2 + 2 on ties.  RMT3D_MIXED_VCYCLE="down,up"
// fixes it.
void MGSolver::tune_schedule(hipStream_t s) {
    if (int f = mixed_vcycle_env()) {
        v_down = f / 10;
        v_up = f % 10;
        return;
    }
    // 0 + 3 only for the FEM operator: for the red-black MAC smoother it
    // scored close to 2 + 2 on this test but needed half as many iterations
    // again on the warm-started solves of a step (1.18 -> 1.43 ms per
    // 128^3 step)
    static const int cand[4][2] = {{2, 2}, {2, 1}, {1, 2}, {0, 3}};
    const int ncand = fem ? 4 : 3;
    const size_t n = L[0].nn;
    std::vector<double> h(n);
    unsigned long long st = 0x9E3779B97F4A7C15ull;
    double mean = 0;
    for (size_t i = 0; i < n; i++) {
        st = st * 6364136223846793005ull + 1442695040888963407ull;
        h[i] = (double)(st >> 11) * (1.0 / 9007199254740992.0) - 0.5;
        mean += h[i];
    }
    mean /= (double)n;
    for (size_t i = 0; i < n; i++) h[i] -= mean;
    if (hipMemcpy(L[0].b, h.data(), sizeof(double) * n, hipMemcpyHostToDevice) != hipSuccess) {
        (void)hipGetLastError();
        return;
    }
    const double thr = max_norm_tol;
    double best = 1e300;
    int bd = v_down, bu = v_up;
    for (int ci = 0; ci < ncand; ci++) {
        const int *c = cand[ci];
        v_down = c[0];
        v_up = c[1];
        hipMemsetAsync(L[0].x, 0, sizeof(double) * n, s);
        int it = 0;
        double res = resid_max(s);
        while (res > thr && it < 40) {
            cycle_body(s);
            it++;
            res = resid_max(s);
        }
        if (!(res <= thr)) continue;
        std::vector<double> t = time_iterations(s, 1, [&](int) {});
        double cost = (it + 2) * t[0];
        if (t[0] < 1e29 && cost < best * (1 - 1e-3)) {
            best = cost;
            bd = c[0];
            bu = c[1];
        }
    }
    v_down = bd;
    v_up = bu;
    for (int l = 0; l < nlev; l++) {
        hipMemsetAsync(L[l].x, 0, sizeof(double) * L[l].nn, s);
        hipMemsetAsync(L[l].b, 0, sizeof(double) * L[l].nn, s);
        hipMemsetAsync(L[l].r, 0, sizeof(double) * L[l].nn, s);
        hipMemsetAsync(Lf[l].x, 0, sizeof(float) * Lf[l].nn, s);
        hipMemsetAsync(Lf[l].b, 0, sizeof(float) * Lf[l].nn, s);
        hipMemsetAsync(Lf[l].r, 0, sizeof(float) * Lf[l].nn, s);
    }
    if (gexec) {
        hipGraphExecDestroy(gexec);
        gexec = nullptr;
    }
}

void MGSolver::vcycle64(hipStream_t s, int lev) {
    if (fdm && lev == nlev - 1) {
        launch_mg_fdm(s, L[lev], fdm64);
        return;
    }
    if (lev >= tail_lev) {
        launch_mg_tail(s, fem, L + lev, nlev - lev, prd, v_down, v_up, coarse_sweeps, tail_lds64, (const double *)d_pinv);
        return;
    }
    KProf *kp = (lev == 0) ? prof : nullptr;
    smooth(s, fem, L[lev], prd, npass[lev], v_down, kp, smooth_bs[lev]);
    launch_mg_resid_restrict(s, fem, L[lev], L[lev + 1], prd);
    vcycle64(s, lev + 1);
    launch_mg_prolong(s, fem, L[lev], L[lev + 1], prd);
    smooth(s, fem, L[lev], prd, npass[lev], v_up, kp, smooth_bs[lev]);
}

// first / last (finest level only): the first colour pass came with the
// residual, and the last sweep adds the correction into the FP64 solution.
// With no sweep before the coarse correction (v_down 0) the level's x is
// still zero (the residual kernel or the restriction above cleared it), so
// its residual is its right-hand side: that is restricted directly.
void MGSolver::vcycle32(hipStream_t s, int lev, bool first, bool last) {
    if (fdm && lev == nlev - 1) {
        launch_mg_fdm(s, Lf[lev], fdm32);
        return;
    }
    if (lev >= tail_lev) {
        launch_mg_tail(s, fem, Lf + lev, nlev - lev, prd, v_down, v_up, coarse_sweeps, tail_lds32, (const float *)f_pinv);
        return;
    }
    KProf *kp = (lev == 0) ? prof : nullptr;
    if (v_down > 0) {
        smooth(s, fem, Lf[lev], prd, npass[lev], v_down, kp, smooth_bs[lev], first ? 1 : 0);
        launch_mg_resid_restrict(s, fem, Lf[lev], Lf[lev + 1], prd);
    } else {
        MGLevel<float> rb = Lf[lev];
        rb.r = rb.b;
        launch_mg_restrict(s, fem, rb, Lf[lev + 1], prd);
    }
    vcycle32(s, lev + 1, false, false);
    launch_mg_prolong(s, fem, Lf[lev], Lf[lev + 1], prd, v_down == 0 ? 1 : 0);
    smooth(s, fem, Lf[lev], prd, npass[lev], v_up, kp, smooth_bs[lev], 0, last ? L[0].x : nullptr);
}

// The defect-correction step's FP64 pieces ride on the FP32 V-cycle's own
// launches when the finest level is smoothed one colour per launch (and
// has smoothing sweeps before and after): the first colour pass is written
// by the residual kernel, and the last sweep adds the correction into the
// FP64 solution, so neither a separate colour pass nor an update pass runs.
// The values are the same as those of the separate launches, bit for bit.
// RMT3D_MIXED_FUSE=0 runs the separate launches.  fuse_first: the residual
// kernel writes the first colour pass (there is a sweep before the coarse
// correction); fuse_last: the last sweep adds into the FP64 solution.
bool MGSolver::fuse_first() const {
    return mixed && !fuse_off && tail_lev > 0 && !(fdm && nlev == 1) && npass[0] == 1 && v_down > 0;
}
bool MGSolver::fuse_last() const {
    return mixed && !fuse_off && tail_lev > 0 && !(fdm && nlev == 1) && npass[0] == 1 && v_up > 0;
}

// one iteration: an FP64 V-cycle, or (mixed) one FP64 defect-correction step
// preconditioned by an FP32 V-cycle: r = b - A x (FP64), A e ~= r (FP32), x += e
void MGSolver::cycle_body(hipStream_t s) {
    if (!mixed) {
        vcycle64(s, 0);
        return;
    }
    const bool first = fuse_first(), last = fuse_last();
    // with no sweep before the coarse correction the prolongation stores
    // every node of the FP32 correction, so it need not be cleared here
    const bool x_stored = v_down == 0 && tail_lev > 0 && !(fdm && nlev == 1);
    launch_resid_to_f32(s, fem, L[0], prd, Lf[0].b, x_stored ? nullptr : Lf[0].x, first ? 1 : 0);
    vcycle32(s, 0, first, last);
    if (!last) launch_axpy_f2d(s, L[0].x, Lf[0].x, L[0].nn);
}

// The kernel sequence of one iteration and all its arguments are fixed for
// the solver's life, so it is captured once into a HIP graph and replayed:
// the device runs the sixty-odd dependent launches without a host round trip
// between them.  Per-kernel event timing (profiling) runs the same sequence
// eagerly.  Capture is thread-local, so other contexts' threads may allocate
// meanwhile.
void MGSolver::cycle(hipStream_t s) {
    bool eager = graph_failed || no_graph || (prof && prof->enabled);
    if (!eager && !gexec) {
        // no other context allocates, frees or uses the null stream while
        // this stream is being captured
        std::lock_guard<std::mutex> lk(hip_setup_mutex());
        hipGraph_t gr = nullptr;
        if (hipStreamBeginCapture(s, hipStreamCaptureModeRelaxed) != hipSuccess) {
            (void)hipGetLastError();
            graph_failed = true;
        } else {
            cycle_body(s);
            hipError_t e1 = hipStreamEndCapture(s, &gr);
            hipError_t e2 = (e1 == hipSuccess) ? hipGraphInstantiate(&gexec, gr, nullptr, nullptr, 0) : e1;
            if (gr) hipGraphDestroy(gr);
            if (e2 != hipSuccess) {
                // leave no half-open capture behind, then run eagerly
                hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
                if (hipStreamIsCapturing(s, &cs) == hipSuccess && cs != hipStreamCaptureStatusNone) {
                    hipGraph_t g2 = nullptr;
                    (void)hipStreamEndCapture(s, &g2);
                    if (g2) hipGraphDestroy(g2);
                }
                (void)hipGetLastError();
                gexec = nullptr;
                graph_failed = true;
            }
        }
        if (graph_failed) {
            cycle_body(s);
            return;
        }
    }
    if (eager) {
        cycle_body(s);
        return;
    }
    hipGraphLaunch(gexec, s);
}

void MGSolver::timed_cycle(hipStream_t s) {
    if (vprof) vprof->begin(s);
    cycle(s);
    if (vprof) vprof->end(s);
}

// mg_prob::run (Utilities/mg_probs.cc:3-83), profile = false
// The residual before the first cycle is read only in the reference's
// profiling mode (mg_probs.cc:9-13), so it is not taken.  Every later
// residual check is followed by at least one more V-cycle: the first of the
// next extra batch, or the first insurance cycle.  That cycle is queued
// right behind the check and its read-back, and the host waits on the
// read-back's event alone, so the GPU runs the cycle while the host decides.
// The V-cycles and checks run in the reference's order on the same iterates.
double MGSolver::run(hipStream_t s) {
    const int hard_cap = 20;
    double thr = tol_override > 0 ? tol_override : max_norm_tol;
    int cycles = 0;
    int pre = iV / iP;
    for (int i = 0; i < pre; i++) { timed_cycle(s); cycles++; }
    check_queue(s);
    timed_cycle(s); cycles++;             // queued: needed whatever the check says
    double res = check_wait();
    int extra = 0;
    while (res > thr && extra < hard_cap) {
        extra++;
        // the batch's first cycle is already queued
        for (int n = 1; n < extra; n++) { timed_cycle(s); cycles++; }
        check_queue(s);
        timed_cycle(s); cycles++;
        res = check_wait();
    }
    // the first insurance cycle is the one queued behind the last check
    timed_cycle(s); cycles++;
    // mg_pred::update(int) (mg_probs.hh:51-66)
    double dV = iV - V_avg;
    int exreduce = (int)(dV / sqrt(V_var));
    exreduce = exreduce > 0 ? exreduce : 0;
    iV += iP * extra * (extra + 1) / 2 - 1 - exreduce;
    if (iV < iP) iV = 2 * iP - 1;
    V_avg = (1 - alpha) * V_avg + alpha * iV;
    V_var = (1 - alpha) * (V_var + alpha * dV * dV);
    last_cycles = cycles;
    last_res = res;
    return res;
}

// Device bytes a context of this configuration allocates, from the same sizes
// as Sim::create (used for admission before anything is allocated).
double Sim::estimate_bytes(const rmt3d_config &c, const rmt3d_object *o, int prec) {
    int n_obj = (c.sim_type == RMT3D_SIM_FTEST) ? 0 : c.n_obj;
    double m = c.sys_size[0], n = c.sys_size[1], k = c.sys_size[2];
    double Ng = (m + 4) * (n + 4) * (k + 4);
    double per_cell = 8.0 * (36 + 39 + 3) + 4.0 * (1 + 2 + 1 + 1 + 4) + 2;
    double dh = std::min((c.xyz_bounds[1] - c.xyz_bounds[0]) / m,
                         std::min((c.xyz_bounds[3] - c.xyz_bounds[2]) / n, (c.xyz_bounds[5] - c.xyz_bounds[4]) / k));
    double cap = 65536;
    for (int i = 0; i < n_obj; i++) {
        double R0 = o[i].basic[3] + (c.nlayers + 4) * dh;
        cap += 2.0 * (4.0 / 3.0) * M_PI * R0 * R0 * R0 / (dh * dh * dh);
    }
    cap = std::min(cap, Ng * std::max(1, n_obj));
    double per_entry = 2 * (8.0 + 8.0 * 39) + 60 + 48;
    double fm = c.prd[0] ? m : m + 1, fn = c.prd[1] ? n : n + 1, fo = c.prd[2] ? k : k + 1;
    double mg_words = (m * n * k + fm * fn * fo) * 8.0 / 7.0;
    double mg = 3 * 8 * mg_words + (prec == RMT3D_PREC_MIXED ? 3 * 4 * mg_words : 0);
    double diag = 8.0 * (7 * 256 * std::max(1, n_obj) + 4096);
    // fast-diagonalisation operands (bases, transposes, 1/eigenvalues) of
    // the two hierarchies, where they end at such a level
    {
        int dm[3] = {(int)m, (int)n, (int)k}, df[3] = {(int)fm, (int)fn, (int)fo};
        diag += (double)mg_fdm_words(0, dm, c.prd) * (prec == RMT3D_PREC_MIXED ? 12 : 8);
        diag += (double)mg_fdm_words(1, df, c.prd) * (prec == RMT3D_PREC_MIXED ? 12 : 8);
    }
    return Ng * per_cell + cap * per_entry + mg + 16.0 * fm * fn * fo + diag + (1 << 20);
}

// ---------------------------------------------------------------------------
// construction (fluid_3d.cc:12-77, sim_manager.cc:19-81, sim_manager.hh:363-444)
// ---------------------------------------------------------------------------
void *Sim::dalloc(size_t bytes) {
    void *p = nullptr;
    if (bytes == 0) bytes = 8;
    if (hipMalloc(&p, bytes) != hipSuccess) {
        (void)hipGetLastError();
        dalloc_failed = true;
        return nullptr;
    }
    hipMemset(p, 0, bytes);
    allocs.push_back(p);
    dev_bytes += bytes;
    return p;
}

void Sim::alloc_band(Band &B, int cap) {
    int Ng = g.Ng;
    B.cap = cap;
    B.start = (int *)dalloc(sizeof(int) * Ng);
    B.cnt = (int *)dalloc(sizeof(int) * Ng);
    B.cell = (int *)dalloc(sizeof(int) * cap);
    B.c = (uint32_t *)dalloc(sizeof(uint32_t) * cap);
    B.phi = (double *)dalloc(sizeof(double) * cap);
    for (int a = 0; a < 3; a++) {
        B.x[a] = (double *)dalloc(sizeof(double) * cap);
        B.xpred[a] = (double *)dalloc(sizeof(double) * cap);
        B.detF[a] = (double *)dalloc(sizeof(double) * cap);
        B.en[a] = (double *)dalloc(sizeof(double) * cap);
    }
    for (int a = 0; a < 18; a++) B.fx[a] = (double *)dalloc(sizeof(double) * cap);
    for (int a = 0; a < 9; a++) B.gphi[a] = (double *)dalloc(sizeof(double) * cap);
    B.n = (int *)dalloc(sizeof(int));
}

// level 0 off; 1 stage buckets and whole V-cycles (graph replay unaffected);
// 2 also per-kernel classes, which runs the V-cycles eagerly
int Sim::profile(int level) {
    if (level < 0 || level > 2) return RMT3D_ERR_ARG;
    for (int k = 0; k < NPROF; k++) {
        KProf &P = kprof[k];
        if (!P.ev) {
            P.ev = new hipEvent_t[2 * KProf::CAP];
            for (int a = 0; a < 2 * KProf::CAP; a++) hipEventCreate(&P.ev[a]);
        }
        P.n = 0;
        bool kernel_class = k < 4;
        P.enabled = kernel_class ? (level >= 2) : (level >= 1);
    }
    fem.prof = &kprof[0];
    mac.prof = &kprof[1];
    fem.vprof = &kprof[4];
    mac.vprof = &kprof[5];
    return 0;
}

int Sim::stage_times(double *ms, int n) {
    for (int k = 0; k < n && k < NPROF - NPROF_KERNEL; k++) {
        int cnt;
        int rc = profile_read(NPROF_KERNEL + k, &ms[k], &cnt);
        if (rc) return rc;
    }
    return 0;
}

int Sim::profile_read(int which, double *ms, int *launches) {
    if (which < 0 || which >= NPROF) return RMT3D_ERR_ARG;
    HCHK(hipStreamSynchronize(stream));
    KProf &P = kprof[which];
    double tot = 0;
    for (int a = 0; a < P.n; a++) {
        float t = 0;
        hipEventElapsedTime(&t, P.ev[2 * a], P.ev[2 * a + 1]);
        tot += t;
    }
    *ms = tot;
    *launches = P.n;
    return 0;
}

Sim::~Sim() {
    std::lock_guard<std::mutex> lk(hip_setup_mutex());
    for (int k = 0; k < NPROF; k++)
        if (kprof[k].ev) {
            for (int a = 0; a < 2 * KProf::CAP; a++) hipEventDestroy(kprof[k].ev[a]);
            delete[] kprof[k].ev;
        }
    mac.release();
    fem.release();
    for (void *p : allocs) hipFree(p);
    if (h_status) hipHostFree(h_status);
    if (h_int) hipHostFree(h_int);
    if (h_dbl) hipHostFree(h_dbl);
    if (W.h_box) hipHostFree(W.h_box);
    if (W.h_batch) hipHostFree(W.h_batch);
    if (stream) hipStreamDestroy(stream);
}

// Serialises what must not overlap another context's stream capture:
// allocation, freeing and null-stream set-up copies.
std::mutex &hip_setup_mutex() {
    static std::mutex m;
    return m;
}

int Sim::create(const rmt3d_config *c, const rmt3d_object *o, int prec) {
    std::lock_guard<std::mutex> lk(hip_setup_mutex());
    cfg = *c;
    precision = prec;
    int st = cfg.sim_type;
    n_obj = (st == RMT3D_SIM_FTEST) ? 0 : cfg.n_obj;
    if (n_obj > 64) { err = "more than 64 objects"; return RMT3D_ERR_ARG; }
    if (st == RMT3D_SIM_STEST) { err = "sim_type stest is not supported"; return RMT3D_ERR_ARG; }
    objs_in.assign(o, o + n_obj);
    g.m = cfg.sys_size[0]; g.n = cfg.sys_size[1]; g.o = cfg.sys_size[2];
    if (g.m < 4 || g.n < 4 || g.o < 4) { err = "grid too small"; return RMT3D_ERR_ARG; }
    for (int d = 0; d < 3; d++)
        if (cfg.prd[d] && (cfg.sys_size[d] % 2)) {
            err = "periodic dimensions must have an even number of cells";
            return RMT3D_ERR_ARG;
        }
    g.m4 = g.m + 4; g.n4 = g.n + 4; g.o4 = g.o + 4;
    g.mn4 = g.m4 * g.n4;
    g.Ng = g.mn4 * g.o4;
    g.G0 = 2 * (1 + g.m4 * (1 + g.n4));
    for (int d = 0; d < 3; d++) g.prd[d] = cfg.prd[d];
    g.fm = g.prd[0] ? g.m : g.m + 1;
    g.fn = g.prd[1] ? g.n : g.n + 1;
    g.fo = g.prd[2] ? g.o : g.o + 1;
    const double *b = cfg.xyz_bounds;
    double Lx = b[1] - b[0], Ly = b[3] - b[2], Lz = b[5] - b[4];
    g.dx = Lx / g.m; g.dy = Ly / g.n; g.dz = Lz / g.o;
    g.dxsp = 1.0 / g.dx; g.dysp = 1.0 / g.dy; g.dzsp = 1.0 / g.dz;
    if (g.dx != g.dy || g.dx != g.dz) { err = "Spatial discretization must be homogenous."; return RMT3D_ERR_ARG; }
    hx.resize(g.m4); hy.resize(g.n4); hz.resize(g.o4);
    for (int i = 0; i < g.m4; i++) hx[i] = b[0] + (i - 2 + 0.5) * g.dx;
    for (int j = 0; j < g.n4; j++) hy[j] = b[2] + (j - 2 + 0.5) * g.dy;
    for (int k = 0; k < g.o4; k++) hz[k] = b[4] + (k - 2 + 0.5) * g.dz;

    HCHK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    double *dlx = (double *)dalloc(sizeof(double) * g.m4);
    double *dly = (double *)dalloc(sizeof(double) * g.n4);
    double *dlz = (double *)dalloc(sizeof(double) * g.o4);
    if (!dlx || !dly || !dlz) { err = "hipMalloc failed"; return RMT3D_ERR_NOMEM; }
    HCHK(hipMemcpy(dlx, hx.data(), sizeof(double) * g.m4, hipMemcpyHostToDevice));
    HCHK(hipMemcpy(dly, hy.data(), sizeof(double) * g.n4, hipMemcpyHostToDevice));
    HCHK(hipMemcpy(dlz, hz.data(), sizeof(double) * g.o4, hipMemcpyHostToDevice));
    g.lx = dlx; g.ly = dly; g.lz = dlz;

    // sim_manager: materials, objects, setup_const
    double dh = std::min(std::min(g.dx, g.dy), g.dz);
    vmax = 1;
    h_obj.resize(std::max(1, n_obj));
    for (int i = 0; i < n_obj; i++) {
        const rmt3d_object &oo = objs_in[i];
        if (oo.type != RMT3D_BASIC_SPHERE && oo.type != RMT3D_SPHERE) { err = "object type not supported"; return RMT3D_ERR_ARG; }
        ObjC &q = h_obj[i];
        for (int a = 0; a < 3; a++) q.c[a] = oo.basic[a];
        q.R = oo.basic[3];
        q.G = oo.shear_mod;
        q.ex_mu = cfg.ex_visc_mult * cfg.fmu;   // sim_manager.cc:72
        q.ev_trans_mult = cfg.ev_trans_mult;
        q.avg_vel[0] = q.avg_vel[1] = q.avg_vel[2] = 0;
        init.type[i] = oo.type;
        for (int a = 0; a < 3; a++) { init.eig[i][a] = 1; init.omega[i][a] = 0; }
        if (oo.type == RMT3D_SPHERE) {
            int sd = (int)oo.extra[0];
            double sl = oo.extra[1];
            for (int a = 0; a < 3; a++) init.eig[i][a] = sl;
            init.eig[i][sd] = 1. / (sl * sl);
            for (int a = 0; a < 3; a++) init.omega[i][a] = oo.extra[2 + a];
        }
    }
    {
        double xxsp3 = 3.0 / (dh * dh), dt_s = 100000;
        for (int i = 0; i < n_obj; i++) {
            double rho = objs_in[i].srho, G = objs_in[i].shear_mod, cc = sqrt(G / rho);
            double exmu = cfg.ex_visc_mult * cfg.fmu;
            double x3 = 3 / (dh * dh);
            double dt1 = 0.5 * rho / (exmu * (cfg.ev_trans_mult + 1) * x3) * cfg.dt_ex_pad;
            double dt0 = dh / cc * cfg.sdt_pad;
            double odt = (dt0 < dt1) ? dt0 : dt1;
            if (odt < dt_s) dt_s = odt;
        }
        double dt_f = 0.5 * cfg.frho / (cfg.fmu * xxsp3) * cfg.fdt_pad;
        dt_reg = (dt_s < dt_f) ? dt_s : dt_f;
    }
    eps = cfg.wt_n * dh;
    P.eps = eps;
    P.tf1 = 0.5 / eps;
    P.tf2 = 0.5 / M_PI;
    P.tf3 = M_PI / eps;
    if (cfg.dt < dt_reg) dt_reg = cfg.dt;
    K_stiff = 0.0025 / (dt_reg * dt_reg);
    P.fmu = cfg.fmu;
    P.sim_type = st;
    P.n_obj = n_obj;
    P.gravity = 0;
    P.min_wall_acc = 0; P.wall_acc_mult = 0; P.wall_dist = 1; P.winv = 1;
    for (int f = 0; f < 6; f++) { P.walls[f] = 0; P.wall_pos[f] = 0; }
    if (st == RMT3D_SIM_OBJECTS || st == RMT3D_SIM_MILLS) {
        for (int f = 0; f < 6; f++) { P.walls[f] = cfg.walls[f]; P.wall_pos[f] = cfg.wall_pos[f]; }
        P.wall_dist = cfg.wall_dist;
        P.winv = 1 / cfg.wall_dist;
        P.gravity = cfg.gravity;
        P.min_wall_acc = 20 * P.winv;
        double mvsq = -1, mvol = -1, rho_f = cfg.frho;
        for (int i = 0; i < n_obj; i++) {
            double rho_s = objs_in[i].srho, R = objs_in[i].basic[3];
            double vsq = 0.5 * fabs((rho_s - rho_f) * P.gravity) / rho_f;
            double volume = M_PI * 4 / 3 * pow(R, 3);
            double mvf = volume / R / R;
            vsq *= mvf;
            if (vsq > mvsq) mvsq = vsq;
            if (mvf > mvol) mvol = mvf;
            double sa = fabs((rho_s - rho_f) * P.gravity * P.winv * mvf / rho_s);
            if (sa > P.min_wall_acc) P.min_wall_acc = sa;
        }
        P.wall_acc_mult = 0.5 * mvol * pow(P.winv, 2);
        vmax = sqrt(mvsq);
    }
    // setup_faces (fluid_3d.cc:2951-2974)
    for (int f = 0; f < 6; f++) {
        for (int cc = 0; cc < 3; cc++)
            if (cfg.bct[cc][f] == 1 && cfg.bcv[cc][f] > vmax) vmax = cfg.bcv[cc][f];
        if (!g.prd[f / 2]) {
            FaceSpec F;
            F.f = f;
            for (int cc = 0; cc < 3; cc++) { F.bct[cc] = cfg.bct[cc][f]; F.bcv[cc] = cfg.bcv[cc][f]; }
            wall_faces.push_back(F);
        }
    }
    P.K_stiff = K_stiff;
    init.has_exact = cfg.has_exact;
    init.vel_prof_num = cfg.vel_prof_num;
    for (int a = 0; a < 16; a++) init.vel_profile[a] = cfg.vel_profile[a];

    dt = cfg.dt;
    time = 0;
    d_obj = (ObjC *)dalloc(sizeof(ObjC) * h_obj.size());
    HCHK(hipMemcpy(d_obj, h_obj.data(), sizeof(ObjC) * h_obj.size(), hipMemcpyHostToDevice));

    // device state
    int Ng = g.Ng;
    for (int a = 0; a < 3; a++) {
        U.vel[a] = (double *)dalloc(sizeof(double) * Ng);
        U.dvel[a] = (double *)dalloc(sizeof(double) * Ng);
    }
    U.p = (double *)dalloc(sizeof(double) * Ng);
    U.dp = (double *)dalloc(sizeof(double) * Ng);
    U.q = (double *)dalloc(sizeof(double) * Ng);
    for (int a = 0; a < 18; a++) U.fvel[a] = (double *)dalloc(sizeof(double) * Ng);
    for (int a = 0; a < 9; a++) U.sigma[a] = (double *)dalloc(sizeof(double) * Ng);
    R.c = (uint32_t *)dalloc(sizeof(uint32_t) * Ng);
    R.phi = (double *)dalloc(sizeof(double) * Ng);
    for (int a = 0; a < 3; a++) {
        R.x[a] = (double *)dalloc(sizeof(double) * Ng);
        R.xpred[a] = (double *)dalloc(sizeof(double) * Ng);
        R.detF[a] = (double *)dalloc(sizeof(double) * Ng);
        R.en[a] = (double *)dalloc(sizeof(double) * Ng);
    }
    for (int a = 0; a < 18; a++) R.fx[a] = (double *)dalloc(sizeof(double) * Ng);
    for (int a = 0; a < 9; a++) R.gphi[a] = (double *)dalloc(sizeof(double) * Ng);
    if (dalloc_failed) { err = "hipMalloc failed"; return RMT3D_ERR_NOMEM; }

    // band capacity: twice the padded shells of all objects, at least 64k
    double cap_est = 65536;
    for (int i = 0; i < n_obj; i++) {
        double R0 = objs_in[i].basic[3] + (cfg.nlayers + 4) * dh;
        cap_est += 2.0 * (4.0 / 3.0) * M_PI * R0 * R0 * R0 / (dh * dh * dh);
    }
    int cap = (int)std::min<double>(cap_est, (double)Ng * std::max(1, n_obj));
    alloc_band(ext, cap);
    alloc_band(tmp, cap);
    W.mask = (uint32_t *)dalloc(sizeof(uint32_t) * Ng);
    W.flag = (unsigned char *)dalloc(Ng);
    W.lid = (signed char *)dalloc(Ng);
    W.box = (int *)dalloc(sizeof(int) * 8 * std::max(1, n_obj));
    W.batch = (BoxBatch *)dalloc(sizeof(BoxBatch) * std::max(1, n_obj));
    HCHK(hipHostMalloc((void **)&W.h_box, sizeof(int) * 8 * std::max(1, n_obj), hipHostMallocDefault));
    HCHK(hipHostMalloc((void **)&W.h_batch, sizeof(BoxBatch) * std::max(1, n_obj), hipHostMallocDefault));
    for (int a = 0; a < 3; a++) W.val[a] = (double *)dalloc(sizeof(double) * Ng);
    W.list[0] = (int *)dalloc(sizeof(int) * Ng);
    W.list[1] = (int *)dalloc(sizeof(int) * Ng);
    W.cnt = (int *)dalloc(sizeof(int) * 2);
    W.st_cap = cap;
    W.st_cell = (int *)dalloc(sizeof(int) * cap);
    W.st_c = (uint32_t *)dalloc(sizeof(uint32_t) * cap);
    for (int a = 0; a < 3; a++) {
        W.st_x[a] = (double *)dalloc(sizeof(double) * cap);
        W.st_xpred[a] = (double *)dalloc(sizeof(double) * cap);
        W.st_en[a] = (double *)dalloc(sizeof(double) * cap);
        W.st_detF[a] = (double *)dalloc(sizeof(double) * cap);
    }
    W.st_n = (int *)dalloc(sizeof(int));
    W.perm = (int *)dalloc(sizeof(int) * cap);
    W.fill = (int *)dalloc(sizeof(int) * Ng);
    W.scan_tmp = (int *)dalloc(sizeof(int) * ((Ng + 1023) / 1024 + 1));
    W.err = (int *)dalloc(sizeof(int));
    // status word, then the three branch counters of the step (Phys::counters)
    d_status = (int *)dalloc(4 * sizeof(int));
    P.counters = d_status + 1;
    d_solid_list = (int *)dalloc(sizeof(int) * Ng);
    d_solid_n = (int *)dalloc(sizeof(int));
    d_u64 = (unsigned long long *)dalloc(sizeof(unsigned long long) * 16 * INIT_VMAX_SLOTS);
    d_partial = (double *)dalloc(sizeof(double) * 1024);
    d_diag = (double *)dalloc(sizeof(double) * diag_scratch_doubles(n_obj));
    d_mean = (double *)dalloc(sizeof(double) * 4);
    d_faces = dalloc(face_struct_size());
    d_rhs_tmp = (double *)dalloc(sizeof(double) * (size_t)g.fm * g.fn * g.fo * 2);
    {
        double wp[64];
        for (int d = 0; d < 64; d++) wp[d] = pow(cfg.weight_fac, d);
        d_wpow = (double *)dalloc(sizeof(wp));
        HCHK(hipMemcpy(d_wpow, wp, sizeof(wp), hipMemcpyHostToDevice));
    }
    if (dalloc_failed) { err = "hipMalloc failed"; return RMT3D_ERR_NOMEM; }
    HCHK(hipHostMalloc((void **)&h_status, sizeof(int) * 4, hipHostMallocDefault));
    HCHK(hipHostMalloc((void **)&h_int, sizeof(int) * 16, hipHostMallocDefault));
    HCHK(hipHostMalloc((void **)&h_dbl, sizeof(double) * 16 * INIT_VMAX_SLOTS, hipHostMallocDefault));
    // initial gphi = -1e200 for all primary maps (ref_map ctor, fields.hh:324-352)
    {
        std::vector<double> neg((size_t)Ng, -1.e200);
        for (int a = 0; a < 9; a++) HCHK(hipMemcpy(R.gphi[a], neg.data(), sizeof(double) * Ng, hipMemcpyHostToDevice));
    }
    int dm[3] = {g.m, g.n, g.o}, df[3] = {g.fm, g.fn, g.fo};
    bool mixed = (prec == RMT3D_PREC_MIXED);
    mac.setup(0, dm, g.prd, mixed);
    fem.setup(1, df, g.prd, mixed);
    if (dalloc_failed) { err = "hipMalloc failed"; return RMT3D_ERR_NOMEM; }
    if (mac.alloc_failed || fem.alloc_failed) { err = "hipMalloc failed (multigrid levels)"; return RMT3D_ERR_NOMEM; }
    // the level arrays' zero fills ran on the null stream
    HCHK(hipStreamSynchronize(nullptr));
    mac.tune(stream);
    fem.tune(stream);
    dev_bytes += mac.bytes + fem.bytes;
    if (mixed) {
        // defect correction stops on the FP64 residual relative to the rhs
        // scale; the rule itself is mg_prob::run's
        mac.tol_override = -1;
        fem.tol_override = -1;
    }
    // the set-up copies and fills ran on the null stream; the context's own
    // stream is non-blocking, so wait for both (not for other contexts)
    HCHK(hipStreamSynchronize(nullptr));
    HCHK(hipStreamSynchronize(stream));
    return 0;
}

Maps Sim::maps() const {
    Maps M;
    M.r = R;
    M.b = ext;
    M.Ng = g.Ng;
    return M;
}

int Sim::check_status(const char *where) {
    HCHK(hipMemcpyAsync(h_status, d_status, sizeof(int), hipMemcpyDeviceToHost, stream));
    HCHK(hipStreamSynchronize(stream));
    if (!hip_pending.empty()) {
        err = std::string(where) + ": " + hip_pending;
        hip_pending.clear();
        return RMT3D_ERR_HIP;
    }
    int s = *h_status;
    if (!s) return 0;
    err = std::string(where) + ": ";
    if (s & ST_FATAL_NVD) err += "normal_derivatives_refmap: missing necessary cells to carry out finite differences. ";
    if (s & ST_FATAL_UPWIND) err += "godunov_upwinding_set: missing neighbour map. ";
    if (s & ST_FATAL_EXTRAP) err += "Extrapolation failed. ";
    if (s & ST_FATAL_REDRAW) err += "redraw_boundary(): Two deep maps try to be inside a cell at the same time! ";
    if (s & ST_BAND_OVERFLOW) err += "band capacity exceeded. ";
    if (s & ST_FATAL_INIT) err += "init_refmap: overlapping primary points! ";
    return RMT3D_ERR_FATAL;
}

// ---------------------------------------------------------------------------
// ghost fills
// ---------------------------------------------------------------------------
void Sim::comm_fluid(int flags) {
    if (!(g.prd[0] || g.prd[1] || g.prd[2])) return;
    double *d[40];
    int nd = 0;
    if (flags & ORC_VEL) for (int a = 0; a < 3; a++) d[nd++] = U.vel[a];
    if (flags & ORC_PRES) { d[nd++] = U.p; d[nd++] = U.dp; d[nd++] = U.q; }
    if (flags & ORC_FVEL) for (int a = 0; a < 18; a++) d[nd++] = U.fvel[a];
    if (flags & ORC_DVEL) for (int a = 0; a < 3; a++) d[nd++] = U.dvel[a];
    if (flags & ORC_FX) for (int a = 0; a < 18; a++) d[nd++] = R.fx[a];
    uint32_t *u[2];
    int nu = 0;
    launch_periodic(stream, g, d, nd, u, nu, flags == ORC_PRES);
}

void Sim::comm_rm(int flags) {
    if (!(g.prd[0] || g.prd[1] || g.prd[2])) return;
    double *d[40];
    int nd = 0;
    if (flags & ORC_X) for (int a = 0; a < 3; a++) d[nd++] = R.x[a];
    if (flags & ORC_XPRED) for (int a = 0; a < 3; a++) d[nd++] = R.xpred[a];
    uint32_t *u[2];
    int nu = 0;
    if (flags & ORC_C) u[nu++] = R.c;
    launch_periodic(stream, g, d, nd, u, nu, 0);
}

// fluid_3d::set_bc_cc / set_bc_cf (fluid_3d.cc:145-165)
void Sim::fill_bc_cc() {
    launch_bc_cc_faces(stream, g, wall_faces.data(), (int)wall_faces.size(), U);
    comm_fluid(ORC_VEL);
}
void Sim::fill_bc_cf() {
    launch_bc_cf_faces(stream, g, wall_faces.data(), (int)wall_faces.size(), U);
    comm_fluid(ORC_FVEL);
}

// ---------------------------------------------------------------------------
// initialisation (fluid_3d.cc:192-400)
// ---------------------------------------------------------------------------
void Sim::cfl_recheck() {
    double dh = std::min(std::min(g.dx, g.dy), g.dz);
    double dt_min = cfg.fdt_pad * dh / vmax;
    if (dt_reg > dt_min) {
        dt_reg = dt_min;
        if (cfg.dt < dt_reg) dt_reg = cfg.dt;
    }
    K_stiff = 0.0025 / (dt_reg * dt_reg);
    P.K_stiff = K_stiff;
}

void Sim::init_fluid() {
    P.time = time;
    launch_zero_u64(stream, d_u64, 16 * INIT_VMAX_SLOTS);
    launch_init_fluid(stream, g, P, d_obj, init, U, d_u64, fields_written ? 1 : 0);
    comm_fluid(ORC_VEL);
    hipMemcpyAsync(h_dbl, d_u64, sizeof(unsigned long long) * 16 * INIT_VMAX_SLOTS, hipMemcpyDeviceToHost, stream);
    hipStreamSynchronize(stream);
    // the slots hold bit patterns of non-negative doubles (0 where no wavefront
    // wrote), which order like the values
    unsigned long long bits = 0;
    for (int a = 0; a < INIT_VMAX_SLOTS; a++) {
        unsigned long long b;
        memcpy(&b, (const char *)h_dbl + sizeof(b) * 16 * a, sizeof b);
        if (b > bits) bits = b;
    }
    double m;
    memcpy(&m, &bits, sizeof m);
    if (m > vmax) vmax = m;   // sim_manager::velocity (sim_manager.cc:129-130)
    cfl_recheck();
    dt = dt_reg;
}

void Sim::init_pres() {
    launch_init_pres(stream, g, P, U);
    comm_fluid(ORC_PRES);
}

int Sim::init_refmap() {
    launch_init_refmap(stream, g, P, d_obj, init, R, d_status);
    comm_rm(ORC_C);
    comm_rm(ORC_X | ORC_XPRED);
    return 0;
}

void Sim::extrapolate() {
    W.tband = tmp;
    last_batches = band_extrapolate(stream, g, P, d_obj, R, ext, W, cfg.nlayers, cfg.max_extrap_rs, d_wpow, nmask,
                                    d_status, n_obj);
    if (last_batches < 0) hip_pending = "band_extrapolate: copying the object boxes failed";
    band_build_from_staging(stream, g, ext, W, cfg.nlayers, d_status);
}

void Sim::init_extrapolate() {
    extrapolate();
    launch_band_x_eq_xpred(stream, ext);
}

void Sim::compute_grad_basis() {
    double fac = 3 * g.dx / dt;
    for (int i = 0; i < 8; i++) {
        gradphi[i][0] = (2 * (i % 2) - 1) * fac;
        gradphi[i][1] = (2 * ((i / 2) % 2) - 1) * fac;
        gradphi[i][2] = (2 * (i / 4) - 1) * fac;
    }
}

// fluid_3d::initialize_from_chk_point (fluid_3d_chk_pt.cc:74-163): the eight
// files are read on the host, scattered into the constructor-state fields on
// the device, their ghosts filled, and the band rebuilt by a fresh
// extrapolation.  Returns 0, or 1 / 2 (file missing / empty) with the state
// untouched, or a library error.
int Sim::init_from_checkpoint() {
    printf("fluid_3d:: initializing from files in %s/\n", cfg.chk_dirname);
    size_t L = (size_t)(g.m + 1) * (g.n + 1), n8 = 8 * (size_t)g.o * L;
    std::vector<double> img(n8);
    int e = chk_read_files(cfg.chk_dirname, img.data(), g.o, L, err);
    if (e) return e;
    time = cfg.current_time;
    nt = cfg.chk_step;
    double *d = nullptr;
    HCHK(hipMalloc(&d, sizeof(double) * n8));
    hipError_t he = hipMemcpyAsync(d, img.data(), sizeof(double) * n8, hipMemcpyHostToDevice, stream);
    if (he == hipSuccess) {
        launch_zero_i32(stream, d_status, 1);
        fields_written = true;
        launch_chk_reset(stream, g, U, R);
        hipEvent_t e0 = nullptr, e1 = nullptr;
        kprof[7].pair(e0, e1);
        launch_chk_unpack(stream, g, U, R, d, e0, e1);
        he = hipStreamSynchronize(stream);
    }
    hipFree(d);
    HCHK(he);
    comm_fluid(ORC_VEL);
    comm_fluid(ORC_PRES);
    comm_rm(ORC_C);
    comm_rm(ORC_X | ORC_XPRED);
    launch_zero_i32(stream, ext.cnt, g.Ng);
    launch_zero_i32(stream, ext.n, 1);
    init_extrapolate();
    return 0;
}

int Sim::initialize() {
    int init_err = 1;
    if (cfg.chk_num > 0 && cfg.chk_dirname[0]) {
        init_err = init_from_checkpoint();
        if (init_err < 0) return init_err;
        if (init_err == 0) {
            fill_bc_cc();
            if ((cfg.dump_code & 1) && cfg.ntracers > 0) {
                int e = init_tracers();
                if (e) return e;
            }
            compute_grad_basis();
            return check_status("initialize");
        }
    }
    if (cfg.chk_num > 0) printf("# Initializing from scratch, setting chk_num to 0.\n");
    cfg.chk_num = 0;
    time = 0;
    launch_zero_i32(stream, d_status, 1);
    init_fluid();
    init_pres();
    init_refmap();
    launch_zero_i32(stream, ext.cnt, g.Ng);
    launch_zero_i32(stream, ext.n, 1);
    init_extrapolate();
    fill_bc_cc();
    // fluid_3d::alloc_tracers (fluid_3d.cc:115-117, 217-219)
    if ((cfg.dump_code & 1) && cfg.ntracers > 0) {
        int e = init_tracers();
        if (e) return e;
    }
    compute_grad_basis();
    int e = check_status("initialize");
    if (e) return e;
    return init_err;
}

// ---------------------------------------------------------------------------
// step stages
// ---------------------------------------------------------------------------
// stage buckets of fluid_3d::watch (fluid_3d.cc:59): 0 extrapolation,
// 1 pressure Poisson, 2 communication, 3 stresses, 4 MAC, 5 advection
void Sim::mark(int k) { kprof[NPROF_KERNEL + k].begin(stream); }
void Sim::mark_end(int k) { kprof[NPROF_KERNEL + k].end(stream); }

int Sim::mac_solve() {
    launch_mac_rhs(stream, g, U, mac.L[0].b, g.dx / (2 * dt));
    double res = mac.run(stream);
    if (std::isnan(res)) {
        err = "fluid_3d::mac_solve(): NaN detected in multigrid solve";
        return RMT3D_ERR_FATAL;
    }
    launch_mac_extract(stream, g, U, mac.L[0].x);
    comm_fluid(ORC_PRES);
    launch_sub_qgrad(stream, g, U, 2 * dt * g.dxsp);
    return 0;
}

int Sim::pp_solve() {
    launch_pp_rhs(stream, g, U, gradphi, wall_faces.data(), (int)wall_faces.size(), d_faces, fem.L[0].b);
    double res = fem.run(stream);
    if (std::isnan(res)) return 1;
    int nn = g.fm * g.fn * g.fo;
    launch_sum_blocks(stream, fem.L[0].x, nn, 1 / (double)nn, d_partial);
    launch_finish_sum(stream, d_partial, d_mean);
    launch_pp_extract(stream, g, U, fem.L[0].x, d_mean);
    comm_fluid(ORC_PRES);
    launch_sub_pgrad(stream, g, U, dt);
    return 0;
}

int Sim::update_reference_map() {
    launch_xpred(stream, g, P, U, R);
    comm_rm(ORC_C);
    comm_rm(ORC_X | ORC_XPRED);
    std::swap(ext, tmp);
    extrapolate();
    launch_zero_i32(stream, W.err, 1);
    launch_corrector(stream, g, P, d_obj, R, ext, tmp, W.err);
    // the status word travels with the error flag, so that a fatal flag of an
    // earlier stage of this step takes precedence over the error code
    hipMemcpyAsync(h_int, W.err, sizeof(int), hipMemcpyDeviceToHost, stream);
    hipMemcpyAsync(h_status, d_status, sizeof(int), hipMemcpyDeviceToHost, stream);
    hipStreamSynchronize(stream);
    return *h_int;
}

// fluid_3d::step_forward (fluid_3d.cc:429-577), non-TANG, no tracers
int Sim::step(int debug) {
    (void)debug;
    fields_written = true;
    nt++;
    P.dt = dt;
    P.time = time;
    launch_zero_i32(stream, d_status, 1);
    // tracers::predict on the velocities at the start of the step
    // (fluid_3d.cc:458-461)
    if (trace) launch_tracer_predict(stream, TQ, TR, U, dt);
    launch_zero_i32(stream, P.counters, 3);
    Maps M = maps();
    mark(3);
    kprof[2].begin(stream);
    launch_stress(stream, g, P, d_obj, U, M, d_solid_list, d_solid_n);
    kprof[2].end(stream);
    mark_end(3);
    mark(5);
    kprof[3].begin(stream);
    launch_advect(stream, g, P, d_obj, U, M, d_solid_list, d_solid_n, d_status);
    kprof[3].end(stream);
    mark_end(5);
    time += 0.5 * dt;
    P.time = time;
    mark(2);
    fill_bc_cf();
    comm_fluid(ORC_FVEL | ORC_FX);
    launch_periodic_band_fx(stream, g, ext);
    mark_end(2);
    mark(5);
    launch_upwind(stream, g, U, M, d_status);
    mark_end(5);
    mark(2);
    fill_bc_cf();
    mark_end(2);
    // The fatal flags of the advection and upwinding stages are read with
    // the step's final status instead of here: after such a flag the
    // remaining stages still index only valid cells and maps, and the step
    // returns the same fatal error, one host round trip fewer per step.
    mark(4);
    int e = mac_solve();
    mark_end(4);
    if (e) return e;
    mark(2);
    fill_bc_cf();
    mark_end(2);
    mark(0);
    int prob = update_reference_map();
    mark_end(0);
    M = maps();
    if (*h_status) return check_status("step");
    if (prob) {
        launch_revert(stream, g, R, ext, tmp);
        hipStreamSynchronize(stream);
        return RMT3D_REFMAP_UPDATE_ERR;
    }
    mark(3);
    launch_stress(stream, g, P, d_obj, U, M, d_solid_list, d_solid_n);
    mark_end(3);
    mark(5);
    launch_ustar(stream, g, P, d_obj, U, M, d_solid_list, d_solid_n);
    mark_end(5);
    mark(2);
    fill_bc_cc();
    comm_fluid(ORC_DVEL);
    mark_end(2);
    time += 0.5 * dt;
    P.time = time;
    mark(1);
    int pe = pp_solve();
    mark_end(1);
    if (pe) {
        e = check_status("step");
        return e ? e : RMT3D_PROJECTION_ERR;
    }
    mark(0);
    launch_full_update(stream, g, R, ext);
    launch_redraw(stream, g, P, d_obj, R, ext, W, cfg.nlayers, d_status);
    band_build_from_staging(stream, g, ext, W, cfg.nlayers, d_status, 1);
    mark_end(0);
    mark(2);
    fill_bc_cc();
    comm_fluid(ORC_DVEL);
    mark_end(2);
    // tracers::advect on the end-of-step velocities (fluid_3d.cc:569-572)
    if (trace) launch_tracer_advect(stream, TQ, TR, U, dt);
    e = check_status("step");
    if (e) return e;
    return RMT3D_NORMAL_EXIT;
}

// fluid_3d::div_u, L2 part (fluid_3d_utils.cc:507-544)
double Sim::div_u() {
    int nn = g.fm * g.fn * g.fo;
    double *r = d_rhs_tmp, *t = d_rhs_tmp + nn;
    launch_pp_rhs(stream, g, U, gradphi, wall_faces.data(), (int)wall_faces.size(), d_faces, r);
    launch_divu_terms(stream, g, r, t);
    launch_sum_blocks(stream, t, nn, 1.0, d_partial);
    launch_finish_sum(stream, d_partial, d_mean);
    hipMemcpyAsync(h_dbl, d_mean, sizeof(double), hipMemcpyDeviceToHost, stream);
    hipStreamSynchronize(stream);
    double l2 = h_dbl[0];
    l2 /= nn;
    return sqrt(l2);
}

// fluid_3d::init_iter (fluid_3d.cc:235-287)
int Sim::init_iter(int init_err) {
    if (init_err > 0) {
        double tol = 1e-4;
        for (int it = 0; it < cfg.num_iters; it++) {
            time = 0;
            fill_bc_cc();
            int e = step(cfg.debug_flag);
            if (e < 0) return e;
            double d = div_u();
            init_fluid();
            init_refmap();
            launch_zero_i32(stream, ext.cnt, g.Ng);
            launch_zero_i32(stream, ext.n, 1);
            init_extrapolate();
            if (d < tol) break;
        }
        nt = 0;
        time = 0.0;
    }
    P.dt = dt;
    P.time = time;
    launch_stress(stream, g, P, d_obj, U, maps(), d_solid_list, d_solid_n);
    return check_status("init_iter");
}

// run_sim.cc:98-106 + stepper ctor (stepper.hh:211-225) + stepper::run set_dt
int Sim::prepare_run(int init_err, double *dt_out, int *total) {
    double sdt = cfg.dt;
    if (dt_reg < sdt) sdt = dt_reg;
    cfg.dt = sdt;
    double df = cfg.T / (cfg.frames - 1);
    int steps = (int)(df / sdt);
    double rem = df - steps * sdt;
    if (rem > 0) sdt -= (sdt - rem) / (++steps);
    // a resumed run has its first chk_num frames behind it (stepper.hh:
    // 214-216, 358-361); stop_short counts from the resumed step
    int tot = (cfg.frames - 1 - cfg.chk_num) * steps;
    if (tot < 0) tot = 0;
    dt = sdt;
    cfg.dt = sdt;
    int e = init_iter(init_err);
    if (e) return e;
    dt = sdt;
    if (dt_out) *dt_out = sdt;
    if (total) *total = (cfg.stop_short > 0 && cfg.stop_short < tot) ? cfg.stop_short : tot;
    return 0;
}

// ---------------------------------------------------------------------------
// state transfer
// ---------------------------------------------------------------------------
// Copies run on the context's own stream and wait only for it, so other
// contexts on the device are neither waited for nor disturbed.
int Sim::upload(const rmt3d_state *st) {
    fields_written = true;
    size_t n = (size_t)g.Ng;
    auto up = [&](void *dst, const void *src, size_t bytes) {
        if (src) hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, stream);
    };
    size_t w = sizeof(double) * n;
    up(U.vel[0], st->u, w); up(U.vel[1], st->v, w); up(U.vel[2], st->w, w);
    up(U.p, st->p, w); up(U.dp, st->dp, w); up(U.q, st->q, w);
    if (st->dvel) for (int a = 0; a < 3; a++) up(U.dvel[a], st->dvel + a * n, w);
    if (st->fvel) for (int a = 0; a < 18; a++) up(U.fvel[a], st->fvel + a * n, w);
    if (st->sigma) for (int a = 0; a < 9; a++) up(U.sigma[a], st->sigma + a * n, w);
    up(R.c, st->rm_c, sizeof(uint32_t) * n);
    if (st->rm_x) for (int a = 0; a < 3; a++) up(R.x[a], st->rm_x + a * n, w);
    if (st->rm_xpred) for (int a = 0; a < 3; a++) up(R.xpred[a], st->rm_xpred + a * n, w);
    HCHK(hipStreamSynchronize(stream));
    if (st->band_cell || st->band_c || st->band_x || st->band_xpred) return upload_band(st);
    return 0;
}

// Replaces the band by st's nband entries (cells ascending, the entries of a
// cell in multimap order), as a freshly added extrap_mmap entry: face values
// zero, face normals not computed (-1e200, fields.hh:311-330), det F and
// energies zero.  All four band arrays are required.
int Sim::upload_band(const rmt3d_state *st) {
    int nb = st->nband;
    if (!st->band_cell || !st->band_c || !st->band_x || !st->band_xpred || nb < 0) {
        err = "band upload needs band_cell, band_c, band_x and band_xpred";
        return RMT3D_ERR_ARG;
    }
    if (nb > ext.cap) {
        err = "band upload exceeds the band capacity";
        return RMT3D_ERR_ARG;
    }
    std::vector<int> cnt((size_t)g.Ng, 0), start((size_t)g.Ng, 0);
    for (int e = 0; e < nb; e++) {
        int c = st->band_cell[e];
        if (c < 0 || c >= g.Ng || (e > 0 && c < st->band_cell[e - 1])) {
            err = "band cells must be ascending and inside the padded grid";
            return RMT3D_ERR_ARG;
        }
        cnt[c]++;
    }
    for (int c = 0, run = 0; c < g.Ng; c++) { start[c] = run; run += cnt[c]; }
    size_t wb = sizeof(double) * (size_t)nb;
    std::vector<double> nc((size_t)nb, -1.e200);
    HCHK(hipMemcpyAsync(ext.start, start.data(), sizeof(int) * (size_t)g.Ng, hipMemcpyHostToDevice, stream));
    HCHK(hipMemcpyAsync(ext.cnt, cnt.data(), sizeof(int) * (size_t)g.Ng, hipMemcpyHostToDevice, stream));
    HCHK(hipMemcpyAsync(ext.n, &nb, sizeof(int), hipMemcpyHostToDevice, stream));
    if (nb > 0) {
        HCHK(hipMemcpyAsync(ext.cell, st->band_cell, sizeof(int) * (size_t)nb, hipMemcpyHostToDevice, stream));
        HCHK(hipMemcpyAsync(ext.c, st->band_c, sizeof(uint32_t) * (size_t)nb, hipMemcpyHostToDevice, stream));
        for (int a = 0; a < 3; a++) {
            HCHK(hipMemcpyAsync(ext.x[a], st->band_x + (size_t)a * nb, wb, hipMemcpyHostToDevice, stream));
            HCHK(hipMemcpyAsync(ext.xpred[a], st->band_xpred + (size_t)a * nb, wb, hipMemcpyHostToDevice, stream));
            HCHK(hipMemsetAsync(ext.detF[a], 0, wb, stream));
            HCHK(hipMemsetAsync(ext.en[a], 0, wb, stream));
        }
        for (int a = 0; a < 18; a++) HCHK(hipMemsetAsync(ext.fx[a], 0, wb, stream));
        for (int a = 0; a < 9; a++)
            HCHK(hipMemcpyAsync(ext.gphi[a], nc.data(), wb, hipMemcpyHostToDevice, stream));
    }
    HCHK(hipStreamSynchronize(stream));
    return 0;
}

// the branch counters of the last step (Phys::counters)
int Sim::step_counters(int out[3]) {
    HCHK(hipMemcpyAsync(h_int, P.counters, 3 * sizeof(int), hipMemcpyDeviceToHost, stream));
    HCHK(hipStreamSynchronize(stream));
    for (int a = 0; a < 3; a++) out[a] = h_int[a];
    return 0;
}

int Sim::band_count() {
    hipMemcpyAsync(h_int, ext.n, sizeof(int), hipMemcpyDeviceToHost, stream);
    hipStreamSynchronize(stream);
    return h_int[0];
}

int Sim::download(rmt3d_state *st) {
    size_t n = (size_t)g.Ng;
    auto dn = [&](void *dst, const void *src, size_t bytes) {
        if (dst) hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, stream);
    };
    size_t w = sizeof(double) * n;
    dn(st->u, U.vel[0], w); dn(st->v, U.vel[1], w); dn(st->w, U.vel[2], w);
    dn(st->p, U.p, w); dn(st->dp, U.dp, w); dn(st->q, U.q, w);
    if (st->dvel) for (int a = 0; a < 3; a++) dn(st->dvel + a * n, U.dvel[a], w);
    if (st->fvel) for (int a = 0; a < 18; a++) dn(st->fvel + a * n, U.fvel[a], w);
    if (st->sigma) for (int a = 0; a < 9; a++) dn(st->sigma + a * n, U.sigma[a], w);
    dn(st->rm_c, R.c, sizeof(uint32_t) * n);
    if (st->rm_x) for (int a = 0; a < 3; a++) dn(st->rm_x + a * n, R.x[a], w);
    if (st->rm_xpred) for (int a = 0; a < 3; a++) dn(st->rm_xpred + a * n, R.xpred[a], w);
    int nb = band_count();
    if (st->band_cell || st->band_c || st->band_x || st->band_xpred) {
        if (st->nband < nb) {
            st->nband = nb;
            err = "band arrays too small";
            return RMT3D_ERR_ARG;
        }
        size_t wb = sizeof(double) * nb;
        dn(st->band_cell, ext.cell, sizeof(int) * nb);
        dn(st->band_c, ext.c, sizeof(uint32_t) * nb);
        for (int a = 0; a < 3; a++) {
            if (st->band_x) dn(st->band_x + (size_t)a * nb, ext.x[a], wb);
            if (st->band_xpred) dn(st->band_xpred + (size_t)a * nb, ext.xpred[a], wb);
        }
    }
    st->nband = nb;
    HCHK(hipStreamSynchronize(stream));
    return 0;
}

// ---------------------------------------------------------------------------
// macro diagnostics (stepper.hh:378-475, fluid_3d_utils.cc:507-1346)
// ---------------------------------------------------------------------------
int Sim::macro_data(rmt3d_macro *m) {
    double l2 = div_u();
    size_t nd = diag_scratch_doubles(n_obj);
    double *out = d_diag + nd - 16;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    kprof[8].pair(e0, e1);
    launch_macro(stream, g, U, maps(), cfg.gravity, cfg.xyz_bounds[4], d_diag, out, e0, e1);
    double r[10];
    HCHK(hipMemcpyAsync(h_dbl, out, sizeof r, hipMemcpyDeviceToHost, stream));
    HCHK(hipStreamSynchronize(stream));
    memcpy(r, h_dbl, sizeof r);
    double dv = g.dx * g.dy * g.dz;
    m->divu_l2 = l2;
    m->divu_max = r[7];
    m->divu_min = -r[8];
    // total_momentum: sum times dx*dy*dz, left to right (fluid_3d_utils.cc:799)
    for (int a = 0; a < 3; a++) m->momentum[a] = r[a] * g.dx * g.dy * g.dz;
    // total_energy (fluid_3d_utils.cc:858-869)
    m->elastic = r[5] * dv;
    m->kinetic = r[3] * dv;
    m->potential = r[4] * dv;
    m->power = 0. * dv;
    m->energy = m->elastic + m->kinetic + m->potential;
    // total_dev_in_detF (fluid_3d_utils.cc:1068-1084)
    m->dev_detF_l2 = sqrt(r[6] * dv);
    double mx = r[9];
    if (mx == -1.0) mx = 0;
    m->max_dev_detF = sqrt(mx);
    return 0;
}

int Sim::centroids(double *out) {
    if (n_obj <= 0) return 0;
    size_t nd = diag_scratch_doubles(n_obj);
    double *res = d_diag + nd - 16 - 7 * (size_t)n_obj;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    kprof[9].pair(e0, e1);
    launch_centroids(stream, g, P, d_obj, U, R, n_obj, d_diag, res, e0, e1);
    std::vector<double> r(7 * (size_t)n_obj);
    HCHK(hipMemcpyAsync(r.data(), res, sizeof(double) * r.size(), hipMemcpyDeviceToHost, stream));
    HCHK(hipStreamSynchronize(stream));
    // sim_manager::compute_avgs (sim_manager.hh:223-246); the average
    // velocity feeds solid_acceleration's wall repulsion from now on
    for (int i = 0; i < n_obj; i++) {
        const double *q = &r[7 * (size_t)i];
        double N = q[0];
        ObjC &o = h_obj[i];
        for (int a = 0; a < 3; a++) o.avg_vel[a] = (N != 0.) ? q[1 + a] / N : 0.;
        if (out) {
            for (int a = 0; a < 3; a++) out[6 * i + a] = (N != 0.) ? q[4 + a] / N : 0.;
            for (int a = 0; a < 3; a++) out[6 * i + 3 + a] = o.avg_vel[a];
        }
    }
    HCHK(hipMemcpyAsync(d_obj, h_obj.data(), sizeof(ObjC) * n_obj, hipMemcpyHostToDevice, stream));
    HCHK(hipStreamSynchronize(stream));
    return 0;
}

// ---------------------------------------------------------------------------
// checkpoint and slice output (fluid_3d_chk_pt.cc:3-72, fluid_3d_io.cc:3-418)
// ---------------------------------------------------------------------------
// ---------------------------------------------------------------------------
// tracers (Utilities/tracers.{hh,cc})
// ---------------------------------------------------------------------------
namespace {
// gsl_rng_taus (GSL 2.x, taus.c): L'Ecuyer's maximally equidistributed
// combined Tausworthe generator (Math. Comp. 65 (1996) 203-213) with GSL's
// seeding by the LCG 69069 and six warm-up draws; uniform() is the 32-bit
// output over 2^32
struct Taus {
    uint32_t s1, s2, s3;
    explicit Taus(uint32_t seed) {
        if (seed == 0) seed = 1;
        s1 = 69069u * seed;
        if (s1 < 2) s1 += 2;
        s2 = 69069u * s1;
        if (s2 < 8) s2 += 8;
        s3 = 69069u * s2;
        if (s3 < 16) s3 += 16;
        for (int i = 0; i < 6; i++) next();
    }
    uint32_t next() {
        s1 = (((s1 & 4294967294u) << 12) ^ (((s1 << 13) ^ s1) >> 19));
        s2 = (((s2 & 4294967288u) << 4) ^ (((s2 << 2) ^ s2) >> 25));
        s3 = (((s3 & 4294967280u) << 17) ^ (((s3 << 3) ^ s3) >> 11));
        return s1 ^ s2 ^ s3;
    }
    double uniform() { return next() / 4294967296.0; }
};
}  // namespace

// tracers::tracers (tracers.cc:7-66): ntracers uniform random positions in
// the box (taus seeded with 1), or the positions of the checkpoint's
// tr.<chk_num>; tracers::set (tracers.hh:160-172) places each in its cube
int Sim::init_tracers() {
    int n = cfg.ntracers;
    const double h[3] = {g.dx, g.dy, g.dz};
    const int cells[3] = {g.m, g.n, g.o};
    const double low[3] = {hx[0], hy[0], hz[0]};
    for (int f = 0; f < 3; f++) {
        TQ.h[f] = h[f];
        TQ.low[f] = low[f];
        double glow = cfg.xyz_bounds[2 * f] + (0 + 0.5) * h[f];   // gx[0] (fluid_3d.cc:132)
        TQ.lo[f] = glow - 0.5 * h[f];
        TQ.hi[f] = TQ.lo[f] + cells[f] * h[f];
        TQ.prd[f] = g.prd[f];
        for (int lr = 0; lr < 2; lr++) {
            TQ.bct[f][lr] = cfg.bct[f][2 * f + lr];
            TQ.bcv[f][lr] = cfg.bcv[f][2 * f + lr];
        }
    }
    TQ.lim[0] = g.m4; TQ.lim[1] = g.n4; TQ.lim[2] = g.o4;
    std::vector<double> pos(3 * (size_t)n);
    if (cfg.chk_num <= 0) {
        Taus rng(1);
        double ax = (cfg.xyz_bounds[0] + 0.5 * h[0]) - 0.5 * h[0];
        double ay = (cfg.xyz_bounds[2] + 0.5 * h[1]) - 0.5 * h[1];
        double az = (cfg.xyz_bounds[4] + 0.5 * h[2]) - 0.5 * h[2];
        double Lx = g.m * h[0], Ly = g.n * h[1], Lz = g.o * h[2];
        for (int i = 0; i < n; i++) {
            double x = ax + Lx * rng.uniform();
            double y = ay + Ly * rng.uniform();
            double z = az + Lz * rng.uniform();
            pos[3 * i] = x; pos[3 * i + 1] = y; pos[3 * i + 2] = z;
        }
    } else {
        // tracers::init_from_chkpt_files (tracers.cc:142-172)
        char fn[600];
        snprintf(fn, sizeof fn, "%s/tr.%05d", cfg.chk_dirname, cfg.chk_num);
        printf("tracers:: initializing tracers from %s\n", fn);
        FILE *fh = fopen(fn, "rb");
        if (!fh) { err = std::string("cannot open ") + fn; return RMT3D_ERR_FATAL; }
        int lnum = 0;
        std::vector<float> tout(3 * (size_t)n);
        size_t r1 = fread(&lnum, sizeof(int), 1, fh);
        size_t r2 = (r1 == 1 && lnum == n) ? fread(tout.data(), sizeof(float), 3 * (size_t)n, fh) : 0;
        fclose(fh);
        if (r1 != 1 || lnum != n || r2 != 3 * (size_t)n) {
            err = "tracer:: Reading check point file failed, or read size doesn't match that specified in cfg file.";
            return RMT3D_ERR_FATAL;
        }
        for (size_t a = 0; a < 3 * (size_t)n; a++) pos[a] = tout[a];
    }
    std::vector<int> alive(n), pc[3];
    std::vector<double> rp[3];
    for (int f = 0; f < 3; f++) { pc[f].assign(n, 0); rp[f].assign(n, 0.0); }
    for (int i = 0; i < n; i++) {
        bool ok = true;
        for (int f = 0; f < 3 && ok; f++) {
            double t_ind = (pos[3 * i + f] - low[f]) / h[f] - 1;
            int p = (int)t_ind;
            double r = t_ind - p - 0.5;
            pc[f][i] = p;
            rp[f][i] = r;
            bool lo = (p < 0) || (p == 0 && r < 0);
            bool hi = (p > TQ.lim[f] - 4) || (p == TQ.lim[f] - 4 && r >= 0);
            if (lo || hi) ok = false;
        }
        alive[i] = ok;
    }
    if (!TR.alive || TR.n != n) {
        TR.n = n;
        TR.alive = (int *)dalloc(sizeof(int) * n);
        for (int f = 0; f < 3; f++) {
            TR.p[f] = (int *)dalloc(sizeof(int) * n);
            TR.rp0[f] = (double *)dalloc(sizeof(double) * n);
            TR.rpf[f] = (double *)dalloc(sizeof(double) * n);
            TR.v0[f] = (double *)dalloc(sizeof(double) * n);
            TR.vf[f] = (double *)dalloc(sizeof(double) * n);
        }
        if (dalloc_failed) { err = "hipMalloc failed"; return RMT3D_ERR_NOMEM; }
    }
    HCHK(hipMemcpyAsync(TR.alive, alive.data(), sizeof(int) * n, hipMemcpyHostToDevice, stream));
    for (int f = 0; f < 3; f++) {
        HCHK(hipMemcpyAsync(TR.p[f], pc[f].data(), sizeof(int) * n, hipMemcpyHostToDevice, stream));
        HCHK(hipMemcpyAsync(TR.rp0[f], rp[f].data(), sizeof(double) * n, hipMemcpyHostToDevice, stream));
        HCHK(hipMemcpyAsync(TR.rpf[f], rp[f].data(), sizeof(double) * n, hipMemcpyHostToDevice, stream));
    }
    HCHK(hipStreamSynchronize(stream));
    trace = true;
    return 0;
}

// tracers::print (tracers.cc:110-139): single-precision positions by id,
// 0 for a tracer that has left the box
int Sim::tracer_positions(std::vector<float> &xyz) {
    int n = TR.n;
    xyz.assign(3 * (size_t)n, 0.f);
    if (!trace || n <= 0) return 0;
    std::vector<int> alive(n), pc[3];
    std::vector<double> rp[3];
    HCHK(hipMemcpyAsync(alive.data(), TR.alive, sizeof(int) * n, hipMemcpyDeviceToHost, stream));
    for (int f = 0; f < 3; f++) {
        pc[f].resize(n);
        rp[f].resize(n);
        HCHK(hipMemcpyAsync(pc[f].data(), TR.p[f], sizeof(int) * n, hipMemcpyDeviceToHost, stream));
        HCHK(hipMemcpyAsync(rp[f].data(), TR.rp0[f], sizeof(double) * n, hipMemcpyDeviceToHost, stream));
    }
    HCHK(hipStreamSynchronize(stream));
    for (int i = 0; i < n; i++)
        if (alive[i])
            for (int f = 0; f < 3; f++) xyz[3 * (size_t)i + f] = (float)(TQ.low[f] + (1.5 + rp[f][i] + pc[f][i]) * TQ.h[f]);
    return 0;
}

int Sim::write_tracers(const char *filename, int binary) {
    std::vector<float> xyz;
    int e = tracer_positions(xyz);
    if (e) return e;
    FILE *fh = fopen(filename, binary ? "wb" : "w");
    if (!fh) { err = std::string("cannot open ") + filename; return RMT3D_ERR_ARG; }
    int num = TR.n;
    if (binary) {
        fwrite(&num, sizeof(int), 1, fh);
        fwrite(xyz.data(), sizeof(float), xyz.size(), fh);
    } else {
        for (size_t a = 0; a < xyz.size(); a += 3) fprintf(fh, "%g %g %g\n", xyz[a], xyz[a + 1], xyz[a + 2]);
    }
    return fclose(fh) == 0 ? 0 : RMT3D_ERR_ARG;
}

// ---------------------------------------------------------------------------
// contours (fluid_3d_io.cc:595-958)
// ---------------------------------------------------------------------------
int Sim::contour(int cnum, std::vector<float> &pts, std::vector<float> &nrm, std::vector<int> &tri) {
    size_t nn = (size_t)(g.m + 1) * (g.n + 1) * (g.o + 1);
    std::vector<void *> mem;
    auto get = [&](size_t bytes) -> void * {
        void *q = nullptr;
        if (hipMalloc(&q, bytes ? bytes : 8) != hipSuccess) { (void)hipGetLastError(); return nullptr; }
        mem.push_back(q);
        return q;
    };
    auto release = [&]() { for (void *q : mem) hipFree(q); };
    CtBufs B{};
    B.phi = (double *)get(sizeof(double) * g.Ng);
    B.vcnt = (int *)get(sizeof(int) * nn);
    B.voff = (int *)get(sizeof(int) * nn);
    B.tcnt = (int *)get(sizeof(int) * nn);
    B.toff = (int *)get(sizeof(int) * nn);
    B.vidx = (int *)get(sizeof(int) * 3 * nn);
    B.bsum = (int *)get(sizeof(int) * (nn / 1024 + 2));
    B.tot = (int *)get(sizeof(int) * 2);
    // cell-centre coordinates gx, gy, gz (fluid_3d.cc:132-134), continued to
    // index m, where the reference reads past its arrays
    std::vector<double> hc[3];
    const int cells[3] = {g.m, g.n, g.o};
    const double hh[3] = {g.dx, g.dy, g.dz};
    double *dc[3];
    for (int a = 0; a < 3; a++) {
        hc[a].resize(cells[a] + 1);
        for (int i = 0; i <= cells[a]; i++) hc[a][i] = cfg.xyz_bounds[2 * a] + (i + 0.5) * hh[a];
        dc[a] = (double *)get(sizeof(double) * (cells[a] + 1));
    }
    for (void *q : mem)
        if (!q) { release(); err = "hipMalloc failed"; return RMT3D_ERR_NOMEM; }
    for (int a = 0; a < 3; a++)
        hipMemcpyAsync(dc[a], hc[a].data(), sizeof(double) * (cells[a] + 1), hipMemcpyHostToDevice, stream);
    B.gx = dc[0]; B.gy = dc[1]; B.gz = dc[2];
    launch_contour_count(stream, g, d_obj, n_obj, maps(), cnum, B);
    int tot[2] = {0, 0};
    hipMemcpyAsync(h_int, B.tot, sizeof(int) * 2, hipMemcpyDeviceToHost, stream);
    hipError_t he = hipStreamSynchronize(stream);
    if (he != hipSuccess) { release(); HCHK(he); }
    tot[0] = h_int[0];
    tot[1] = h_int[1];
    B.pts = (float *)get(sizeof(float) * 3 * (size_t)tot[0]);
    B.nrm = (float *)get(sizeof(float) * 3 * (size_t)tot[0]);
    B.tri = (int *)get(sizeof(int) * 3 * (size_t)tot[1]);
    if (!B.pts || !B.nrm || !B.tri) { release(); err = "hipMalloc failed"; return RMT3D_ERR_NOMEM; }
    launch_contour_emit(stream, g, B);
    pts.resize(3 * (size_t)tot[0]);
    nrm.resize(3 * (size_t)tot[0]);
    tri.resize(3 * (size_t)tot[1]);
    hipMemcpyAsync(pts.data(), B.pts, sizeof(float) * pts.size(), hipMemcpyDeviceToHost, stream);
    hipMemcpyAsync(nrm.data(), B.nrm, sizeof(float) * nrm.size(), hipMemcpyDeviceToHost, stream);
    hipMemcpyAsync(tri.data(), B.tri, sizeof(int) * tri.size(), hipMemcpyDeviceToHost, stream);
    he = hipStreamSynchronize(stream);
    release();
    HCHK(he);
    return 0;
}

// the ctr file: the object count, then per object the vertex and triangle
// counts and, when there are vertices, positions, normals (floats) and the
// triangles' vertex indices (fluid_3d_io.cc:595-848)
int Sim::write_contours(const char *filename) {
    FILE *fh = fopen(filename, "wb");
    if (!fh) { err = std::string("cannot open ") + filename; return RMT3D_ERR_ARG; }
    int no = n_obj;
    fwrite(&no, sizeof(int), 1, fh);
    for (int i = 0; i < n_obj; i++) {
        std::vector<float> pts, nrm;
        std::vector<int> tri;
        int e = contour(i + 1, pts, nrm, tri);
        if (e) { fclose(fh); return e; }
        int od[2] = {(int)(pts.size() / 3), (int)(tri.size() / 3)};
        fwrite(od, sizeof(int), 2, fh);
        if (od[0] == 0) continue;
        fwrite(pts.data(), sizeof(float), pts.size(), fh);
        fwrite(nrm.data(), sizeof(float), nrm.size(), fh);
        fwrite(tri.data(), sizeof(int), tri.size(), fh);
    }
    return fclose(fh) == 0 ? 0 : RMT3D_ERR_ARG;
}

int Sim::write_checkpoint(int snum, int fnum, const char *dir) {
    size_t L = (size_t)(g.m + 1) * (g.n + 1), n8 = 8 * (size_t)g.o * L;
    std::vector<double> img(n8);
    double *d = nullptr;
    HCHK(hipMalloc(&d, sizeof(double) * n8));
    hipEvent_t e0 = nullptr, e1 = nullptr;
    kprof[6].pair(e0, e1);
    launch_chk_pack(stream, g, d_obj, n_obj, U, maps(), d, e0, e1);
    hipError_t he = hipMemcpyAsync(img.data(), d, sizeof(double) * n8, hipMemcpyDeviceToHost, stream);
    if (he == hipSuccess) he = hipStreamSynchronize(stream);
    hipFree(d);
    HCHK(he);
    // the run's tracer file of this frame goes with the checkpoint
    // (fluid_3d_chk_pt.cc:28-32)
    if (cfg.dirname[0]) {
        char src[600], dst[600];
        snprintf(src, sizeof src, "%s/tr.%05d", cfg.dirname, fnum);
        snprintf(dst, sizeof dst, "%s/tr.%05d", dir, fnum);
        if (FILE *in = fopen(src, "rb")) {
            FILE *out = fopen(dst, "wb");
            char buf[65536];
            size_t k;
            while (out && (k = fread(buf, 1, sizeof buf, in)) > 0) fwrite(buf, 1, k, out);
            if (out) fclose(out);
            fclose(in);
        }
    }
    int e = chk_write_files(dir, img.data(), g.o, L, err);
    if (e) return e;
    cfg.chk_num = fnum;
    cfg.chk_step = snum;
    cfg.current_time = time;
    e = rmt3d_write_params(&cfg, objs_in.data(), dir);
    if (e) err = std::string("cannot write ") + dir + "/recover.cfg";
    return e;
}

// write_params::corner_field (IO/write_params.hh:66-69)
static bool corner_field(int t) { return (t >= 3 && t < 6) || t == 7 || t == 18; }

int Sim::slice_dims(int dim, int t, int &d1, int &d2) const {
    if (dim < 0 || dim > 2 || t < 0 || t > 21) return RMT3D_ERR_ARG;
    int gd[3] = {g.m, g.n, g.o};
    if (corner_field(t)) { gd[0] = g.fm; gd[1] = g.fn; gd[2] = g.fo; }
    gd[dim] = 1;
    static const int it1[3] = {1, 0, 0}, it2[3] = {2, 2, 1};
    d1 = gd[it1[dim]];
    d2 = gd[it2[dim]];
    return 0;
}

int Sim::slice(int dim, int point, int t, int obj_id, double *out, int *len) { return slice_into(dim, point, t, obj_id, out, len, false); }

int Sim::slice_noalloc(int dim, int point, int t, int obj_id, double *out, int *len) {
    return slice_into(dim, point, t, obj_id, out, len, true);
}

// fluid_3d::slice on the device.  The slice is gathered into a device buffer
// and copied out: a buffer allocated for the call, or (noalloc) the
// library's fixed per-GPU scratch, in which case nothing is allocated and a
// slice larger than the scratch is refused.
int Sim::slice_into(int dim, int point, int t, int obj_id, double *out, int *len, bool noalloc) {
    int d1, d2;
    if (slice_dims(dim, t, d1, d2)) { err = "slice: bad dimension or output type"; return RMT3D_ERR_ARG; }
    size_t n = (size_t)d1 * d2;
    if (!out || !len || (size_t)*len < n) {
        if (len) *len = (int)n;
        err = "slice: output array too small";
        return RMT3D_ERR_ARG;
    }
    *len = (int)n;
    int cells[3] = {g.m, g.n, g.o};
    if (point < 0 || point >= cells[dim]) {
        // copy_slice_to_buf refuses the point and the zeroed slice is kept
        printf("copy_slice_to_buf(): Error! Index out of bound. Nothing has been done.\n");
        for (size_t a = 0; a < n; a++) out[a] = 0.;
        return 0;
    }
    double *d = nullptr;
    if (noalloc) {
        if (n > SLICE_SCRATCH_MAX) {
            err = "slice_noalloc: " + std::to_string(n) + " values exceed the fixed scratch of " +
                  std::to_string(SLICE_SCRATCH_MAX);
            return RMT3D_ERR_ARG;
        }
        d = slice_scratch(err);
        if (!d) return RMT3D_ERR_HIP;
    } else {
        HCHK(hipMalloc(&d, sizeof(double) * n));
    }
    const double *r0 = nullptr;
    if (t == 18) {
        launch_pp_rhs(stream, g, U, gradphi, wall_faces.data(), (int)wall_faces.size(), d_faces, d_rhs_tmp);
        r0 = d_rhs_tmp;
    }
    launch_slice(stream, g, d_obj, n_obj, U, maps(), r0, t, obj_id, dim, point, d1, d2, d);
    hipError_t he = hipMemcpyAsync(out, d, sizeof(double) * n, hipMemcpyDeviceToHost, stream);
    if (he == hipSuccess) he = hipStreamSynchronize(stream);
    if (!noalloc) hipFree(d);
    HCHK(he);
    return 0;
}

int Sim::write_slice(int dim, int point, int t, int obj_id, int format, const char *filename) {
    int d1, d2;
    if (slice_dims(dim, t, d1, d2) || format < 0 || format > 2 || !filename) {
        err = "write_slice: bad arguments";
        return RMT3D_ERR_ARG;
    }
    std::vector<double> v((size_t)d1 * d2);
    int len = (int)v.size();
    int e = slice(dim, point, t, obj_id, v.data(), &len);
    if (e) return e;
    // cell-centre coordinates gx, gy, gz (fluid_3d.cc:132-134), continued
    // past the last cell for the node counts of the corner fields
    int cells[3] = {g.m, g.n, g.o};
    double h[3] = {g.dx, g.dy, g.dz};
    std::vector<double> crd[3];
    for (int a = 0; a < 3; a++) {
        int nn = cells[a] + 1;
        crd[a].resize(nn);
        for (int i = 0; i < nn; i++) crd[a][i] = cfg.xyz_bounds[2 * a] + (i + 0.5) * h[a];
    }
    const double *a1 = crd[0].data(), *a2 = crd[1].data();
    if (dim == 0) { a1 = crd[1].data(); a2 = crd[2].data(); }
    else if (dim == 1) a2 = crd[2].data();
    e = slice_write_file(filename, format, v.data(), d1, d2, a1, a2, g.m);
    if (e) err = std::string("cannot write ") + filename;
    return e;
}

int Sim::mg_solve(int which, const double *rhs, double *x, double tol, int *ncyc, double *res) {
    MGSolver &G = which ? fem : mac;
    size_t nn = (size_t)G.L[0].nn;
    HCHK(hipMemcpyAsync(G.L[0].b, rhs, sizeof(double) * nn, hipMemcpyHostToDevice, stream));
    HCHK(hipMemcpyAsync(G.L[0].x, x, sizeof(double) * nn, hipMemcpyHostToDevice, stream));
    double save = G.tol_override;
    if (tol > 0) G.tol_override = tol;
    double r = G.run(stream);
    G.tol_override = save;
    HCHK(hipMemcpyAsync(x, G.L[0].x, sizeof(double) * nn, hipMemcpyDeviceToHost, stream));
    HCHK(hipStreamSynchronize(stream));
    if (ncyc) *ncyc = G.last_cycles;
    if (res) *res = r;
    return 0;
}

}  // namespace rmt
