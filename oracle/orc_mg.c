/*
 * ORACLE -- test infrastructure only (see orc.h).
 *
 * Geometric multigrid for the two projection problems of the step.
 *
 * Operators (reference, closed form):
 *  MAC  simp_mac::setup (Utilities/mg_probs.hh:505-527): 7-point stencil,
 *       centre -6, faces +1, cell-centred unknowns m x n x o; at a wall the
 *       missing face entry is added to the centre (stencil_set::set_partials
 *       with macfix, mg_probs.hh:281-343) = homogeneous Neumann.
 *  FEM  simp_fem::setup (mg_probs.hh:466-492): 27-point nodal Q1 stencil,
 *       centre 32, edges -2, corners -1, faces 0; at a wall node every entry
 *       whose offset is 0 in that wall's direction is halved and outward
 *       entries are absent (set_partials(false)).
 *  Periodic dimensions wrap.
 *
 * Solve loop: mg_prob::run (Utilities/mg_probs.cc:3-83) with the mg_pred
 * iteration predictor (mg_probs.hh:37-78), max-norm stopping rule, hard cap
 * of 20 extra batches and two insurance cycles; v_up = v_down = 2
 * (mg_opt defaults, mg_probs.hh:92-93).
 *
 * V-cycle (this project's design; PGMG's is absent, so its iterates are
 * parity unpinned): MAC -- red-black SOR (omega 1.25), restriction = sum of the
 * 8 children / 2, cell-centred trilinear prolongation; FEM -- four-colour
 * Gauss-Seidel by (j,k) parity (whole x-rows per colour, valid because the
 * Q1 face weights vanish), full-weighting restriction (= P^T) with the
 * Galerkin coarse operator, which for nested Q1 spaces equals twice the fine
 * stencil (coarse rhs = R r / 2), vertex trilinear prolongation.  The
 * hierarchy ends at the first level with at most ORC_MG_DIRECT_NN unknowns,
 * solved exactly: x = A^+ b with the Moore-Penrose pseudo-inverse (the
 * minimum-norm least-squares solution, so the singular all-Neumann and
 * periodic problems need no special case).  A hierarchy whose first coarse
 * level (or a later one) has more than ORC_MG_SMALL_NN and at most
 * ORC_MG_FDM_NN unknowns ends there instead, solved exactly by fast
 * diagonalisation (Lynch, Rice & Thomas, Numer. Math. 6, 1964): both
 * operators are sums of Kronecker products of 1D matrices -- MAC
 * T x I x I + I x T x I + I x I x T with T the 1D second difference (Neumann
 * ends folded), FEM 12 (K x M x M + M x K x M + M x M x K) with K, M the 1D
 * Q1 stiffness and mass matrices of unit elements (Neumann ends halved) -- so
 * with per-dimension bases S (S^T M S = I, S^T K S = diag(lam); M = I for
 * MAC) x = (Sx.Sy.Sz) D^+ (Sx.Sy.Sz)^T b, D = sum of the three lam (times
 * 12 for FEM), one contraction per dimension each way.  The 1D bases are
 * closed forms: DCT-I on the Q1 nodes of a wall dimension, DCT-II on the
 * cells of a MAC wall dimension, the real Fourier basis on a periodic one.
 * The null mode (all three lam zero) is dropped.  A hierarchy that reaches
 * neither ends with 4 Gauss-Seidel sweeps on its coarsest level.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <float.h>
#include "orc.h"
#include "orc_detmath.h"

#define NCOARSE_SWEEPS 4
/* over-relaxation of the MAC red-black sweeps: 1.25 brings the V-cycles per
 * solve from about 13 to 9-10 on the sphere and Taylor-Green workloads */
#define MAC_OMEGA 1.25

/* a periodic dimension is halved only while the half stays even, so that
 * the parity colourings of the smoothers stay valid across the wrap */
static int can_coarsen(int fem, int d, int prd)
{
    if (prd) return (d % 4 == 0) && d / 2 >= 2;
    if (!fem) return (d % 2 == 0) && d / 2 >= 2;
    return (d % 2 == 1) && (d - 1) / 2 + 1 >= 2 && d >= 3;
}

static int coarse_dim(int fem, int d, int prd)
{
    if (!fem || prd) return d / 2;
    return (d - 1) / 2 + 1;
}

/* The 1D basis of one dimension of the fast-diagonalisation level:
 * S[q*d + p] = mode p at point q, lam[p] its eigenvalue.  Angles are reduced
 * exactly in integers before det_cos / det_sin.  Modes:
 *   FEM wall (d nodes, n = d-1 unit elements): cos(pi p q / n),
 *     lam = 6 (1 - cos(pi p / n)) / (2 + cos(pi p / n))  (K v = lam M v);
 *   MAC wall (d cells): cos(pi p (2q+1) / 2d), lam = -2 (1 - cos(pi p / d));
 *   periodic: p = 0 constant, p = 2k-1 cos(2 pi k q / d), p = 2k
 *     sin(2 pi k q / d), the FEM or MAC lam of the angle 2 pi k / d.
 * Each mode is scaled to unit M-norm (FEM: M = tridiag(1, 4, 1) / 6 with
 * wall ends 2/6; MAC: M = I), summed in point order. */
static void fdm_basis(int fem, int d, int prd, double *S, double *lam)
{
    const double pi = 3.14159265358979323846;
    double *v = (double *)malloc(sizeof(double) * (size_t)d);
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
            if (!fem) { nrm += v[q] * v[q]; continue; }
            int lo = q - 1, hi = q + 1;
            int end = 0;
            if (prd) { lo = (lo + d) % d; hi = hi % d; }
            else { end = (q == 0 || q == d - 1); }
            double mv = (end ? 2.0 : 4.0) * v[q];
            if (lo >= 0) mv += v[lo];
            if (hi < d) mv += v[hi];
            nrm += v[q] * (mv / 6.0);
        }
        double sq = sqrt(nrm);
        for (int q = 0; q < d; q++) S[(size_t)q * d + p] = v[q] / sq;
    }
    free(v);
}

static void fdm_setup(orc_mg *G, const orc_mglev *L)
{
    double *lam[3];
    for (int a = 0; a < 3; a++) {
        int d = L->d[a];
        G->fS[a] = (double *)malloc(sizeof(double) * (size_t)d * d);
        G->fST[a] = (double *)malloc(sizeof(double) * (size_t)d * d);
        lam[a] = (double *)malloc(sizeof(double) * (size_t)d);
        fdm_basis(G->fem, d, G->prd[a], G->fS[a], lam[a]);
        for (int q = 0; q < d; q++)
            for (int p = 0; p < d; p++) G->fST[a][(size_t)p * d + q] = G->fS[a][(size_t)q * d + p];
    }
    G->fdinv = (double *)malloc(sizeof(double) * (size_t)L->nn);
    double scale = G->fem ? 12.0 : 1.0;
    for (int k = 0; k < L->d[2]; k++)
        for (int j = 0; j < L->d[1]; j++)
            for (int i = 0; i < L->d[0]; i++) {
                double D = scale * ((lam[0][i] + lam[1][j]) + lam[2][k]);
                G->fdinv[i + L->d[0] * (j + L->d[1] * k)] = (D == 0) ? 0.0 : 1.0 / D;
            }
    for (int a = 0; a < 3; a++) free(lam[a]);
}

void orc_mg_setup(orc_mg *G, int fem, const int d[3], const int prd[3])
{
    memset(G, 0, sizeof *G);
    G->fem = fem;
    for (int a = 0; a < 3; a++) G->prd[a] = prd[a];
    int cur[3] = {d[0], d[1], d[2]};
    int l = 0;
    for (;;) {
        orc_mglev *L = &G->lev[l];
        for (int a = 0; a < 3; a++) L->d[a] = cur[a];
        L->nn = cur[0] * cur[1] * cur[2];
        L->x = (double *)calloc((size_t)L->nn, sizeof(double));
        L->b = (double *)calloc((size_t)L->nn, sizeof(double));
        L->r = (double *)calloc((size_t)L->nn, sizeof(double));
        l++;
        if (L->nn <= ORC_MG_DIRECT_NN) break;
        if (l > 1 && L->nn > ORC_MG_SMALL_NN && L->nn <= ORC_MG_FDM_NN && cur[0] <= ORC_MG_FDM_DMAX &&
            cur[1] <= ORC_MG_FDM_DMAX && cur[2] <= ORC_MG_FDM_DMAX) {
            G->fdm = 1;
            break;
        }
        int ok = l < 16;
        for (int a = 0; a < 3; a++) ok = ok && can_coarsen(fem, cur[a], prd[a]);
        if (!ok) break;
        for (int a = 0; a < 3; a++) cur[a] = coarse_dim(fem, cur[a], prd[a]);
    }
    G->nlev = l;
    G->pinv = NULL;
    if (G->fdm) {
        fdm_setup(G, &G->lev[l - 1]);
    } else if (G->lev[l - 1].nn <= ORC_MG_DIRECT_NN) {
        /* the coarsest operator column by column, then its pseudo-inverse */
        int n = G->lev[l - 1].nn;
        double *A = (double *)calloc((size_t)n * n, sizeof(double));
        double *e = (double *)calloc((size_t)n, sizeof(double));
        double *y = (double *)calloc((size_t)n, sizeof(double));
        for (int c = 0; c < n; c++) {
            e[c] = 1;
            orc_mg_apply(G, l - 1, e, y);
            for (int r = 0; r < n; r++) A[r * n + c] = y[r];
            e[c] = 0;
        }
        G->pinv = (double *)malloc(sizeof(double) * (size_t)n * n);
        orc_sym_pinv(n, A, G->pinv);
        free(A); free(e); free(y);
    }
    double eps = DBL_EPSILON;
    double c = fem ? 32.0 : 6.0;
    G->tol = c * c * eps * eps * 1e4;
    G->max_norm_tol = c * eps * 1e2;
    G->log_tol = log10(G->tol);
    G->v_up = 2;
    G->coarse_sweeps = NCOARSE_SWEEPS;
    G->v_down = 2;
    G->iP = 16;
    G->iV = 2 * G->iP;
    G->V_avg = G->iV;
    G->V_var = 1e6;
    G->alpha = 1 / (3 * G->iP); /* integer division in the reference: 0 */
}

void orc_mg_free(orc_mg *G)
{
    for (int l = 0; l < G->nlev; l++) {
        free(G->lev[l].x); free(G->lev[l].b); free(G->lev[l].r);
    }
    free(G->pinv);
    G->pinv = NULL;
    for (int a = 0; a < 3; a++) {
        free(G->fS[a]); free(G->fST[a]);
        G->fS[a] = G->fST[a] = NULL;
    }
    free(G->fdinv);
    G->fdinv = NULL;
    G->fdm = 0;
    G->nlev = 0;
}

/* Moore-Penrose pseudo-inverse of a symmetric matrix from its eigen-
 * decomposition by cyclic Jacobi rotations (Golub & Van Loan, Matrix
 * Computations, 4th ed., sec. 8.5): rotations in (p, q) row order until the
 * off-diagonal mass falls below 1e-28 of the diagonal's, then
 * P = sum over eigenvalues |l_i| > 1e-10 max|l| of v_i v_i^T / l_i, summed
 * in eigenvalue order; P is computed on and above the diagonal and mirrored,
 * so it is exactly symmetric. */
void orc_sym_pinv(int n, const double *A0, double *P)
{
    size_t nn = (size_t)n * n;
    double *A = (double *)malloc(sizeof(double) * nn);
    double *V = (double *)calloc(nn, sizeof(double));
    memcpy(A, A0, sizeof(double) * nn);
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
    for (int i = 0; i < n; i++) {
        double a = fabs(A[(size_t)i * n + i]);
        if (a > lmax) lmax = a;
    }
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
    free(A);
    free(V);
}

/* neighbour index along one dimension: -1 if outside a wall */
static inline int nb(int i, int di, int d, int prd)
{
    int j = i + di;
    if (j < 0) return prd ? j + d : -1;
    if (j >= d) return prd ? j - d : -1;
    return j;
}

/* ---- MAC operator ---- */
static inline double mac_Ax(const orc_mglev *L, const int *prd, const double *x, int i, int j, int k, double *diag, double *offsum)
{
    int D0 = L->d[0], D1 = L->d[1], D2 = L->d[2];
    int c = i + D0 * (j + D1 * k);
    double s = 0;
    int miss = 0;
    int ii;
    ii = nb(i, -1, D0, prd[0]); if (ii < 0) miss++; else s += x[ii + D0 * (j + D1 * k)];
    ii = nb(i, 1, D0, prd[0]);  if (ii < 0) miss++; else s += x[ii + D0 * (j + D1 * k)];
    ii = nb(j, -1, D1, prd[1]); if (ii < 0) miss++; else s += x[i + D0 * (ii + D1 * k)];
    ii = nb(j, 1, D1, prd[1]);  if (ii < 0) miss++; else s += x[i + D0 * (ii + D1 * k)];
    ii = nb(k, -1, D2, prd[2]); if (ii < 0) miss++; else s += x[i + D0 * (j + D1 * ii)];
    ii = nb(k, 1, D2, prd[2]);  if (ii < 0) miss++; else s += x[i + D0 * (j + D1 * ii)];
    double dg = -6.0 + miss;
    if (diag) *diag = dg;
    if (offsum) *offsum = s;
    return s + dg * x[c];
}

/* ---- FEM operator ---- */
static inline double fem_w(int oi, int oj, int ok)
{
    int nz = (oi != 0) + (oj != 0) + (ok != 0);
    if (nz == 0) return 32.0;
    if (nz == 1) return 0.0;
    if (nz == 2) return -2.0;
    return -1.0;
}

static inline double fem_Ax(const orc_mglev *L, const int *prd, const double *x, int i, int j, int k, double *diag, double *offsum)
{
    int D0 = L->d[0], D1 = L->d[1], D2 = L->d[2];
    int c[3] = {i, j, k}, D[3] = {D0, D1, D2};
    int wall[3];
    for (int a = 0; a < 3; a++) wall[a] = !prd[a] && (c[a] == 0 || c[a] == D[a] - 1);
    double s = 0, dg = 0;
    for (int ok = -1; ok <= 1; ok++) {
        int kk = nb(k, ok, D2, prd[2]);
        if (kk < 0) continue;
        for (int oj = -1; oj <= 1; oj++) {
            int jj = nb(j, oj, D1, prd[1]);
            if (jj < 0) continue;
            for (int oi = -1; oi <= 1; oi++) {
                int ii = nb(i, oi, D0, prd[0]);
                if (ii < 0) continue;
                double w = fem_w(oi, oj, ok);
                if (w == 0.0) continue;
                double f = 1;
                if (wall[0] && oi == 0) f *= 2;
                if (wall[1] && oj == 0) f *= 2;
                if (wall[2] && ok == 0) f *= 2;
                w /= f;
                if (oi == 0 && oj == 0 && ok == 0) dg = w;
                else s += w * x[ii + D0 * (jj + D1 * kk)];
            }
        }
    }
    if (diag) *diag = dg;
    if (offsum) *offsum = s;
    return s + dg * x[i + D0 * (j + D1 * k)];
}

void orc_mg_apply(const orc_mg *G, int lev, const double *x, double *y)
{
    const orc_mglev *L = &G->lev[lev];
#pragma omp parallel for schedule(static)
    for (int k = 0; k < L->d[2]; k++)
        for (int j = 0; j < L->d[1]; j++)
            for (int i = 0; i < L->d[0]; i++)
                y[i + L->d[0] * (j + L->d[1] * k)] = G->fem ? fem_Ax(L, G->prd, x, i, j, k, NULL, NULL)
                                                            : mac_Ax(L, G->prd, x, i, j, k, NULL, NULL);
}

static void residual(orc_mg *G, int lev)
{
    orc_mglev *L = &G->lev[lev];
#pragma omp parallel for schedule(static)
    for (int k = 0; k < L->d[2]; k++)
        for (int j = 0; j < L->d[1]; j++)
            for (int i = 0; i < L->d[0]; i++) {
                int c = i + L->d[0] * (j + L->d[1] * k);
                double ax = G->fem ? fem_Ax(L, G->prd, L->x, i, j, k, NULL, NULL)
                                   : mac_Ax(L, G->prd, L->x, i, j, k, NULL, NULL);
                L->r[c] = L->b[c] - ax;
            }
}

double orc_mg_resid_max(orc_mg *G, int lev)
{
    residual(G, lev);
    orc_mglev *L = &G->lev[lev];
    double mx = 0;
    for (int c = 0; c < L->nn; c++) {
        double a = fabs(L->r[c]);
        if (a > mx || isnan(a)) mx = isnan(a) ? NAN : a;
        if (isnan(mx)) return mx;
    }
    return mx;
}

/* one Gauss-Seidel sweep: MAC red-black by (i+j+k) parity; FEM four colours
 * by (j,k) parity -- the Q1 stencil's face weights are zero, so nodes that
 * differ only in i are not coupled and each colour is a set of whole x-rows */
static void smooth(orc_mg *G, int lev)
{
    orc_mglev *L = &G->lev[lev];
    int ncol = G->fem ? 4 : 2;
    /* the nodes of one colour are not coupled: any order, any thread */
    for (int col = 0; col < ncol; col++)
#pragma omp parallel for schedule(static)
        for (int k = 0; k < L->d[2]; k++)
            for (int j = 0; j < L->d[1]; j++)
                for (int i = 0; i < L->d[0]; i++) {
                    int mycol = G->fem ? ((j & 1) + 2 * (k & 1)) : ((i + j + k) & 1);
                    if (mycol != col) continue;
                    int c = i + L->d[0] * (j + L->d[1] * k);
                    double dg, off;
                    if (G->fem) fem_Ax(L, G->prd, L->x, i, j, k, &dg, &off);
                    else mac_Ax(L, G->prd, L->x, i, j, k, &dg, &off);
                    double xn = (L->b[c] - off) / dg;
                    if (G->fem) L->x[c] = xn;
                    else L->x[c] = (1 - MAC_OMEGA) * L->x[c] + MAC_OMEGA * xn;
                }
}

/* restriction of level lev's residual into level lev+1's rhs */
static void restrict_res(orc_mg *G, int lev)
{
    orc_mglev *F = &G->lev[lev], *C = &G->lev[lev + 1];
#pragma omp parallel for schedule(static)
    for (int K = 0; K < C->d[2]; K++)
        for (int J = 0; J < C->d[1]; J++)
            for (int I = 0; I < C->d[0]; I++) {
                double s = 0;
                if (!G->fem) {
                    for (int c = 0; c < 2; c++)
                        for (int b = 0; b < 2; b++)
                            for (int a = 0; a < 2; a++)
                                s += F->r[(2 * I + a) + F->d[0] * ((2 * J + b) + F->d[1] * (2 * K + c))];
                    s *= 0.5;
                } else {
                    for (int dc = -1; dc <= 1; dc++) {
                        int kk = nb(2 * K, dc, F->d[2], G->prd[2]);
                        if (kk < 0) continue;
                        for (int db = -1; db <= 1; db++) {
                            int jj = nb(2 * J, db, F->d[1], G->prd[1]);
                            if (jj < 0) continue;
                            for (int da = -1; da <= 1; da++) {
                                int ii = nb(2 * I, da, F->d[0], G->prd[0]);
                                if (ii < 0) continue;
                                double w = (da ? 0.5 : 1.0) * (db ? 0.5 : 1.0) * (dc ? 0.5 : 1.0);
                                s += w * F->r[ii + F->d[0] * (jj + F->d[1] * kk)];
                            }
                        }
                    }
                    s *= 0.5;
                }
                int ci = I + C->d[0] * (J + C->d[1] * K);
                C->b[ci] = s;
                C->x[ci] = 0;
            }
}

/* cell-centred 1D prolongation weights: fine index f = 2I + a */
static inline void cc_w(int f, int D, int prd, int *i0, int *i1, double *w0, double *w1)
{
    int I = f >> 1, a = f & 1;
    int o = a ? I + 1 : I - 1;
    int oo = nb(I, o - I, D, prd);
    *i0 = I; *w0 = 0.75;
    if (oo < 0) { *i1 = I; *w1 = 0.25; } else { *i1 = oo; *w1 = 0.25; }
}

/* vertex 1D prolongation: fine index f */
static inline void vx_w(int f, int D, int prd, int *i0, int *i1, double *w0, double *w1)
{
    int I = f >> 1;
    if ((f & 1) == 0) { *i0 = I; *i1 = I; *w0 = 1.0; *w1 = 0.0; return; }
    int J = I + 1;
    if (J >= D) J = prd ? J - D : D - 1;
    *i0 = I; *i1 = J; *w0 = 0.5; *w1 = 0.5;
}

static void prolong_add(orc_mg *G, int lev)
{
    orc_mglev *F = &G->lev[lev], *C = &G->lev[lev + 1];
#pragma omp parallel for schedule(static)
    for (int k = 0; k < F->d[2]; k++) {
        int k0, k1; double wk0, wk1;
        if (G->fem) vx_w(k, C->d[2], G->prd[2], &k0, &k1, &wk0, &wk1);
        else cc_w(k, C->d[2], G->prd[2], &k0, &k1, &wk0, &wk1);
        for (int j = 0; j < F->d[1]; j++) {
            int j0, j1; double wj0, wj1;
            if (G->fem) vx_w(j, C->d[1], G->prd[1], &j0, &j1, &wj0, &wj1);
            else cc_w(j, C->d[1], G->prd[1], &j0, &j1, &wj0, &wj1);
            for (int i = 0; i < F->d[0]; i++) {
                int i0, i1; double wi0, wi1;
                if (G->fem) vx_w(i, C->d[0], G->prd[0], &i0, &i1, &wi0, &wi1);
                else cc_w(i, C->d[0], G->prd[0], &i0, &i1, &wi0, &wi1);
                const double *e = C->x;
                int D0 = C->d[0], D1 = C->d[1];
#define E(a, b, c) e[(a) + D0 * ((b) + D1 * (c))]
                double v = wk0 * (wj0 * (wi0 * E(i0, j0, k0) + wi1 * E(i1, j0, k0)) +
                                  wj1 * (wi0 * E(i0, j1, k0) + wi1 * E(i1, j1, k0))) +
                           wk1 * (wj0 * (wi0 * E(i0, j0, k1) + wi1 * E(i1, j0, k1)) +
                                  wj1 * (wi0 * E(i0, j1, k1) + wi1 * E(i1, j1, k1)));
#undef E
                F->x[i + F->d[0] * (j + F->d[1] * k)] += v;
            }
        }
    }
}

/* out[.., p, ..] = sum over q of W[q*d + p] in[.., q, ..] along dimension a,
 * each output summed over ascending q from +0, times dinv[.] when given.
 * Along y and z the loops run over whole x-lines (or xy-planes) per term,
 * which keeps each output's order of operations. */
static void fdm_contract(const orc_mglev *L, int a, const double *W, const double *in, double *out,
                         const double *dinv)
{
    int d0 = L->d[0], d1 = L->d[1], d2 = L->d[2], da = L->d[a];
    if (a == 0) {
#pragma omp parallel for schedule(static)
        for (int row = 0; row < d1 * d2; row++) {
            const double *src = in + (size_t)row * d0;
            double *dst = out + (size_t)row * d0;
            for (int p = 0; p < d0; p++) {
                double s = 0;
                for (int q = 0; q < d0; q++) s += W[(size_t)q * d0 + p] * src[q];
                if (dinv) s *= dinv[(size_t)row * d0 + p];
                dst[p] = s;
            }
        }
        return;
    }
    int inner = a == 1 ? d0 : d0 * d1, nouter = a == 1 ? d2 : 1;
#pragma omp parallel for collapse(2) schedule(static)
    for (int o = 0; o < nouter; o++)
        for (int p = 0; p < da; p++) {
            const double *src = in + (size_t)o * da * inner;
            size_t off = ((size_t)o * da + p) * inner;
            double *dst = out + off;
            for (int ii = 0; ii < inner; ii++) dst[ii] = 0;
            for (int q = 0; q < da; q++) {
                double w = W[(size_t)q * da + p];
                const double *sq = src + (size_t)q * inner;
                for (int ii = 0; ii < inner; ii++) dst[ii] += w * sq[ii];
            }
            if (dinv)
                for (int ii = 0; ii < inner; ii++) dst[ii] *= dinv[off + ii];
        }
}

/* x = S D^+ S^T b: forward along x, y, z (the last scaled by D^+), back
 * along z, y, x, through the level's r and x arrays */
static void fdm_solve(orc_mg *G, orc_mglev *L)
{
    fdm_contract(L, 0, G->fS[0], L->b, L->r, NULL);
    fdm_contract(L, 1, G->fS[1], L->r, L->x, NULL);
    fdm_contract(L, 2, G->fS[2], L->x, L->r, G->fdinv);
    fdm_contract(L, 2, G->fST[2], L->r, L->x, NULL);
    fdm_contract(L, 1, G->fST[1], L->x, L->r, NULL);
    fdm_contract(L, 0, G->fST[0], L->r, L->x, NULL);
}

/* Test hooks on a single d[0] x d[1] x d[2] level: y = A x, and the
 * fast-diagonalisation solve x of A x = b. */
static void one_level(orc_mg *G, int fem, const int d[3], const int prd[3])
{
    memset(G, 0, sizeof *G);
    G->fem = fem;
    G->nlev = 1;
    orc_mglev *L = &G->lev[0];
    for (int a = 0; a < 3; a++) { G->prd[a] = prd[a]; L->d[a] = d[a]; }
    L->nn = d[0] * d[1] * d[2];
    L->x = (double *)calloc((size_t)L->nn, sizeof(double));
    L->b = (double *)calloc((size_t)L->nn, sizeof(double));
    L->r = (double *)calloc((size_t)L->nn, sizeof(double));
}

void orc_level_apply(int fem, const int d[3], const int prd[3], const double *x, double *y)
{
    orc_mg G;
    one_level(&G, fem, d, prd);
    orc_mg_apply(&G, 0, x, y);
    orc_mg_free(&G);
}

void orc_fdm_solve_level(int fem, const int d[3], const int prd[3], const double *b, double *x)
{
    orc_mg G;
    one_level(&G, fem, d, prd);
    orc_mglev *L = &G.lev[0];
    fdm_setup(&G, L);
    G.fdm = 1;
    memcpy(L->b, b, sizeof(double) * (size_t)L->nn);
    fdm_solve(&G, L);
    memcpy(x, L->x, sizeof(double) * (size_t)L->nn);
    orc_mg_free(&G);
}

static void vcycle_rec(orc_mg *G, int lev)
{
    if (lev == G->nlev - 1) {
        orc_mglev *L = &G->lev[lev];
        if (G->fdm) {
            fdm_solve(G, L);
        } else if (G->pinv) {
            /* x = A^+ b, each row summed in column order */
            int n = L->nn;
            for (int r = 0; r < n; r++) {
                double s = 0;
                for (int c = 0; c < n; c++) s += G->pinv[(size_t)r * n + c] * L->b[c];
                L->x[r] = s;
            }
        } else {
            for (int s = 0; s < G->coarse_sweeps; s++) smooth(G, lev);
        }
        return;
    }
    for (int s = 0; s < G->v_down; s++) smooth(G, lev);
    residual(G, lev);
    restrict_res(G, lev);
    vcycle_rec(G, lev + 1);
    prolong_add(G, lev);
    for (int s = 0; s < G->v_up; s++) smooth(G, lev);
}

void orc_mg_vcycle(orc_mg *G) { vcycle_rec(G, 0); }

/* mg_pred::update(int) (mg_probs.hh:51-66) */
static void pred_update(orc_mg *G, int extras)
{
    double dV = G->iV - G->V_avg;
    int exreduce = (int)(dV / sqrt(G->V_var));
    exreduce = exreduce > 0 ? exreduce : 0;
    G->iV += G->iP * extras * (extras + 1) / 2 - 1 - exreduce;
    if (G->iV < G->iP) G->iV = 2 * G->iP - 1;
    G->V_avg = (1 - G->alpha) * G->V_avg + G->alpha * G->iV;
    G->V_var = (1 - G->alpha) * (G->V_var + G->alpha * dV * dV);
}

/* mg_prob::run (mg_probs.cc:3-83), profile = false.  Returns the max-norm
 * residual measured before the two insurance cycles, as the reference. */
double orc_mg_run(orc_mg *G)
{
    const int hard_cap = 20;
    int cycles = 0;
    double res = orc_mg_resid_max(G, 0);
    int pre = G->iV / G->iP;
    for (int i = 0; i < pre; i++) { orc_mg_vcycle(G); cycles++; }
    res = orc_mg_resid_max(G, 0);
    int extra = 0;
    while (res > G->max_norm_tol && extra < hard_cap) {
        extra++;
        for (int n = 0; n < extra; n++) { orc_mg_vcycle(G); cycles++; }
        res = orc_mg_resid_max(G, 0);
    }
    for (int i = 0; i < 2; i++) { orc_mg_vcycle(G); cycles++; }
    pred_update(G, extra);
    G->last_cycles = cycles;
    G->last_res = res;
    return res;
}
