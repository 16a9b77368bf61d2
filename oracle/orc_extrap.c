/*
 * ORACLE -- test infrastructure only (see orc.h).
 *
 * Band ("extraps") container, reference-map lookups and the layered
 * least-squares extrapolation of the reference map:
 *   get_refmap / get_refmap_prev       F3D_Source/fluid_3d_utils.cc:31-84
 *   extrapolate / first_layer /        F3D_Source/fluid_3d_ex.cc:31-81
 *   gather_and_list / transfer         F3D_Source/fluid_3d_ex.cc:86-328
 *   redraw_boundary                    F3D_Source/fluid_3d_ex.cc:336-394
 *   extrap_helper (masks, BFS lists,   Extrapolation/extrap_helper.hh:166-412,
 *   least_square, lin_sys, dot)        Extrapolation/extrap_helper.cc:6-225
 *   mat3::inverse                      Utilities/mat3.hh:61-67
 */
#include <limits.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "orc.h"

#define OMASK 1023u
#define LDIG 10

/* ------------------------------------------------------------------ */
/* band container                                                       */
/* ------------------------------------------------------------------ */
void orc_band_init(orc_band *B, int Ng)
{
    B->n = 0;
    B->cap = 1024;
    B->e = (orc_bentry *)calloc((size_t)B->cap, sizeof(orc_bentry));
    B->start = (int *)calloc((size_t)Ng, sizeof(int));
    B->cnt = (int *)calloc((size_t)Ng, sizeof(int));
}

void orc_band_free(orc_band *B)
{
    free(B->e); free(B->start); free(B->cnt);
    memset(B, 0, sizeof *B);
}

/* extrap_mmap::reset (extrap_mmap.hh:72-75) */
void orc_band_reset(orc_band *B, int Ng)
{
    B->n = 0;
    memset(B->cnt, 0, sizeof(int) * (size_t)Ng);
}

static void band_reserve(orc_band *B, int n)
{
    if (n <= B->cap) return;
    while (B->cap < n) B->cap *= 2;
    B->e = (orc_bentry *)realloc(B->e, sizeof(orc_bentry) * (size_t)B->cap);
}

/* extrap_mmap::copy (extrap_mmap.hh:55-64): per-cell order is preserved */
void orc_band_copy(orc_band *dst, const orc_band *src, int Ng)
{
    band_reserve(dst, src->n);
    dst->n = src->n;
    memcpy(dst->e, src->e, sizeof(orc_bentry) * (size_t)src->n);
    memcpy(dst->start, src->start, sizeof(int) * (size_t)Ng);
    memcpy(dst->cnt, src->cnt, sizeof(int) * (size_t)Ng);
}

/* Rebuild a band from entries whose per-cell order is their order in the
 * array: stable counting sort by cell. */
static void band_build(orc_band *B, const orc_bentry *e, int n, int Ng)
{
    band_reserve(B, n);
    memset(B->cnt, 0, sizeof(int) * (size_t)Ng);
    for (int a = 0; a < n; a++) B->cnt[e[a].cell]++;
    int s = 0;
    for (int c = 0; c < Ng; c++) { B->start[c] = s; s += B->cnt[c]; }
    int *fill = (int *)calloc((size_t)Ng, sizeof(int));
    for (int a = 0; a < n; a++) {
        int c = e[a].cell;
        B->e[B->start[c] + fill[c]++] = e[a];
    }
    free(fill);
    B->n = n;
}

/* ------------------------------------------------------------------ */
/* lookups (fluid_3d_utils.cc:31-84)                                    */
/* ------------------------------------------------------------------ */
static orc_rmap *band_find(orc_band *B, int ind, unsigned oid)
{
    for (int a = 0; a < B->cnt[ind]; a++) {
        orc_rmap *r = &B->e[B->start[ind] + a].r;
        if ((r->c & OMASK) == oid) return r;
    }
    return NULL;
}

orc_rmap *orc_get_refmap(orc_sim *S, int ind, unsigned oid)
{
    if (ind < 0 || ind >= S->Ng) return NULL;
    if (S->rm[ind].c == oid) return S->rm + ind;
    return band_find(&S->ext, ind, oid);
}

orc_rmap *orc_get_refmap_pos(orc_sim *S, int ind, unsigned oid, const int pos[3])
{
    if (!(pos[0] >= 0 && pos[0] <= S->m4 - 1 && pos[1] >= 0 && pos[1] <= S->n4 - 1 &&
          pos[2] >= 0 && pos[2] <= S->o4 - 1))
        return NULL;
    return orc_get_refmap(S, ind, oid);
}

orc_rmap *orc_get_refmap_prev(orc_sim *S, int ind, unsigned oid)
{
    if (ind < 0 || ind >= S->Ng) return NULL;
    if (S->rm[ind].c == oid) return S->rm + ind;
    return band_find(&S->tmp, ind, oid);
}

/* ------------------------------------------------------------------ */
/* extrap_helper                                                        */
/* ------------------------------------------------------------------ */
static inline void unpack3(const orc_sim *S, int ind, int *i, int *j, int *k)
{
    *k = ind / S->mn4;
    *j = (ind % S->mn4) / S->m4;
    *i = ind % S->m4;
}

/* extrap_helper::within_extrapolation_domain (extrap_helper.hh:295-404) on
 * one rank: a ghost region is skipped when it has a (periodic, self)
 * neighbour; the far corner tag 26 tests neigh[26] > 0, which rank 0 never
 * satisfies, so that region is always extrapolated. */
static int within_extrap_domain(const orc_sim *S, int ind)
{
    int c[3];
    unpack3(S, ind, &c[0], &c[1], &c[2]);
    int s[3] = {S->m, S->n, S->o};
    int t[3], any = 0, has = 1;
    for (int d = 0; d < 3; d++) {
        t[d] = (c[d] < 2) ? 0 : ((c[d] > s[d] + 1) ? 2 : 1);
        if (t[d] != 1) { any = 1; if (!S->prd[d]) has = 0; }
    }
    if (!any) return 1;
    if (t[0] == 2 && t[1] == 2 && t[2] == 2) return 1;
    return has ? 0 : 1;
}

static inline int within_inner_ghost(const orc_sim *S, int ind)
{
    int i, j, k;
    unpack3(S, ind, &i, &j, &k);
    return i >= 1 && i <= S->m4 - 2 && j >= 1 && j <= S->n4 - 2 && k >= 1 && k <= S->o4 - 2;
}

static int cmp_int(const void *a, const void *b)
{
    int x = *(const int *)a, y = *(const int *)b;
    return (x > y) - (x < y);
}

/* extrap_helper::update_list / unique_indices (extrap_helper.cc:6-29) */
static void update_list(orc_sim *S)
{
    qsort(S->next, (size_t)S->nnext, sizeof(int), cmp_int);
    int u = 0;
    for (int a = 0; a < S->nnext; a++)
        if (u == 0 || S->next[a] != S->next[u - 1]) S->next[u++] = S->next[a];
    memcpy(S->list, S->next, sizeof(int) * (size_t)u);
    S->nlist = u;
    S->nnext = 0;
}

/* add_to_next_list with the orthogonal rule (extrap_helper.hh:166-184,246-255) */
static inline void add_to_next_list(orc_sim *S, int ind, int p, int q, int r)
{
    int step = p + S->m4 * q + S->mn4 * r;
    int cand = ind + step;
    int diff = abs(step);
    if ((diff == 1 || diff == S->m4 || diff == S->mn4) && S->mask[cand] < S->cur_mask) {
        S->next[S->nnext++] = cand;
        S->mask[cand] = S->nmask;
    }
}

/* extrap_helper::scan_ghost_region (extrap_helper.cc:304-327): the 26 ghost
 * regions together are every cell with a coordinate in a ghost layer */
static void scan_ghost_region(orc_sim *S)
{
    for (int k = 0; k < S->o4; k++)
        for (int j = 0; j < S->n4; j++)
            for (int i = 0; i < S->m4; i++) {
                int g = (i < 2 || i > S->m + 1 || j < 2 || j > S->n + 1 || k < 2 || k > S->o + 1);
                if (!g) continue;
                int ind = i + S->m4 * (j + S->n4 * k);
                if (S->mask[ind] == S->nmask) S->next[S->nnext++] = ind;
            }
}

/* mat3::inverse (mat3.hh:61-67) */
void orc_mat3_inverse(const double a[9], double out[9], double *dt)
{
    double a11 = a[0], a12 = a[1], a13 = a[2], a21 = a[3], a22 = a[4], a23 = a[5],
           a31 = a[6], a32 = a[7], a33 = a[8];
    *dt = a11 * (a22 * a33 - a23 * a32) + a21 * (a13 * a32 - a12 * a33) + a31 * (a12 * a23 - a13 * a22);
    double idet = 1 / (*dt);
    out[0] = (a22 * a33 - a23 * a32) * idet;
    out[1] = (a13 * a32 - a12 * a33) * idet;
    out[2] = (a12 * a23 - a13 * a22) * idet;
    out[3] = (a23 * a31 - a21 * a33) * idet;
    out[4] = (a11 * a33 - a13 * a31) * idet;
    out[5] = (a13 * a21 - a11 * a23) * idet;
    out[6] = (a21 * a32 - a22 * a31) * idet;
    out[7] = (a12 * a31 - a11 * a32) * idet;
    out[8] = (a11 * a22 - a12 * a21) * idet;
}

/* extrap_helper::least_square + lin_sys + dot (extrap_helper.cc:34-169) for
 * N = 3.  Returns 1 on success; fails when there is no point or
 * |det| < 1e-17. */
int orc_least_square_fit(int npt, const double *pos, const double *val,
                         const double *w, double wsum, double davg[3],
                         double ravg[3], double coeff[9], double *det)
{
    if (npt == 0) return 0;
    for (int k = 0; k < 3; k++) { ravg[k] = 0; davg[k] = 0; }
    double a[6] = {0, 0, 0, 0, 0, 0}, rhs[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int p = 0; p < npt; p++) {
        double Q = w[p];
        for (int i = 0; i < 3; i++) ravg[i] += val[3 * p + i] * Q;
        for (int i = 0; i < 3; i++) davg[i] += pos[3 * p + i] * Q;
    }
    for (int k = 0; k < 3; k++) ravg[k] /= wsum;
    for (int k = 0; k < 3; k++) davg[k] /= wsum;
    for (int p = 0; p < npt; p++) {
        const double *dp = pos + 3 * p, *rp = val + 3 * p;
        double Q = w[p];
        for (int i = 0; i < 3; i++) a[i] += Q * (dp[i] - davg[i]) * (dp[i] - davg[i]);
        a[3] += Q * (dp[0] - davg[0]) * (dp[1] - davg[1]);
        a[4] += Q * (dp[0] - davg[0]) * (dp[2] - davg[2]);
        a[5] += Q * (dp[1] - davg[1]) * (dp[2] - davg[2]);
        for (int i = 0; i < 3; i++) {
            double lr = Q * (rp[i] - ravg[i]);
            for (int j = 0; j < 3; j++) rhs[3 * i + j] += (dp[j] - davg[j]) * lr;
        }
    }
    double A[9] = {a[0], a[3], a[4], a[3], a[1], a[5], a[4], a[5], a[2]}, Ai[9];
    orc_mat3_inverse(A, Ai, det);
    if (fabs(*det) < 1e-17) return 0;
    for (int i = 0; i < 3; i++) {
        int x = 3 * i, y = x + 1, z = x + 2;
        coeff[x] = Ai[0] * rhs[x] + Ai[1] * rhs[y] + Ai[2] * rhs[z];
        coeff[y] = Ai[3] * rhs[x] + Ai[4] * rhs[y] + Ai[5] * rhs[z];
        coeff[z] = Ai[6] * rhs[x] + Ai[7] * rhs[y] + Ai[8] * rhs[z];
    }
    return 1;
}

/* scratch of one object's maps: lid (-1 none, 0 primary, L layer), value */
typedef struct {
    signed char *lid;
    double *val;       /* 3 x Ng interleaved per cell */
} obj_scratch;

/* averaged unit normal from the previous step's face normals
 * (fluid_3d_ex.cc:101-167) */
static void prev_normal(orc_sim *S, int ind, unsigned oid, double gp[3], double *gnorm)
{
    int ii, jj, kk;
    unpack3(S, ind, &ii, &jj, &kk);
    gp[0] = gp[1] = gp[2] = 0;
    *gnorm = 0;
    int nf = 0;
    if (!(ii >= 2 && ii <= S->m4 - 3 && jj >= 2 && jj <= S->n4 - 3 && kk >= 2 && kk <= S->o4 - 3))
        return;
    orc_rmap *prev[4];
    prev[0] = orc_get_refmap_prev(S, ind, oid);
    prev[1] = orc_get_refmap_prev(S, ind + 1, oid);
    prev[2] = orc_get_refmap_prev(S, ind + S->m4, oid);
    prev[3] = orc_get_refmap_prev(S, ind + S->mn4, oid);
    for (int pp = 0; pp < 4; pp++) {
        if (!prev[pp]) continue;
        if (pp == 0) {
            for (int dd = 0; dd < 3; dd++) {
                if (prev[0]->gphi[dd][dd] != -1.e200) {
                    int add = 0;
                    for (int ff = 0; ff < 3; ff++) {
                        if (prev[0]->gphi[dd][ff] != 0.0) add = 1;
                        gp[ff] += prev[0]->gphi[dd][ff];
                    }
                    if (add) nf++;
                }
            }
        } else if (prev[pp]->gphi[pp - 1][pp - 1] != -1.e200) {
            int add = 0;
            for (int ff = 0; ff < 3; ff++) {
                if (prev[pp]->gphi[pp - 1][ff] != 0.0) add = 1;
                gp[ff] += prev[pp]->gphi[pp - 1][ff];
            }
            if (add) nf++;
        }
    }
    if (nf == 6) {
        double nrm = 0;
        for (int pp = 0; pp < 3; pp++) { gp[pp] /= nf; nrm += gp[pp] * gp[pp]; }
        if (nrm != 0) {
            nrm = sqrt(nrm);
            for (int pp = 0; pp < 3; pp++) gp[pp] /= nrm;
            *gnorm = nrm;
        } else {
            gp[0] = gp[1] = gp[2] = 0;
        }
    } else {
        gp[0] = gp[1] = gp[2] = 0;
    }
}

/* gather_and_list + least_square + transfer for one cell
 * (fluid_3d_ex.cc:94-271, 293-328).  Returns 1 and the value on success,
 * 0 if the value is NaN (no entry), -1 on extrapolation failure (fatal). */
static int extrap_cell(orc_sim *S, obj_scratch *sc, int ind, unsigned oid, int L,
                       double out[3], double *wpow, double *pos, double *val, double *w)
{
    int ii, jj, kk;
    unpack3(S, ind, &ii, &jj, &kk);
    double gp[3], gnorm;
    prev_normal(S, ind, oid, gp, &gnorm);
    int sr = 2, ok = 0, npt = 0;
    double wsum = 0, davg[3], ravg[3], coeff[9], det;
    while (!ok && sr <= S->cfg.max_extrap_rs) {
        npt = 0; wsum = 0;
        for (int r = -sr; r <= sr; r++)
            for (int q = -sr; q <= sr; q++)
                for (int p = -sr; p <= sr; p++) {
                    if (ii + p > S->m4 - 1 || ii + p < 0) continue;
                    if (jj + q > S->n4 - 1 || jj + q < 0) continue;
                    if (kk + r > S->o4 - 1 || kk + r < 0) continue;
                    int nb = ind + p + S->m4 * q + S->mn4 * r;
                    int lid = sc->lid[nb];
                    if (lid < 0 || lid >= L) continue;
                    double lc[3] = {S->lx[ii + p], S->ly[jj + q], S->lz[kk + r]};
                    double wt;
                    int dist = abs(p) + abs(q) + abs(r);
                    if (gnorm != 0. && L <= 2) {
                        double ev[3] = {S->lx[ii] - lc[0], S->ly[jj] - lc[1], S->lz[kk] - lc[2]};
                        double dw = 0., en = 0.;
                        for (int a = 0; a < 3; a++) { dw += ev[a] * gp[a]; en += ev[a] * ev[a]; }
                        en = sqrt(en);
                        dw /= en;
                        dw *= wpow[dist];
                        if (!(dw > 0)) continue;
                        wt = dw;
                    } else {
                        wt = wpow[dist];
                    }
                    pos[3 * npt] = lc[0]; pos[3 * npt + 1] = lc[1]; pos[3 * npt + 2] = lc[2];
                    val[3 * npt] = sc->val[3 * nb]; val[3 * npt + 1] = sc->val[3 * nb + 1];
                    val[3 * npt + 2] = sc->val[3 * nb + 2];
                    w[npt] = wt;
                    wsum += wt;
                    npt++;
                }
        ok = orc_least_square_fit(npt, pos, val, w, wsum, davg, ravg, coeff, &det);
        sr++;
    }
    if (!ok) return -1;
    double ps[3] = {S->lx[ii] - davg[0], S->ly[jj] - davg[1], S->lz[kk] - davg[2]};
    for (int a = 0; a < 3; a++) {
        out[a] = ravg[a];
        for (int b = 0; b < 3; b++) out[a] += coeff[3 * a + b] * ps[b];
    }
    if (isnan(out[0]) || isnan(out[1]) || isnan(out[2])) return 0;
    return 1;
}

/* fluid_3d::extrapolate (fluid_3d_ex.cc:31-47) with first_layer,
 * extrap_a_layer, gather_and_list and communicate_a_layer.  The band is
 * assumed reset by the caller.  Returns 0, or -1 on extrapolation failure. */
static int extrapolate_impl(orc_sim *S)
{
    int Ng = S->Ng;
    int nstage = 0, capstage = 1024;
    orc_bentry *stage = (orc_bentry *)malloc(sizeof(orc_bentry) * (size_t)capstage);
    obj_scratch sc;
    sc.lid = (signed char *)malloc((size_t)Ng);
    sc.val = (double *)malloc(sizeof(double) * 3 * (size_t)Ng);
    int rmax = S->cfg.max_extrap_rs;
    int maxpt = (2 * rmax + 1) * (2 * rmax + 1) * (2 * rmax + 1);
    double *pos = (double *)malloc(sizeof(double) * 3 * (size_t)maxpt);
    double *val = (double *)malloc(sizeof(double) * 3 * (size_t)maxpt);
    double *w = (double *)malloc(sizeof(double) * (size_t)maxpt);
    double wpow[64];
    for (int d = 0; d < 64; d++) wpow[d] = pow(S->cfg.weight_fac, d);
    int err = 0;

    for (int oid = 1; oid <= S->n_obj && !err; oid++) {
        /* scratch: primary maps of this object (get_refmap: c == oid) */
        for (int a = 0; a < Ng; a++) {
            if (S->rm[a].c == (uint32_t)oid) {
                sc.lid[a] = 0;
                memcpy(sc.val + 3 * a, S->rm[a].xpred, 3 * sizeof(double));
            } else {
                sc.lid[a] = -1;
            }
        }
        /* first_layer (fluid_3d_ex.cc:52-70) -> reset, mark_primary,
         * add_first_layer (extrap_helper.hh:214-227, .cc:207-225) */
        S->n_layers = 0;
        if (S->nmask > UINT_MAX - 50) {
            S->nmask = 1;
            memset(S->mask, 0, sizeof(uint32_t) * (size_t)Ng);
        }
        S->cur_mask = S->nmask;
        for (int a = 0; a < Ng; a++)
            if (S->rm[a].c == (uint32_t)oid) S->mask[a] = S->nmask;
        S->nmask++;
        S->nnext = 0;
        for (int k = 1; k < S->o4 - 1; k++)
            for (int j = 1; j < S->n4 - 1; j++)
                for (int i = 1; i < S->m4 - 1; i++) {
                    int ind = i + S->m4 * (j + S->n4 * k);
                    if (S->mask[ind] == S->cur_mask)
                        for (int kk = -1; kk < 2; kk++)
                            for (int jj = -1; jj < 2; jj++)
                                for (int ii = -1; ii < 2; ii++) add_to_next_list(S, ind, ii, jj, kk);
                }
        orc_comm_mask(S);
        scan_ghost_region(S);
        update_list(S);

        for (int L = 1; L <= S->cfg.nlayers; L++) {
            S->n_layers = L;
            S->nmask++;  /* increase_nmask */
            for (int a = 0; a < S->nlist; a++) {
                int ind = S->list[a];
                if (within_extrap_domain(S, ind)) {
                    double v[3];
                    int rc = extrap_cell(S, &sc, ind, (unsigned)oid, L, v, wpow, pos, val, w);
                    if (rc < 0) {
                        int ii, jj, kk;
                        unpack3(S, ind, &ii, &jj, &kk);
                        fprintf(stderr, "oracle: extrapolation failed at (%d %d %d) layer %d\n",
                                ii - 2, jj - 2, kk - 2, L);
                        err = -1;
                        break;
                    }
                    if (rc == 1) {
                        /* deferred: the value becomes visible only to later
                         * layers (lid < n_layers in gather_and_list) */
                        if (nstage == capstage) {
                            capstage *= 2;
                            stage = (orc_bentry *)realloc(stage, sizeof(orc_bentry) * (size_t)capstage);
                        }
                        orc_bentry *e = &stage[nstage++];
                        memset(e, 0, sizeof *e);
                        e->cell = ind;
                        e->r.c = (uint32_t)oid | ((uint32_t)L << LDIG);
                        for (int b = 0; b < 3; b++) e->r.x[b] = e->r.xpred[b] = v[b];
                        for (int f = 0; f < 3; f++) {
                            for (int b = 0; b < 3; b++) e->r.gphi[f][b] = -1.e200;
                            e->r.detF[f] = -1.e200;
                            e->r.energy[f] = -1.e200;
                        }
                    }
                }
                if (L < S->cfg.nlayers && within_inner_ghost(S, ind))
                    for (int r = -1; r < 2; r++)
                        for (int q = -1; q < 2; q++)
                            for (int p = -1; p < 2; p++) add_to_next_list(S, ind, p, q, r);
            }
            if (err) break;
            /* make this layer's values visible to the next layer */
            for (int a = nstage - 1; a >= 0; a--) {
                orc_bentry *e = &stage[a];
                if ((e->r.c & OMASK) != (uint32_t)oid || (int)(e->r.c >> LDIG) != L) break;
                sc.lid[e->cell] = (signed char)L;
                memcpy(sc.val + 3 * e->cell, e->r.xpred, 3 * sizeof(double));
            }
            if (L < S->cfg.nlayers) {
                orc_comm_mask(S);
                scan_ghost_region(S);
            }
            update_list(S);
            /* communicate_a_layer(c = L<<10 | oid, xpred) on one rank: a periodic
             * ghost takes its source's entry of this layer (fluid_3d_comm.cc:203-280,
             * extrap_mmap.hh:79-146 with type 1) */
            if (S->prd[0] || S->prd[1] || S->prd[2]) {
                int base = nstage;
                for (int k = 0; k < S->o4; k++)
                    for (int j = 0; j < S->n4; j++)
                        for (int i = 0; i < S->m4; i++) {
                            int src;
                            if (orc_periodic_src(S, i, j, k, 0, &src)) continue;
                            if (sc.lid[src] != L) continue;
                            int dst = i + S->m4 * (j + S->n4 * k);
                            if (nstage == capstage) {
                                capstage *= 2;
                                stage = (orc_bentry *)realloc(stage, sizeof(orc_bentry) * (size_t)capstage);
                            }
                            orc_bentry *e = &stage[nstage++];
                            memset(e, 0, sizeof *e);
                            e->cell = dst;
                            e->r.c = (uint32_t)oid | ((uint32_t)L << LDIG);
                            for (int b = 0; b < 3; b++) e->r.x[b] = e->r.xpred[b] = sc.val[3 * src + b];
                            for (int f = 0; f < 3; f++) {
                                for (int b = 0; b < 3; b++) e->r.gphi[f][b] = -1.e200;
                                e->r.detF[f] = -1.e200;
                                e->r.energy[f] = -1.e200;
                            }
                        }
                for (int a = base; a < nstage; a++) {
                    sc.lid[stage[a].cell] = (signed char)L;
                    memcpy(sc.val + 3 * stage[a].cell, stage[a].r.xpred, 3 * sizeof(double));
                }
            }
        }
    }
    if (!err) band_build(&S->ext, stage, nstage, Ng);
    free(stage); free(sc.lid); free(sc.val); free(pos); free(val); free(w);
    return err;
}

void orc_extrapolate(orc_sim *S)
{
    if (extrapolate_impl(S) != 0) {
        fprintf(stderr, "oracle: Extrapolation failed. Oops\n");
        abort();
    }
}

/* fluid_3d::redraw_boundary (fluid_3d_ex.cc:336-394).  The loop over the
 * N entries read before the loop keeps the reference's index semantics: a
 * removal shifts later entries down, so the entry that moves into position l
 * is skipped; indices past the shortened list are skipped (the reference
 * reads a stale slot there). */
void orc_redraw_boundary(orc_sim *S)
{
    orc_band *B = &S->ext;
    int Ng = S->Ng;
    uint32_t lflag = (uint32_t)(S->cfg.nlayers + 1) << LDIG;
    int cap = B->n + 1024;
    orc_bentry *out = (orc_bentry *)malloc(sizeof(orc_bentry) * (size_t)cap);
    int nout = 0;
    orc_bentry loc[64];
    for (int ind = 0; ind < Ng; ind++) {
        int nl = 0;
        for (int a = 0; a < B->cnt[ind] && nl < 63; a++) loc[nl++] = B->e[B->start[ind] + a];
        int occupied = 0;
        orc_rmap *prime = S->rm + ind;
        int oidp = (int)(prime->c & OMASK);
        if (oidp > 0) {
            double pphi = orc_phi(S, oidp - 1, prime->x);
            if (pphi > 0) {
                prime->c = lflag | prime->c;
                loc[nl].cell = ind;
                loc[nl].r = *prime;
                nl++;
                memset(prime, 0, sizeof *prime);
                S->n_redraw_out++;
            } else {
                occupied = 1;
            }
        }
        int N = nl;
        for (int l = 0; l < N; l++) {
            if (l >= nl) continue;
            orc_rmap *sf = &loc[l].r;
            unsigned oid = sf->c & OMASK;
            double sphi = orc_phi(S, (int)oid - 1, sf->x);
            int inside = (sphi < 0 && (sf->c >> LDIG) == 1);
            if (!inside) continue;
            int po = (int)(prime->c & OMASK);
            double pphi = po > 0 ? orc_phi(S, po - 1, prime->x) : NAN;
            if ((occupied && sphi < pphi) || !occupied) {
                if (pphi < -S->dx) {
                    fprintf(stderr, "oracle: redraw_boundary(): Two deep maps try to be inside a cell at the same time!\n");
                    abort();
                }
                *prime = *sf;
                prime->c = oid;
                occupied = 1;
                S->n_redraw_in++;
                /* extrap_mmap::rm_mem (extrap_mmap.cc:207-257): first match */
                int j0 = -1;
                for (int b = 0; b < nl; b++)
                    if ((loc[b].r.c & OMASK) == oid && j0 == -1) j0 = b;
                if (j0 >= 0) {
                    for (int b = j0; b < nl - 1; b++) loc[b] = loc[b + 1];
                    nl--;
                }
            }
        }
        if (nout + nl > cap) {
            cap = 2 * (nout + nl);
            out = (orc_bentry *)realloc(out, sizeof(orc_bentry) * (size_t)cap);
        }
        for (int a = 0; a < nl; a++) out[nout++] = loc[a];
    }
    band_build(B, out, nout, Ng);
    free(out);
}
