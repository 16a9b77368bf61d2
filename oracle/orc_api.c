/*
 * ORACLE -- test infrastructure only (see orc.h).
 *
 * Flat entry points loaded by tests/ and bench.py's cpu_baseline leg through
 * ctypes.  State arrays use the rmt3d_state layout of include/rmt3d.h.
 */
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif
#include "orc.h"
#include "orc_detmath.h"

double orc_mono_diff(double u0, double u1, double u2, double u3, double u4);
void orc_godunov_set_velocity(double *nl, double *vl, double *wl, double *nr, double *vr, double *wr);
void orc_godunov_set_refmap(double nl, double nr, double *ml, double *mvl, double *mwl,
                            double *mr, double *mvr, double *mwr);
void orc_mat_inverse(const double v[9], double m2[9]);
void orc_mat_aat(const double F[9], double s[6]);
void orc_mat_ata(const double F[9], double s[6]);
double orc_mat_det(const double v[9]);
void orc_mat3_inverse(const double a[9], double out[9], double *dt);

orc_sim *oracle_create(const rmt3d_config *cfg, const rmt3d_object *objs) { return orc_create(cfg, objs); }
void oracle_destroy(orc_sim *S) { orc_destroy(S); }
int oracle_initialize(orc_sim *S) { return orc_initialize(S); }
int oracle_init_iter(orc_sim *S, int e) { return orc_init_iter(S, e); }
int oracle_prepare_run(orc_sim *S, int e, double *dt, int *tot) { return orc_prepare_run(S, e, dt, tot); }
int oracle_step(orc_sim *S, int debug) { return orc_step(S, debug); }
double oracle_get_dt(orc_sim *S) { return S->dt; }
double oracle_get_time(orc_sim *S) { return S->time; }
void oracle_set_dt(orc_sim *S, double dt) { S->dt = dt; S->cfg.dt = dt; }
void oracle_set_time(orc_sim *S, double t) { S->time = t; }
double oracle_div_u(orc_sim *S) { return orc_div_u(S); }
int oracle_band_count(orc_sim *S) { return S->ext.n; }
void oracle_compute_stress(orc_sim *S) { orc_compute_stress(S); }
void oracle_collision_face(orc_sim *S, int eid, int F, double *fs) { orc_collision_face(S, eid, F, fs); }
void oracle_redraw(orc_sim *S) { S->n_redraw_out = S->n_redraw_in = 0; orc_redraw_boundary(S); }
void oracle_set_bc_cc(orc_sim *S) { orc_set_bc_cc(S); }
void oracle_set_bc_cf(orc_sim *S) { orc_set_bc_cf(S); }
/* One extrapolation pass from the current primary xpred, with the current
 * band kept as the previous one (get_refmap_prev), as update_reference_map
 * runs it (fluid_3d.cc:1182-1190). */
void oracle_extrap_now(orc_sim *S)
{
    orc_band_copy(&S->tmp, &S->ext, S->Ng);
    orc_band_reset(&S->ext, S->Ng);
    orc_extrapolate(S);
}
/* Face normals grad_phi[F][a] (written by compute_stress): which 0 -> the
 * primary maps, out[(3F+a) Ng + cell]; which 1 -> the previous band (tmp),
 * out[(3F+a) n + e] with cell[e], c[e].  Returns the entry count. */
int oracle_gphi(orc_sim *S, int which, double *out, int *cell, uint32_t *c)
{
    if (which == 0) {
        for (int a = 0; a < S->Ng; a++)
            for (int q = 0; q < 9; q++) out[(size_t)q * S->Ng + a] = S->rm[a].gphi[q / 3][q % 3];
        return S->Ng;
    }
    int n = S->tmp.n;
    if (!out) return n;
    for (int e = 0; e < n; e++) {
        cell[e] = S->tmp.e[e].cell;
        c[e] = S->tmp.e[e].r.c;
        for (int q = 0; q < 9; q++) out[(size_t)q * n + e] = S->tmp.e[e].r.gphi[q / 3][q % 3];
    }
    return n;
}

/* scalars: 0 dt_reg, 1 eps, 2 K_stiff, 3 vmax, 4 mac cycles, 5 fem cycles,
 * 6 mac res, 7 fem res, 8 nmask; counters of the last step: 9 collision
 * contributions, 10 primary maps redrawn into the band, 11 band maps redrawn
 * onto the primary grid */
double oracle_scalar(orc_sim *S, int which)
{
    switch (which) {
    case 0: return S->dt_reg;
    case 1: return S->eps;
    case 2: return S->K_stiff;
    case 3: return S->vmax;
    case 4: return S->mac_cycles;
    case 5: return S->fem_cycles;
    case 6: return S->mac_res;
    case 7: return S->fem_res;
    case 8: return (double)S->nmask;
    case 9: return (double)S->n_coll;
    case 10: return (double)S->n_redraw_out;
    case 11: return (double)S->n_redraw_in;
    }
    return 0;
}

int oracle_download(orc_sim *S, rmt3d_state *st)
{
    int Ng = S->Ng;
    for (int a = 0; a < Ng; a++) {
        orc_field *f = S->u + a;
        orc_rmap *r = S->rm + a;
        if (st->u) st->u[a] = f->vel[0];
        if (st->v) st->v[a] = f->vel[1];
        if (st->w) st->w[a] = f->vel[2];
        if (st->p) st->p[a] = f->p;
        if (st->dp) st->dp[a] = f->dp;
        if (st->q) st->q[a] = f->q;
        if (st->dvel) for (int c = 0; c < 3; c++) st->dvel[(size_t)c * Ng + a] = f->dvel[c];
        if (st->fvel) for (int c = 0; c < 18; c++) st->fvel[(size_t)c * Ng + a] = f->fvel[c / 3][c % 3];
        if (st->sigma) for (int c = 0; c < 9; c++) st->sigma[(size_t)c * Ng + a] = f->sigma[c / 3][c % 3];
        if (st->rm_c) st->rm_c[a] = r->c;
        if (st->rm_x) for (int c = 0; c < 3; c++) st->rm_x[(size_t)c * Ng + a] = r->x[c];
        if (st->rm_xpred) for (int c = 0; c < 3; c++) st->rm_xpred[(size_t)c * Ng + a] = r->xpred[c];
    }
    int nb = S->ext.n;
    if (st->band_cell || st->band_c || st->band_x || st->band_xpred) {
        if (st->nband < nb) { st->nband = nb; return -1; }
        for (int a = 0; a < nb; a++) {
            orc_bentry *e = &S->ext.e[a];
            if (st->band_cell) st->band_cell[a] = e->cell;
            if (st->band_c) st->band_c[a] = e->r.c;
            if (st->band_x) for (int c = 0; c < 3; c++) st->band_x[(size_t)c * nb + a] = e->r.x[c];
            if (st->band_xpred) for (int c = 0; c < 3; c++) st->band_xpred[(size_t)c * nb + a] = e->r.xpred[c];
        }
    }
    st->nband = nb;
    return 0;
}

int oracle_upload(orc_sim *S, const rmt3d_state *st)
{
    int Ng = S->Ng;
    for (int a = 0; a < Ng; a++) {
        orc_field *f = S->u + a;
        orc_rmap *r = S->rm + a;
        if (st->u) f->vel[0] = st->u[a];
        if (st->v) f->vel[1] = st->v[a];
        if (st->w) f->vel[2] = st->w[a];
        if (st->p) f->p = st->p[a];
        if (st->dp) f->dp = st->dp[a];
        if (st->q) f->q = st->q[a];
        if (st->dvel) for (int c = 0; c < 3; c++) f->dvel[c] = st->dvel[(size_t)c * Ng + a];
        if (st->fvel) for (int c = 0; c < 18; c++) f->fvel[c / 3][c % 3] = st->fvel[(size_t)c * Ng + a];
        if (st->sigma) for (int c = 0; c < 9; c++) f->sigma[c / 3][c % 3] = st->sigma[(size_t)c * Ng + a];
        if (st->rm_c) r->c = st->rm_c[a];
        if (st->rm_x) for (int c = 0; c < 3; c++) r->x[c] = st->rm_x[(size_t)c * Ng + a];
        if (st->rm_xpred) for (int c = 0; c < 3; c++) r->xpred[c] = st->rm_xpred[(size_t)c * Ng + a];
    }
    return 0;
}

/* Replaces the band by n entries (sorted by cell; entries of one cell in
 * multimap order), for format tests on hand-made states. */
int oracle_set_band(orc_sim *S, int n, const int *cell, const uint32_t *c, const double *x, const double *xpred)
{
    orc_band *B = &S->ext;
    orc_band_reset(B, S->Ng);
    if (n > B->cap) {
        while (B->cap < n) B->cap *= 2;
        B->e = (orc_bentry *)realloc(B->e, sizeof(orc_bentry) * (size_t)B->cap);
    }
    for (int a = 0; a < n; a++) {
        if (cell[a] < 0 || cell[a] >= S->Ng || (a > 0 && cell[a] < cell[a - 1])) return -1;
        orc_bentry *e = &B->e[a];
        memset(e, 0, sizeof *e);
        e->cell = cell[a];
        e->r.c = c[a];
        for (int b = 0; b < 3; b++) { e->r.x[b] = x[(size_t)b * n + a]; e->r.xpred[b] = xpred[(size_t)b * n + a]; }
        for (int f = 0; f < 3; f++)
            for (int b = 0; b < 3; b++) e->r.gphi[f][b] = -1.e200;
        if (B->cnt[cell[a]] == 0) B->start[cell[a]] = a;
        B->cnt[cell[a]]++;
    }
    B->n = n;
    return 0;
}

/* checkpoint and slice formats (orc_io.c) */
void oracle_chk_images(orc_sim *S, double *img) { orc_chk_images(S, img); }
void oracle_chk_restore(orc_sim *S, const double *img) { orc_chk_restore(S, img); }
int oracle_chk_read(const char *dir, int o, size_t L, double *img) { return orc_chk_read(dir, o, L, img); }
int oracle_write_chk_pt(orc_sim *S, const char *dir) { return orc_write_chk_pt(S, dir); }
int oracle_slice(orc_sim *S, int dim, int point, int t, int obj_id, double *g) { return orc_slice(S, dim, point, t, obj_id, g); }
int oracle_get_nt(orc_sim *S) { return S->nt; }
int oracle_chk_num(orc_sim *S) { return S->cfg.chk_num; }

/* Multigrid solve on a stand-alone problem: which 0 MAC, 1 FEM.  Uses the
 * simulation's solver objects (initial guess = x on entry). */
int oracle_mg_solve(orc_sim *S, int which, const double *rhs, double *x, double tol, int *ncyc, double *res)
{
    orc_mg *G = which ? &S->fem : &S->mac;
    orc_mglev *L = &G->lev[0];
    memcpy(L->b, rhs, sizeof(double) * (size_t)L->nn);
    memcpy(L->x, x, sizeof(double) * (size_t)L->nn);
    double save = G->max_norm_tol;
    if (tol > 0) G->max_norm_tol = tol;
    double r = orc_mg_run(G);
    G->max_norm_tol = save;
    memcpy(x, L->x, sizeof(double) * (size_t)L->nn);
    if (ncyc) *ncyc = G->last_cycles;
    if (res) *res = r;
    return 0;
}

/* y = A x on the finest level of a problem */
int oracle_mg_apply(orc_sim *S, int which, const double *x, double *y)
{
    orc_mg *G = which ? &S->fem : &S->mac;
    orc_mg_apply(G, 0, x, y);
    return 0;
}

int oracle_mg_levels(orc_sim *S, int which, int *dims)
{
    orc_mg *G = which ? &S->fem : &S->mac;
    for (int l = 0; l < G->nlev; l++)
        for (int a = 0; a < 3; a++) dims[3 * l + a] = G->lev[l].d[a];
    return G->nlev;
}

/* unit-level entry points for known-answer tests */
double oracle_mono_diff(double u0, double u1, double u2, double u3, double u4)
{
    return orc_mono_diff(u0, u1, u2, u3, u4);
}
void oracle_godunov_vel(double *v6) { orc_godunov_set_velocity(v6, v6 + 1, v6 + 2, v6 + 3, v6 + 4, v6 + 5); }
void oracle_godunov_refmap(double nl, double nr, double *m6)
{
    orc_godunov_set_refmap(nl, nr, m6, m6 + 1, m6 + 2, m6 + 3, m6 + 4, m6 + 5);
}
void oracle_mat_inverse(const double *v, double *o) { orc_mat_inverse(v, o); }
void oracle_mat_aat(const double *v, double *o) { orc_mat_aat(v, o); }
void oracle_mat_ata(const double *v, double *o) { orc_mat_ata(v, o); }
double oracle_mat_det(const double *v) { return orc_mat_det(v); }
void oracle_mat3_inverse(const double *v, double *o, double *det) { orc_mat3_inverse(v, o, det); }
double oracle_heaviside(orc_sim *S, double phi) { return orc_heaviside(S, phi); }
int oracle_lsq(int npt, const double *pos, const double *val, const double *w, double wsum,
               double *davg, double *ravg, double *coeff, double *det)
{
    return orc_least_square_fit(npt, pos, val, w, wsum, davg, ravg, coeff, det);
}

void oracle_mg_set_coarse_sweeps(orc_sim *S, int n) { S->mac.coarse_sweeps = n; S->fem.coarse_sweeps = n; }
void oracle_sym_pinv(int n, const double *A, double *P) { orc_sym_pinv(n, A, P); }
void orc_level_apply(int fem, const int d[3], const int prd[3], const double *x, double *y);
void orc_fdm_solve_level(int fem, const int d[3], const int prd[3], const double *b, double *x);
/* y = A x and the fast-diagonalisation solve on one level of given extents */
void oracle_level_apply(int fem, const int *d, const int *prd, const double *x, double *y)
{
    orc_level_apply(fem, d, prd, x, y);
}
void oracle_fdm_solve(int fem, const int *d, const int *prd, const double *b, double *x)
{
    orc_fdm_solve_level(fem, d, prd, b, x);
}
/* 1 when the coarsest level of problem `which` is solved by fast
 * diagonalisation */
int oracle_mg_fdm(orc_sim *S, int which) { return (which ? S->fem.fdm : S->mac.fdm); }
/* 1 when the coarsest level of problem `which` (0 MAC, 1 FEM) is solved
 * exactly by its pseudo-inverse */
int oracle_mg_direct(orc_sim *S, int which) { return (which ? S->fem.pinv : S->mac.pinv) != NULL; }

/* the deterministic sine/cosine shared with the HIP path */
double oracle_det_sin(double x) { return det_sin(x); }
double oracle_det_cos(double x) { return det_cos(x); }
double oracle_det_log(double x) { return det_log(x); }
double oracle_det_exp(double x) { return det_exp(x); }
void oracle_macro(orc_sim *S, double *out) { orc_macro(S, out); }
int oracle_tracer_count(orc_sim *S) { return S->trace ? S->ntr : 0; }
int orc_mc_table(int key, signed char *tri);
int orc_contour(orc_sim *S, int cnum, float **pts, float **nrm, int **tri, int *ntri);
int oracle_mc_table(int key, signed char *tri) { return orc_mc_table(key, tri); }
/* one object's contour into caller arrays of capacity cap vertices / tcap
 * triangles; returns the vertex count (or -1 when it does not fit) */
int oracle_contour(orc_sim *S, int cnum, float *pts, float *nrm, int cap, int *tri, int tcap, int *ntri)
{
    float *p, *n;
    int *t, nt;
    int nv = orc_contour(S, cnum, &p, &n, &t, &nt);
    *ntri = nt;
    int ok = nv <= cap && nt <= tcap;
    if (ok) {
        memcpy(pts, p, sizeof(float) * 3 * (size_t)nv);
        memcpy(nrm, n, sizeof(float) * 3 * (size_t)nv);
        memcpy(tri, t, sizeof(int) * 3 * (size_t)nt);
    }
    free(p); free(n); free(t);
    return ok ? nv : -1;
}
void orc_tracer_interp(const orc_sim *S, const int p[3], const double r[3], double v[3]);
void oracle_tracer_interp(orc_sim *S, const int *p, const double *r, double *v) { orc_tracer_interp(S, p, r, v); }
int oracle_tracers_init(orc_sim *S) { return orc_tracers_init(S); }
void oracle_tracers_move(orc_sim *S) { orc_tracers_predict(S); orc_tracers_advect(S); }
void oracle_tracer_positions(orc_sim *S, float *xyz) { orc_tracer_positions(S, xyz); }
uint32_t orc_taus_stream(uint32_t seed, int n, uint32_t *out);
void oracle_taus(uint32_t seed, int n, uint32_t *out) { orc_taus_stream(seed, n, out); }
void oracle_centroids(orc_sim *S, double *out) { orc_centroids(S, out); }

/* threads the oracle's parallel loops use (OpenMP; 1 without it) */
int oracle_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
