/*
 * ORACLE -- test infrastructure only.
 *
 * Tracer particles restated on the oracle's state (Utilities/tracers.{hh,cc}):
 * the constructor's random placement (tracers.cc:7-66) with GSL's taus
 * generator, tracers::set, predict / imp_euler / advect, dx and proc_adjust
 * (tracers.hh:160-442, tracers.cc:259-352) on one rank, where the hand-off
 * to a neighbour is a periodic wrap or, without one, the loss of the tracer.
 * The velocity interpolant is the tricubic polynomial through the 4x4x4
 * cube of cell velocities, evaluated as a tensor product of cubic Lagrange
 * weights (the reference builds the same polynomial from a stored inverse
 * Vandermonde matrix, so the two agree to rounding; GSL itself is absent, so
 * the taus seeding is restated from its published description and its bit
 * stream is "parity unpinned").
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "orc.h"
#include "orc_detmath.h"

/* L'Ecuyer (1996) three-component Tausworthe generator with GSL's LCG
 * seeding and six warm-up draws (gsl_rng_taus) */
typedef struct { uint32_t s1, s2, s3; } taus_t;

static uint32_t taus_next(taus_t *t)
{
    t->s1 = ((t->s1 & 4294967294u) << 12) ^ (((t->s1 << 13) ^ t->s1) >> 19);
    t->s2 = ((t->s2 & 4294967288u) << 4) ^ (((t->s2 << 2) ^ t->s2) >> 25);
    t->s3 = ((t->s3 & 4294967280u) << 17) ^ (((t->s3 << 3) ^ t->s3) >> 11);
    return t->s1 ^ t->s2 ^ t->s3;
}

static void taus_seed(taus_t *t, uint32_t s)
{
    if (s == 0) s = 1;
    t->s1 = 69069u * s;
    if (t->s1 < 2) t->s1 += 2;
    t->s2 = 69069u * t->s1;
    if (t->s2 < 8) t->s2 += 8;
    t->s3 = 69069u * t->s2;
    if (t->s3 < 16) t->s3 += 16;
    for (int i = 0; i < 6; i++) taus_next(t);
}

uint32_t orc_taus_stream(uint32_t seed, int n, uint32_t *out)
{
    taus_t t;
    taus_seed(&t, seed);
    for (int i = 0; i < n; i++) out[i] = taus_next(&t);
    return n;
}

static double tr_h(const orc_sim *S, int f) { return f == 0 ? S->dx : (f == 1 ? S->dy : S->dz); }
static double tr_low(const orc_sim *S, int f) { return f == 0 ? S->lx[0] : (f == 1 ? S->ly[0] : S->lz[0]); }
static int tr_lim(const orc_sim *S, int f) { return f == 0 ? S->m4 : (f == 1 ? S->n4 : S->o4); }
static int tr_cells(const orc_sim *S, int f) { return f == 0 ? S->m : (f == 1 ? S->n : S->o); }

static int too_low(const orc_tracer *t, int f, int st)
{
    return (t->p_crd[f] < 0) || (t->p_crd[f] == 0 && ((st == 0) ? t->rpf[f] : t->rp0[f]) < 0);
}
static int too_high(const orc_sim *S, const orc_tracer *t, int f, int st)
{
    int l = tr_lim(S, f) - 4;
    return (t->p_crd[f] > l) || (t->p_crd[f] == l && ((st == 0) ? t->rpf[f] : t->rp0[f]) >= 0);
}

/* tracers::set (tracers.hh:160-172) */
static int tr_set(const orc_sim *S, orc_tracer *t, const double r[3])
{
    for (int f = 0; f < 3; f++) {
        double t_ind = (r[f] - tr_low(S, f)) / tr_h(S, f) - 1;
        t->p_crd[f] = (int)t_ind;
        t->rp0[f] = t_ind - t->p_crd[f] - 0.5;
        if (too_low(t, f, 1) || too_high(S, t, f, 1)) return 0;
    }
    return 1;
}

int orc_tracers_init(orc_sim *S)
{
    int n = S->cfg.ntracers;
    free(S->tr);
    S->tr = (orc_tracer *)calloc((size_t)(n > 0 ? n : 1), sizeof(orc_tracer));
    S->ntr = n;
    if (S->cfg.chk_num <= 0) {
        taus_t rng;
        taus_seed(&rng, 1);
        double h[3] = {S->dx, S->dy, S->dz};
        const double *b = S->cfg.xyz_bounds;
        double ax = (b[0] + 0.5 * h[0]) - 0.5 * h[0];
        double ay = (b[2] + 0.5 * h[1]) - 0.5 * h[1];
        double az = (b[4] + 0.5 * h[2]) - 0.5 * h[2];
        double Lx = S->m * h[0], Ly = S->n * h[1], Lz = S->o * h[2];
        for (int i = 0; i < n; i++) {
            double r[3];
            r[0] = ax + Lx * (taus_next(&rng) / 4294967296.0);
            r[1] = ay + Ly * (taus_next(&rng) / 4294967296.0);
            r[2] = az + Lz * (taus_next(&rng) / 4294967296.0);
            S->tr[i].alive = tr_set(S, &S->tr[i], r);
        }
    } else {
        /* tracers::init_from_chkpt_files (tracers.cc:142-172) */
        char fn[600];
        snprintf(fn, sizeof fn, "%s/tr.%05d", S->cfg.chk_dirname, S->cfg.chk_num);
        FILE *fh = fopen(fn, "rb");
        if (!fh) return -1;
        int lnum = 0;
        float *tout = (float *)malloc(sizeof(float) * 3 * (size_t)(n > 0 ? n : 1));
        size_t r1 = fread(&lnum, sizeof(int), 1, fh);
        size_t r2 = (r1 == 1 && lnum == n) ? fread(tout, sizeof(float), 3 * (size_t)n, fh) : 0;
        fclose(fh);
        if (r1 != 1 || lnum != n || r2 != 3 * (size_t)n) { free(tout); return -1; }
        for (int i = 0; i < n; i++) {
            double r[3] = {tout[3 * i], tout[3 * i + 1], tout[3 * i + 2]};
            S->tr[i].alive = tr_set(S, &S->tr[i], r);
        }
        free(tout);
    }
    for (int i = 0; i < n; i++) memcpy(S->tr[i].rpf, S->tr[i].rp0, sizeof(double) * 3);
    S->trace = 1;
    return 0;
}

static void lagrange4(double r, double w[4])
{
    w[0] = -(r + 0.5) * (r - 0.5) * (r - 1.5) / 6.0;
    w[1] = (r + 1.5) * (r - 0.5) * (r - 1.5) / 2.0;
    w[2] = -(r + 1.5) * (r + 0.5) * (r - 1.5) / 2.0;
    w[3] = (r + 1.5) * (r + 0.5) * (r - 0.5) / 6.0;
}

/* tracers::interp_vels (tracers.cc:395-455) */
static void interp_vels(const orc_sim *S, const orc_tracer *t, const double r[3], double v[3])
{
    int ti = t->p_crd[0], tj = t->p_crd[1], tk = t->p_crd[2];
    if (ti < 0) ti = 0;
    if (tj < 0) tj = 0;
    if (tk < 0) tk = 0;
    if (ti > S->m4 - 4) ti = S->m4 - 4;
    if (tj > S->n4 - 4) tj = S->n4 - 4;
    if (tk > S->o4 - 4) tk = S->o4 - 4;
    double wx[4], wy[4], wz[4];
    lagrange4(r[0], wx);
    lagrange4(r[1], wy);
    lagrange4(r[2], wz);
    for (int f = 0; f < 3; f++) {
        double sk = 0;
        for (int k = 0; k < 4; k++) {
            double sj = 0;
            for (int j = 0; j < 4; j++) {
                const orc_field *row = S->u + ti + S->m4 * ((tj + j) + S->n4 * (tk + k));
                double si = 0;
                for (int i = 0; i < 4; i++) si += wx[i] * row[i].vel[f];
                sj += wy[j] * si;
            }
            sk += wz[k] * sj;
        }
        v[f] = sk;
    }
}

/* tracers::dx (tracers.hh:331-442) */
static double tr_dx(const orc_sim *S, const orc_tracer *t, int f, int st)
{
    double v = st == 0 ? t->v0[f] : t->vf[f];
    double update = S->dt * v;
    int lr = (v < 0) ? 0 : 1;
    int bct = S->cfg.bct[f][2 * f + lr];
    double bcv = S->cfg.bcv[f][2 * f + lr];
    if (!(bct == 1 && bcv == 0)) return update;
    double h = tr_h(S, f);
    double lo = (S->cfg.xyz_bounds[2 * f] + 0.5 * h) - 0.5 * h;
    double hi = lo + tr_cells(S, f) * h;
    double xx = tr_low(S, f) + (1.5 + (st == 0 ? t->rp0[f] : t->rpf[f]) + t->p_crd[f]) * h;
    const double min_sep = 0.01 * (hi - lo);
    double sep, vn;
    if (v == 0) {
        double v1 = fabs(xx - lo), v2 = fabs(hi - xx);
        sep = v2 < v1 ? (hi - xx) : (xx - lo);
    } else
        sep = (v < 0) ? (xx - lo) : (hi - xx);
    vn = fabs(v);
    if (sep <= min_sep) {
        if (v == 0) {
            double v1 = fabs(xx - lo), v2 = fabs(hi - xx);
            return (v2 < v1) ? (hi - min_sep - xx) : (lo + min_sep - xx);
        }
        return (v < 0) ? (lo + min_sep - xx) : (hi - min_sep - xx);
    }
    if (fabs(update) > 0.1 * sep) update = (det_exp(-S->dt * vn / sep) - 1) * ((v < 0) ? sep : -sep);
    if ((v < 0 && (xx + update) < (lo + min_sep)) || (v > 0 && (xx + update) > (hi - min_sep)))
        return (v < 0) ? (lo + min_sep - xx) : (hi - min_sep - xx);
    return update;
}

/* tracers::proc_adjust + the one-rank communicate / empty_buffer
 * (tracers.cc:220-352): 0 when the tracer is lost */
static int proc_adjust(const orc_sim *S, orc_tracer *t, int st)
{
    int comm[3];
    for (int f = 0; f < 3; f++) {
        double *val = (st == 0) ? &t->rpf[f] : &t->rp0[f];
        while (*val <= -0.5) { t->rp0[f] += 1.0; t->rpf[f] += 1.0; t->p_crd[f]--; }
        while (*val > 0.5) { t->rp0[f] -= 1.0; t->rpf[f] -= 1.0; t->p_crd[f]++; }
        if (too_low(t, f, st)) comm[f] = 0;
        else if (too_high(S, t, f, st)) { comm[f] = 2; t->p_crd[f] -= tr_lim(S, f) - 4; }
        else comm[f] = 1;
    }
    int ind = comm[0] + 3 * comm[1] + 9 * comm[2];
    if (ind == 13) return 1;
    for (int f = 0; f < 3; f++)
        if (comm[f] != 1 && !S->prd[f]) return 0;
    for (int f = 0; f < 3; f++) {
        double val = (st == 0) ? t->rpf[f] : t->rp0[f];
        if (t->p_crd[f] < 0 || (t->p_crd[f] == 0 && val < 0)) t->p_crd[f] += tr_lim(S, f) - 4;
    }
    return 1;
}

void orc_tracers_predict(orc_sim *S)
{
    for (int i = 0; i < S->ntr; i++) {
        orc_tracer *t = &S->tr[i];
        if (!t->alive) continue;
        interp_vels(S, t, t->rp0, t->v0);
        for (int f = 0; f < 3; f++) {
            double dr = tr_dx(S, t, f, 0) / tr_h(S, f);
            t->rpf[f] = t->rp0[f] + dr;
        }
        t->alive = proc_adjust(S, t, 0);
    }
}

void orc_tracers_advect(orc_sim *S)
{
    for (int i = 0; i < S->ntr; i++) {
        orc_tracer *t = &S->tr[i];
        if (!t->alive) continue;
        interp_vels(S, t, t->rpf, t->vf);
        for (int f = 0; f < 3; f++) {
            double dr0 = t->rpf[f] - t->rp0[f];
            double drf = tr_dx(S, t, f, 1) / tr_h(S, f);
            t->rp0[f] += 0.5 * (dr0 + drf);
        }
        t->alive = proc_adjust(S, t, 1);
    }
}

/* tracers::print's single-precision positions by id (tracers.cc:110-127) */
void orc_tracer_positions(const orc_sim *S, float *xyz)
{
    for (int i = 0; i < S->ntr; i++) {
        const orc_tracer *t = &S->tr[i];
        for (int f = 0; f < 3; f++)
            xyz[3 * i + f] = t->alive ? (float)(tr_low(S, f) + (1.5 + t->rp0[f] + t->p_crd[f]) * tr_h(S, f)) : 0.f;
    }
}

/* test hooks: the interpolant at one cube position, and one predict +
 * advect of every tracer on the current velocity field */
void orc_tracer_interp(const orc_sim *S, const int p[3], const double r[3], double v[3])
{
    orc_tracer t;
    memset(&t, 0, sizeof t);
    for (int f = 0; f < 3; f++) t.p_crd[f] = p[f];
    interp_vels(S, &t, r, v);
}
