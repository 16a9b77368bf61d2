/*
 * ORACLE -- test infrastructure only.
 *
 * The macro.dat / centroid.dat reductions of the stepper
 * (Sim/stepper.hh:378-475) restated on the oracle's state:
 *   fluid_3d::div_u (MAC extremes), total_momentum, total_energy,
 *   total_dev_in_detF, compute_v_centroid + sim_manager::compute_avgs
 *   (fluid_3d_utils.cc:507-1346, sim_manager.hh:223-246).
 * Each sum is formed as the reference forms its per-cell contribution, then
 * added over the interior cells in i-fastest order with diag_sum (the fixed
 * blocked order the HIP path uses), not serially.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "orc.h"

static inline int IXd(const orc_sim *S, int i, int j, int k) { return i + S->m4 * (j + S->n4 * k); }

/* The fixed summation order of the HIP path's diagnostics: 2048 blocks of
 * ceil(n/2048) consecutive terms, 256 strided partial sums per block and a
 * binary tree; then 256 sums of 8 consecutive block partials, and a tree. */
static double diag_sum(const double *a, int n)
{
    enum { NB = 2048, NT = 256, NF = NB / NT };
    double part[NB];
    int per = (n + NB - 1) / NB;
    for (int b = 0; b < NB; b++) {
        double sh[NT];
        int s0 = b * per, e0 = s0 + per < n ? s0 + per : n;
        for (int t = 0; t < NT; t++) {
            double v = 0;
            for (int e = s0 + t; e < e0; e += NT) v += a[e];
            sh[t] = v;
        }
        for (int off = NT / 2; off > 0; off >>= 1)
            for (int t = 0; t < off; t++) sh[t] += sh[t + off];
        part[b] = sh[0];
    }
    double sh[NT];
    for (int t = 0; t < NT; t++) {
        double v = part[NF * t];
        for (int i = 1; i < NF; i++) v += part[NF * t + i];
        sh[t] = v;
    }
    for (int off = NT / 2; off > 0; off >>= 1)
        for (int t = 0; t < off; t++) sh[t] += sh[t + off];
    return sh[0];
}

/* out: divu_max, divu_min, divu_l2, momentum[3], energy, kinetic, elastic,
 * potential, power, max_dev_detF, dev_detF_l2 (the rmt3d_macro order) */
void orc_macro(orc_sim *S, double *out)
{
    int m = S->m, n = S->n, o = S->o, N = m * n * o;
    const int strides[3] = {1, S->m4, S->mn4};
    const double dhsp[3] = {S->dxsp, S->dysp, S->dzsp};
    double *T = (double *)calloc((size_t)7 * N, sizeof(double));
    double dmax = -1e200, dmin = 1e200, local_max = -1;
    for (int kk = 0; kk < o; kk++)
        for (int jj = 0; jj < n; jj++)
            for (int ii = 0; ii < m; ii++) {
                int e = ii + m * (jj + n * kk);
                int ind = IXd(S, ii + 2, jj + 2, kk + 2);
                const orc_field *f = S->u + ind;
                orc_rmap *rf = S->rm + ind;
                /* total_momentum (fluid_3d_utils.cc:786-797) */
                for (int s = 0; s < 3; s++) T[(size_t)s * N + e] = f->vel[s];
                /* total_energy (:818-856) */
                double kin = 0;
                for (int s = 0; s < 3; s++) kin += 0.5 * f->vel[s] * f->vel[s];
                T[(size_t)3 * N + e] = kin;
                unsigned oid = rf->c & 1023u;
                if (S->cfg.gravity != 0 && oid > 0)
                    T[(size_t)4 * N + e] = S->cfg.gravity * (S->lz[kk + 2] - S->cfg.xyz_bounds[4]);
                if (oid > 0) {
                    double tmp_elas = 0, tmp_detF = 0;
                    for (int s = 0; s < 3; s++) {
                        tmp_elas += rf->energy[s];
                        tmp_detF += fabs(rf->detF[s] - 1);
                        orc_rmap *rp = orc_get_refmap(S, ind + strides[s], oid);
                        if (rp) {
                            tmp_elas += rp->energy[s];
                            tmp_detF += fabs(rp->detF[s] - 1);
                        }
                    }
                    tmp_elas /= 6.;
                    T[(size_t)5 * N + e] = tmp_elas;
                    /* total_dev_in_detF (:1041-1066) */
                    tmp_detF /= 6.;
                    double tmp_sq = tmp_detF * tmp_detF;
                    T[(size_t)6 * N + e] = tmp_sq;
                    if (tmp_sq > local_max) local_max = tmp_sq;
                }
                /* div_u's MAC extremes (:572-594) */
                double div = 0;
                for (int fc = 0; fc < 3; fc++) div += dhsp[fc] * (f->fvel[2 * fc + 1][fc] - f->fvel[2 * fc][fc]);
                if (div > dmax) dmax = div;
                if (div < dmin) dmin = div;
            }
    double r[7];
    for (int q = 0; q < 7; q++) r[q] = diag_sum(T + (size_t)q * N, N);
    free(T);
    double dv = S->dx * S->dy * S->dz;
    out[0] = dmax;
    out[1] = dmin;
    out[2] = orc_div_u(S);
    for (int s = 0; s < 3; s++) out[3 + s] = r[s] * S->dx * S->dy * S->dz;
    double elas = r[5] * dv, kin = r[3] * dv, pot = r[4] * dv, power = 0. * dv;
    out[6] = elas + kin + pot;
    out[7] = kin;
    out[8] = elas;
    out[9] = pot;
    out[10] = power;
    if (local_max == -1.0) local_max = 0;
    out[11] = sqrt(local_max);
    out[12] = sqrt(r[6] * dv);
}

/* compute_v_centroid + compute_avgs: out[6i..6i+5] = centroid x y z and
 * average velocity; the average velocity is stored for the wall repulsion */
void orc_centroids(orc_sim *S, double *out)
{
    int m = S->m, n = S->n, o = S->o, N = m * n * o;
    if (S->n_obj <= 0) return;
    double *T = (double *)malloc(sizeof(double) * 7 * (size_t)N);
    for (int ob = 0; ob < S->n_obj; ob++) {
        memset(T, 0, sizeof(double) * 7 * (size_t)N);
        for (int kk = 0; kk < o; kk++)
            for (int jj = 0; jj < n; jj++)
                for (int ii = 0; ii < m; ii++) {
                    int e = ii + m * (jj + n * kk);
                    int ind = IXd(S, ii + 2, jj + 2, kk + 2);
                    const orc_rmap *rf = S->rm + ind;
                    if ((int)(rf->c & 1023u) != ob + 1) continue;
                    double phiv = orc_phi(S, ob, rf->x);
                    double sfrac = orc_heaviside(S, phiv);
                    const orc_field *ff = S->u + ind;
                    T[e] = sfrac;
                    T[(size_t)1 * N + e] = ff->vel[0] * sfrac;
                    T[(size_t)2 * N + e] = ff->vel[1] * sfrac;
                    T[(size_t)3 * N + e] = ff->vel[2] * sfrac;
                    T[(size_t)4 * N + e] = S->lx[ii + 2] * sfrac;
                    T[(size_t)5 * N + e] = S->ly[jj + 2] * sfrac;
                    T[(size_t)6 * N + e] = S->lz[kk + 2] * sfrac;
                }
        double r[7];
        for (int q = 0; q < 7; q++) r[q] = diag_sum(T + (size_t)q * N, N);
        double *av = S->avg_vel + 3 * ob;
        for (int a = 0; a < 3; a++) av[a] = (r[0] != 0.) ? r[1 + a] / r[0] : 0.;
        if (out) {
            for (int a = 0; a < 3; a++) out[6 * ob + a] = (r[0] != 0.) ? r[4 + a] / r[0] : 0.;
            for (int a = 0; a < 3; a++) out[6 * ob + 3 + a] = av[a];
        }
    }
    free(T);
}
