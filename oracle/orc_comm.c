/*
 * ORACLE -- test infrastructure only (see orc.h).
 *
 * Ghost-cell fills of the single-domain step:
 *  - face_gm boundary conditions (F3D_Source/fields.hh:661-850) applied face
 *    by face in face order (fluid_3d.cc:145-165);
 *  - the halo exchange communicate<flags> (fluid_3d_comm.cc:77-198) reduced to
 *    one rank: a periodic self-copy of the two-wide ghost regions, region
 *    geometry from mpi_table::setup_table (Utilities/comm_info.hh:91-142);
 *    the periodic shift of the reference map stays disabled as in the
 *    reference (fluid_3d_comm.cc:189);
 *  - the band layer and face copies (communicate_a_layer / _extrap_fx,
 *    fluid_3d_comm.cc:203-365) and the mask copy of extrap_helper::comm_mask
 *    (extrap_helper.cc:261-300).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "orc.h"

/* face_gm::setup (fields.hh:691-740) */
void orc_face_setup(orc_sim *S)
{
    int dims[3] = {S->m, S->n, S->o};
    double hs[3] = {S->dx, S->dy, S->dz};
    for (int f = 0; f < 6; f++) {
        orc_face *F = &S->faces[f];
        for (int c = 0; c < 3; c++) {
            F->bct[c] = S->cfg.bct[c][f];
            F->bcv[c] = S->cfg.bcv[c][f];
        }
        F->n = f / 2;
        F->t[0] = 0; F->t[1] = 2;
        if (F->n == 0) F->t[0] = 1;
        if (F->n == 2) F->t[1] = 1;
        F->len[0] = dims[0]; F->len[1] = dims[1]; F->len[2] = dims[2];
        F->len4[0] = 1; F->len4[1] = S->m4; F->len4[2] = S->mn4;
        F->h = hs[F->n];
        F->u0_off = (f % 2 == 0) ? 0 : (F->len[F->n] - 1) * F->len4[F->n];
        F->len[F->n] = 1;
        F->sgn = 1;
        if (f % 2 == 0) { F->len4[F->n] *= -1; F->sgn = -1; }
        F->ghost_cf[0] = f;
        F->ghost_cf[1] = 2 * F->n + (F->sgn < 0);
        /* neighbour table: a wall face has no neighbour (-1) */
        S->neigh_face[f] = S->prd[F->n] ? 0 : -1;
    }
}

static inline int face_node(const orc_sim *S, const orc_face *F, const int ind[3])
{
    int t = 0;
    for (int i = 0; i < 3; i++) t += ind[i] * F->len4[i];
    return S->G0 + F->u0_off + t;
}

static inline int face_mirror(const orc_sim *S, const orc_face *F, const int ind[3])
{
    int tmp[3] = {ind[0], ind[1], ind[2]};
    tmp[F->n] = 1 - ind[F->n];
    return face_node(S, F, tmp);
}

/* face_gm::set_ghost_cc (fields.hh:765-800) */
static void face_set_ghost_cc(orc_sim *S, const orc_face *F)
{
    int st[3], en[3], in[3];
    for (int i = 0; i < 3; i++) { st[i] = -2; en[i] = F->len[i] + 2; }
    st[F->n] = 1; en[F->n] = 3;
    for (in[2] = st[2]; in[2] < en[2]; in[2]++)
        for (in[1] = st[1]; in[1] < en[1]; in[1]++)
            for (in[0] = st[0]; in[0] < en[0]; in[0]++) {
                orc_field *u = S->u + face_node(S, F, in);
                orc_field *um = S->u + face_mirror(S, F, in);
                for (int c = 0; c < 3; c++) {
                    if (F->bct[c] == 1) {
                        u->vel[c] = 2 * F->bcv[c] - um->vel[c];
                    } else if (F->bct[c] == 2) {
                        double dn = 2 * fabs(0.5 * F->h * (2 * in[F->n] - 1));
                        u->vel[c] = F->sgn * dn * F->bcv[c] + um->vel[c];
                    }
                }
            }
}

/* face_gm::set_ghost_cf + compute_ghost_cf (fields.hh:804-849) */
static void face_set_ghost_cf(orc_sim *S, const orc_face *F)
{
    int st[3], en[3], in[3];
    for (int i = 0; i < 3; i++) { st[i] = -2; en[i] = F->len[i] + 2; }
    st[F->n] = 1; en[F->n] = 2;
    for (in[2] = st[2]; in[2] < en[2]; in[2]++)
        for (in[1] = st[1]; in[1] < en[1]; in[1]++)
            for (in[0] = st[0]; in[0] < en[0]; in[0]++) {
                int tmp[3] = {in[0], in[1], in[2]};
                orc_field *y0 = S->u + face_mirror(S, F, in);
                tmp[F->n] = 2;
                orc_field *y1 = S->u + face_mirror(S, F, tmp);
                double fb[3];
                for (int c = 0; c < 3; c++) {
                    fb[c] = F->bcv[c];
                    if (F->bct[c] == 2)
                        fb[c] = (3 * F->h * F->bcv[c] * F->sgn + 9 * y0->vel[c] - y1->vel[c]) * 0.125;
                }
                orc_field *u = S->u + face_node(S, F, in);
                orc_field *um = y0;
                for (int c = 0; c < 3; c++) {
                    um->fvel[F->ghost_cf[0]][c] = fb[c];
                    u->fvel[F->ghost_cf[1]][c] = fb[c];
                }
            }
}

/* fluid_3d::set_bc_cc (fluid_3d.cc:145-152) */
void orc_set_bc_cc(orc_sim *S)
{
    for (int s = 0; s < 6; s++)
        if (S->neigh_face[s] == -1) face_set_ghost_cc(S, &S->faces[s]);
    orc_communicate(S, ORC_COMM_VEL);
}

/* fluid_3d::set_bc_cf (fluid_3d.cc:155-165) */
void orc_set_bc_cf(orc_sim *S)
{
    for (int s = 0; s < 6; s++)
        if (S->neigh_face[s] == -1) face_set_ghost_cf(S, &S->faces[s]);
    orc_communicate(S, ORC_COMM_FVEL);
}

/* Source cell of a periodic ghost copy for padded cell (i,j,k), or -1.
 * Region tags follow mpi_table::setup_table: in each dimension t=0 for the two
 * low ghost layers, t=2 from s+2 upward, t=1 for [2, s+2) with s the local
 * extent of the table's grid (cells, or FEM nodes for the corner table used by
 * communicate<2>, fluid_3d_comm.cc:81).  A region receives data iff every
 * dimension with t != 1 is periodic; its data come from the opposite side. */
int orc_periodic_src(const orc_sim *S, int i, int j, int k, int corner, int *src)
{
    int c[3] = {i, j, k};
    int s[3];
    if (corner) { s[0] = S->fm; s[1] = S->fn; s[2] = S->fo; }
    else { s[0] = S->m; s[1] = S->n; s[2] = S->o; }
    int per[3] = {S->m, S->n, S->o};
    int any = 0;
    for (int d = 0; d < 3; d++) {
        int t = (c[d] < 2) ? 0 : ((c[d] >= s[d] + 2) ? 2 : 1);
        if (t != 1) {
            if (!S->prd[d]) return -1;
            any = 1;
            c[d] += (t == 0) ? per[d] : -per[d];
        }
    }
    if (!any) return -1;
    *src = c[0] + S->m4 * (c[1] + S->n4 * c[2]);
    return 0;
}

/* communicate<flags> on one rank (fluid_3d_comm.cc:77-198 with
 * copy_to_buffer / copy_from_buffer :488-536).  Sources are interior cells,
 * never targets, so a plain copy reproduces the pack-all-then-unpack order. */
void orc_communicate(orc_sim *S, int flags)
{
    if (!(S->prd[0] || S->prd[1] || S->prd[2])) return;
    int corner = (flags == ORC_COMM_PRES);
    for (int k = 0; k < S->o4; k++)
        for (int j = 0; j < S->n4; j++)
            for (int i = 0; i < S->m4; i++) {
                int src;
                if (orc_periodic_src(S, i, j, k, corner, &src)) continue;
                int dst = i + S->m4 * (j + S->n4 * k);
                orc_field *ud = S->u + dst, *us = S->u + src;
                orc_rmap *rd = S->rm + dst, *rs = S->rm + src;
                if (flags & ORC_COMM_VEL) memcpy(ud->vel, us->vel, sizeof ud->vel);
                if (flags & ORC_COMM_PRES) { ud->p = us->p; ud->dp = us->dp; ud->q = us->q; }
                if (flags & ORC_COMM_FVEL) memcpy(ud->fvel, us->fvel, sizeof ud->fvel);
                if (flags & ORC_COMM_C) rd->c = rs->c;
                if (flags & ORC_COMM_X) memcpy(rd->x, rs->x, sizeof rd->x);
                if (flags & ORC_COMM_XPRED) memcpy(rd->xpred, rs->xpred, sizeof rd->xpred);
                if (flags & ORC_COMM_FX) memcpy(rd->fx, rs->fx, sizeof rd->fx);
                if (flags & ORC_COMM_DVEL) memcpy(ud->dvel, us->dvel, sizeof ud->dvel);
            }
}

/* communicate_extrap_fx (fluid_3d_comm.cc:285-365) with pack_extrap type 2
 * and unpack_extrap type 2 (extrap_mmap.hh:104-159): a ghost entry takes the
 * face values of the source entry with the same c. */
void orc_communicate_band_fx(orc_sim *S)
{
    if (!(S->prd[0] || S->prd[1] || S->prd[2])) return;
    orc_band *B = &S->ext;
    for (int k = 0; k < S->o4; k++)
        for (int j = 0; j < S->n4; j++)
            for (int i = 0; i < S->m4; i++) {
                int src;
                if (orc_periodic_src(S, i, j, k, 0, &src)) continue;
                int dst = i + S->m4 * (j + S->n4 * k);
                for (int a = 0; a < B->cnt[src]; a++) {
                    orc_rmap *rs = &B->e[B->start[src] + a].r;
                    for (int b = 0; b < B->cnt[dst]; b++) {
                        orc_rmap *rd = &B->e[B->start[dst] + b].r;
                        if (rd->c == rs->c) memcpy(rd->fx, rs->fx, sizeof rd->fx);
                    }
                }
            }
}

/* extrap_helper::comm_mask (extrap_helper.cc:261-300) on one rank */
void orc_comm_mask(orc_sim *S)
{
    if (!(S->prd[0] || S->prd[1] || S->prd[2])) return;
    for (int k = 0; k < S->o4; k++)
        for (int j = 0; j < S->n4; j++)
            for (int i = 0; i < S->m4; i++) {
                int src;
                if (orc_periodic_src(S, i, j, k, 0, &src)) continue;
                S->mask[i + S->m4 * (j + S->n4 * k)] = S->mask[src];
            }
}
