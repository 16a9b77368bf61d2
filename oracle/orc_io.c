/*
 * ORACLE -- test infrastructure only.
 *
 * Checkpoint and slice formats restated on the oracle's AoS state, in the
 * reference's own loop order:
 *   fluid_3d::write_chk_pt / initialize_from_chk_point / init_from_slice /
 *   copy_chk_slice_to_buf / copy_chk_to_field (fluid_3d_chk_pt.cc:3-352),
 *   fluid_3d::slice / copy_slice_to_buf / copy_slice_from_buf / min_phi_id
 *   (fluid_3d_io.cc:98-504).
 * One rank holds the whole grid, so every local offset (ai, aj, ak) is 0 and
 * each MPI exchange reduces to the master copying its own buffer.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "orc.h"

#define LMASK (63u << 10)   /* defs.hh:14 */

static const char *const chk_names[8] = {"vx", "vy", "vz", "p", "X", "Y", "Z", "OID"};
static const int chk_types[8] = {0, 1, 2, 7, 9, 10, 11, 12};

static inline int corner_field(int t) { return (t >= 3 && t < 6) || t == 7 || t == 18; }

/* ref_map::phi (fields.hh:430-434) */
static double rm_phi(const orc_sim *S, const orc_rmap *r, int pred)
{
    int o = (int)(r->c & 1023u);
    if (o == 0) return NAN;
    return orc_phi(S, o - 1, pred ? r->xpred : r->x);
}

/* fluid_3d::min_phi_id (fluid_3d_io.cc:420-432); eid is the padded index */
static int min_phi_id(const orc_sim *S, int eid)
{
    int lo = (int)(S->rm[eid].c & 1023u), obj_id = -1;
    double mphi = 1e30, tphi;
    if (lo > 0) { obj_id = lo; mphi = rm_phi(S, &S->rm[eid], 0); }
    for (int n = 0; n < S->ext.cnt[eid]; n++) {
        const orc_rmap *p = &S->ext.e[S->ext.start[eid] + n].r;
        tphi = rm_phi(S, p, 0);
        if (tphi < mphi) {
            mphi = tphi;
            obj_id = (int)(p->c & 1023u);
        }
    }
    return obj_id;
}

/* the body of copy_slice_to_buf's loop (fluid_3d_io.cc:282-411) at local
 * (i,j,k); r0 is the pressure-Poisson rhs for o_type 18 */
double orc_slice_point(orc_sim *S, const double *r0, int t, int wobj, int i, int j, int k)
{
    int sm4 = S->m4, smn4 = S->mn4;
    int eid = S->G0 + i + sm4 * j + smn4 * k;
    int eidlow = eid - smn4;
    const orc_field *u0 = S->u;
    if (t < 3) return u0[eid].vel[t];
    if (t == 3)
        return 0.25 * S->dysp * (u0[eid].vel[2] - u0[eid - sm4].vel[2] + u0[eid - 1].vel[2] - u0[eid - 1 - sm4].vel[2]
                                 + u0[eidlow].vel[2] - u0[eidlow - sm4].vel[2] + u0[eidlow - 1].vel[2]
                                 - u0[eidlow - 1 - sm4].vel[2])
               - 0.25 * S->dzsp * (u0[eid].vel[1] - u0[eidlow].vel[1] + u0[eid - 1].vel[1] - u0[eidlow - 1].vel[1]
                                   + u0[eid - sm4].vel[1] - u0[eidlow - sm4].vel[1] + u0[eid - 1 - sm4].vel[1]
                                   - u0[eidlow - 1 - sm4].vel[1]);
    if (t == 4)
        return 0.25 * S->dzsp * (u0[eid].vel[0] - u0[eidlow].vel[0] + u0[eid - 1].vel[0] - u0[eidlow - 1].vel[0]
                                 + u0[eid - sm4].vel[0] - u0[eidlow - sm4].vel[0] + u0[eid - 1 - sm4].vel[0]
                                 - u0[eidlow - 1 - sm4].vel[0])
               - 0.25 * S->dxsp * (u0[eid].vel[2] - u0[eid - 1].vel[2] + u0[eid - sm4].vel[2] - u0[eid - sm4 - 1].vel[2]
                                   + u0[eidlow].vel[2] - u0[eidlow - 1].vel[2] + u0[eidlow - sm4].vel[2]
                                   - u0[eidlow - sm4 - 1].vel[2]);
    if (t == 5)
        return 0.25 * S->dxsp * (u0[eid].vel[1] - u0[eid - 1].vel[1] + u0[eid - sm4].vel[1] - u0[eid - sm4 - 1].vel[1]
                                 + u0[eidlow].vel[1] - u0[eidlow - 1].vel[1] + u0[eidlow - sm4].vel[1]
                                 - u0[eidlow - sm4 - 1].vel[1])
               - 0.25 * S->dysp * (u0[eid].vel[0] - u0[eid - sm4].vel[0] + u0[eid - 1].vel[0] - u0[eid - 1 - sm4].vel[0]
                                   + u0[eidlow].vel[0] - u0[eidlow - sm4].vel[0] + u0[eidlow - 1].vel[0]
                                   - u0[eidlow - 1 - sm4].vel[0]);
    if (t == 6) {
        double tmp = 0;
        for (int ii = 0; ii < 3; ii++) tmp += u0[eid].vel[ii] * u0[eid].vel[ii];
        return sqrt(tmp);
    }
    if (t == 7) return u0[eid].p;
    if (t == 18) return r0[i + S->fm * j + S->fm * S->fn * k];
    if (t == 19)
        return S->dxsp * (u0[eid].fvel[1][0] - u0[eid].fvel[0][0]) + S->dysp * (u0[eid].fvel[3][1] - u0[eid].fvel[2][1])
               + S->dzsp * (u0[eid].fvel[5][2] - u0[eid].fvel[4][2]);
    if (t == 20) return 0.;
    if (t == 21) return 100.;
    int obj_id;
    if (wobj == -1) {
        obj_id = min_phi_id(S, eid);
        if (obj_id == -1) return t == 12 ? 0. : 100.;
    } else {
        obj_id = wobj;
    }
    const orc_rmap *r = &S->rm[eid];
    if ((int)(r->c & 1023u) == obj_id) {
        switch (t) {
            case 8: return rm_phi(S, r, 0);
            case 9: case 10: case 11: return r->x[t - 9];
            case 12: return (double)(r->c & 1023u);
            case 13: return (double)((r->c & LMASK) >> 10);
            case 14: return rm_phi(S, r, 1);
            default: return r->xpred[t - 15];
        }
    }
    for (int n = 0; n < S->ext.cnt[eid]; n++) {
        const orc_rmap *p = &S->ext.e[S->ext.start[eid] + n].r;
        if ((int)(p->c & 1023u) == obj_id) {
            switch (t) {
                case 8: return rm_phi(S, p, 0);
                case 9: case 10: case 11: return p->x[t - 9];
                case 12: return 0.;
                case 13: return (double)((p->c & LMASK) >> 10);
                case 14: return rm_phi(S, p, 1);
                default: return p->xpred[t - 15];
            }
        }
    }
    return t == 12 ? 0. : 100.;
}

/* fluid_3d::slice on one rank (fluid_3d_io.cc:98-255): copy_slice_to_buf
 * fills the send buffer in k, j, i order, copy_slice_from_buf places it at
 * g_val[i + j*gdim1].  Positions of g_val outside the slice are left alone.
 * Returns the slice length, 0 for a refused point. */
int orc_slice(orc_sim *S, int dim, int point, int t, int obj_id, double *g_val)
{
    const int global_dim[3] = {S->m, S->n, S->o};
    const int p_global_dim[3] = {S->fm, S->fn, S->fo};
    int tmp1 = (dim + 1) % 3, tmp2 = (dim + 2) % 3;
    int iter_dim1 = tmp1 < tmp2 ? tmp1 : tmp2, iter_dim2 = tmp1 < tmp2 ? tmp2 : tmp1;
    int global_len = global_dim[iter_dim1], local_dim1 = global_dim[iter_dim1], local_dim2 = global_dim[iter_dim2];
    if (corner_field(t)) {
        global_len = p_global_dim[iter_dim1];
        local_dim1 = p_global_dim[iter_dim1];
        local_dim2 = p_global_dim[iter_dim2];
    }
    /* copy_slice_to_buf (fluid_3d_io.cc:257-418) */
    if (point < 0 || (dim == 0 && point >= S->m) || (dim == 1 && point >= S->n) || (dim == 2 && point >= S->o)) return 0;
    int sp[3] = {0, 0, 0}, ep[3] = {S->m, S->n, S->o};
    if (corner_field(t)) { ep[0] = S->fm; ep[1] = S->fn; ep[2] = S->fo; }
    sp[dim] = point;
    ep[dim] = point + 1;
    double *r0 = NULL;
    if (t == 18) {
        r0 = (double *)malloc(sizeof(double) * (size_t)S->fm * S->fn * S->fo);
        orc_fill_pp_rhs(S, r0);
    }
    int len = (ep[0] - sp[0]) * (ep[1] - sp[1]) * (ep[2] - sp[2]);
    double *buf = (double *)malloc(sizeof(double) * (size_t)len), *b = buf;
    for (int k = sp[2]; k < ep[2]; k++)
        for (int j = sp[1]; j < ep[1]; j++)
            for (int i = sp[0]; i < ep[0]; i++, b++) *b = orc_slice_point(S, r0, t, obj_id, i, j, k);
    /* copy_slice_from_buf (fluid_3d_io.cc:479-504), slice tag 0 */
    b = buf;
    for (int j = 0; j < local_dim2; j++)
        for (int i = 0; i < local_dim1; i++, b++) g_val[i + j * global_len] = *b;
    free(buf);
    free(r0);
    return len;
}

/* fluid_3d::write_chk_pt (fluid_3d_chk_pt.cc:3-72) without the file writes:
 * img[f][k][...] receives what fwrite would send for field f, slice k.  The
 * one gathering buffer u_global is zeroed once before all fields. */
void orc_chk_images(orc_sim *S, double *img)
{
    size_t gslice_len = (size_t)(S->m + 1) * (S->n + 1);
    double *u_global = (double *)calloc(gslice_len, sizeof(double));
    for (int j = 0; j < 8; j++)
        for (int i = 0; i < S->o; i++) {
            orc_slice(S, 2, i, chk_types[j], -1, u_global);
            memcpy(img + ((size_t)j * S->o + i) * gslice_len, u_global, sizeof(double) * gslice_len);
        }
    free(u_global);
}

/* the files of write_chk_pt, uncompressed */
int orc_write_chk_pt(orc_sim *S, const char *dir)
{
    size_t L = (size_t)(S->m + 1) * (S->n + 1), n = (size_t)S->o * L;
    double *img = (double *)malloc(sizeof(double) * 8 * n);
    orc_chk_images(S, img);
    char fn[1024];
    int rc = 0;
    for (int f = 0; f < 8 && !rc; f++) {
        snprintf(fn, sizeof fn, "%s/%s", dir, chk_names[f]);
        FILE *fh = fopen(fn, "wb");
        if (!fh) { rc = -1; break; }
        if (fwrite(img + (size_t)f * n, sizeof(double), n, fh) != n) rc = -1;
        fclose(fh);
    }
    free(img);
    return rc;
}

/* the reading loop of initialize_from_chk_point (fluid_3d_chk_pt.cc:103-150)
 * into img[f][k][...]: 1 when a file cannot be opened, 2 when a slice read
 * returns nothing, else 0 */
int orc_chk_read(const char *dir, int o, size_t L, double *img)
{
    double *u_global = (double *)calloc(L, sizeof(double));
    char fn[1024];
    int err = 0;
    for (int j = 0; j < 8; j++) {
        snprintf(fn, sizeof fn, "%s/%s", dir, chk_names[j]);
        FILE *fh = fopen(fn, "rb");
        if (!fh) { free(u_global); return 1; }
        for (int i = 0; i < o; i++) {
            size_t bs = fread(u_global, sizeof(double), L, fh);
            if (bs == 0) err = 2;
            memcpy(img + ((size_t)j * o + i) * L, u_global, sizeof(double) * L);
        }
        fclose(fh);
    }
    free(u_global);
    return err;
}

/* init_from_slice + copy_chk_slice_to_buf + copy_chk_to_field for every
 * field and z-slice (fluid_3d_chk_pt.cc:165-352).  copy_chk_slice_to_buf
 * strides the global slice by m for every field, the pressure included. */
void orc_chk_restore(orc_sim *S, const double *img)
{
    size_t L = (size_t)(S->m + 1) * (S->n + 1);
    int m = S->m;
    double *cbuf = (double *)malloc(sizeof(double) * L);
    for (int jf = 0; jf < 8; jf++) {
        int o_type = chk_types[jf];
        int lm = S->m, ln = S->n;
        if (corner_field(o_type)) { lm = S->fm; ln = S->fn; }
        for (int k = 0; k < S->o; k++) {
            const double *gp = img + ((size_t)jf * S->o + k) * L;
            double *cp = cbuf;
            for (int j = 0; j < ln; j++)
                for (int i = 0; i < lm; i++, cp++) *cp = gp[i + j * m];
            cp = cbuf;
            for (int j = 0; j < ln; j++)
                for (int i = 0; i < lm; i++) {
                    int ind = S->G0 + i + S->m4 * j + S->mn4 * k;
                    if (o_type <= 2) S->u[ind].vel[o_type] = *(cp++);
                    else if (o_type == 7) S->u[ind].p = *(cp++);
                    else if (o_type <= 11) {
                        S->rm[ind].x[o_type - 9] = S->rm[ind].xpred[o_type - 9] = *cp;
                        cp++;
                    } else S->rm[ind].c = (uint32_t)(*(cp++));
                }
        }
    }
    free(cbuf);
}

/* fluid_3d::initialize_from_chk_point (fluid_3d_chk_pt.cc:74-163) on a
 * freshly constructed state */
int orc_init_from_chk_point(orc_sim *S)
{
    size_t L = (size_t)(S->m + 1) * (S->n + 1);
    double *img = (double *)malloc(sizeof(double) * 8 * (size_t)S->o * L);
    int err = orc_chk_read(S->cfg.chk_dirname, S->o, L, img);
    if (err) { free(img); return err; }
    S->time = S->cfg.current_time;
    S->nt = S->cfg.chk_step;
    /* constructor state of field and ref_map (fields.hh:57-70, 326-357) */
    memset(S->u, 0, sizeof(orc_field) * (size_t)S->Ng);
    for (int a = 0; a < S->Ng; a++) {
        memset(&S->rm[a], 0, sizeof(orc_rmap));
        for (int f = 0; f < 3; f++)
            for (int c = 0; c < 3; c++) S->rm[a].gphi[f][c] = -1.e200;
    }
    orc_chk_restore(S, img);
    free(img);
    orc_communicate(S, ORC_COMM_VEL);
    orc_communicate(S, ORC_COMM_PRES);
    orc_communicate(S, ORC_COMM_C);
    orc_communicate(S, ORC_COMM_X | ORC_COMM_XPRED);
    orc_band_reset(&S->ext, S->Ng);
    orc_init_extrapolate(S);
    return 0;
}
