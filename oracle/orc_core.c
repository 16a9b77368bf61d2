/*
 * ORACLE -- test infrastructure only (see orc.h).
 *
 * Set-up, initialisation and the stages of fluid_3d::step_forward
 * (F3D_Source/fluid_3d.cc:429-577), each function citing the reference lines
 * it follows.  Arithmetic is FP64 in the reference's order of operations.
 */
#include <float.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "orc.h"
#include "orc_detmath.h"

#define OMASK 1023u
#define LDIG 10
#define SMALL_NUMBER (100. * DBL_EPSILON)   /* defs.hh:45 */

/* Sum of a[e]*scale in the fixed blocked order of the HIP reduction: 256
 * chunks of ceil(n/256) elements, 256 strided partial sums per chunk folded
 * by a binary tree, chunk sums added in order.  The reference sums
 * sequentially (fluid_3d.cc:2698-2699); the blocked order differs only in
 * rounding and makes the HIP path and this oracle bit-identical. */
double orc_blocked_sum(const double *a, int n, double scale)
{
    const int NB = 256, NT = 256;
    int per = (n + NB - 1) / NB;
    double tot = 0;
    for (int b = 0; b < NB; b++) {
        double sh[256];
        int s0 = b * per, e0 = s0 + per < n ? s0 + per : n;
        for (int t = 0; t < NT; t++) {
            double v = 0;
            for (int e = s0 + t; e < e0; e += NT) v += a[e] * scale;
            sh[t] = v;
        }
        for (int off = NT / 2; off > 0; off >>= 1)
            for (int t = 0; t < off; t++) sh[t] += sh[t + off];
        tot += sh[0];
    }
    return tot;
}

static inline int IX(const orc_sim *S, int i, int j, int k) { return i + S->m4 * (j + S->n4 * k); }
static inline double dmin(double a, double b) { return (a < b) ? a : b; }
static inline double dsign(double x) { return (x >= 0) ? 1 : -1; }

/* ------------------------------------------------------------------ */
/* sim_manager / objects                                                */
/* ------------------------------------------------------------------ */

/* sim_manager::phi -> basic_sphere::phi (sim_manager.cc:97-100, object.cc:76-82) */
double orc_phi(const orc_sim *S, int oi, const double xi[3])
{
    if (oi < 0 || oi >= S->n_obj) return NAN;
    const orc_obj *O = &S->obj[oi];
    double sq = 0;
    for (int i = 0; i < 3; i++) sq += (xi[i] - O->c[i]) * (xi[i] - O->c[i]);
    return sqrt(sq) - O->R;
}

static inline double rm_phi(const orc_sim *S, const orc_rmap *r)
{
    int o = (int)(r->c & OMASK);
    if (o == 0) return NAN;
    return orc_phi(S, o - 1, r->x);
}

/* sim_manager::heaviside (sim_manager.hh:259-268) */
double orc_heaviside(const orc_sim *S, double phi)
{
    if (phi >= S->eps) return 0;
    if (phi <= -S->eps) return 1;
    double tmp = 0.5 - phi * S->tf1 - det_sin(S->tf3 * phi) * S->tf2;
    if (tmp < 0) tmp = 0.;
    else if (tmp > 1) tmp = 1;
    return tmp;
}

/* sim_manager.hh:269-273 */
static inline double tderiv_func_in(const orc_sim *S, double x)
{
    double tmp = 0;
    if (x > -S->eps && x < S->eps) tmp = 0.5 + 0.5 * det_cos(x * S->tf3);
    return tmp;
}

/* sim_manager.hh:274-285 */
static inline double delta_func_wall(double x) { return (x <= 0) ? 100 : (1 - x) / x; }
static inline double coll_func(const orc_sim *S, double x) { return x > S->eps ? 0 : (0.5 - x * S->tf1); }
static inline double wall_trans_func(double phiv, double w, double winv)
{
    if (phiv < -w) return 1;
    else if (phiv < w) return 0.5 * (1 - phiv * winv);
    return 0;
}

/* sphere::rm0 / object::rm0 (object.hh:41-43, 110-114) */
static void obj_rm0(const orc_obj *O, const double x[3], double xi[3])
{
    if (O->type == RMT3D_SPHERE) {
        for (int i = 0; i < 3; i++) xi[i] = (x[i] - O->c[i]) * O->eig[i] + O->c[i];
    } else {
        for (int i = 0; i < 3; i++) xi[i] = x[i];
    }
}

/* sim_ftest::velocity (sim_manager.cc:139-146) */
static void ftest_vel(double x, double y, double z, double *u, double *v, double *w)
{
    *u = -det_sin(2 * M_PI * x) * det_cos(2 * M_PI * y) * det_cos(2 * M_PI * z);
    *v = 0.5 * det_cos(2 * M_PI * x) * det_sin(2 * M_PI * y) * det_cos(2 * M_PI * z);
    *w = 0.5 * det_cos(2 * M_PI * x) * det_cos(2 * M_PI * y) * det_sin(2 * M_PI * z);
}

/* sim_manager::velocity / sim_ftest::velocity (sim_manager.cc:102-146) */
static void mgmt_velocity(orc_sim *S, double x, double y, double z, double *u, double *v, double *w)
{
    if (S->cfg.sim_type == RMT3D_SIM_FTEST || S->cfg.sim_type == RMT3D_SIM_MILLS) {
        ftest_vel(x, y, z, u, v, w);
        double tmp = sqrt((*u) * (*u) + (*v) * (*v) + (*w) * (*w));
        if (tmp > S->vmax) S->vmax = tmp;
        return;
    }
    *u = 0; *v = 0; *w = 0;
    const double *vpp = S->cfg.vel_profile;
    for (int i = 0; i < S->cfg.vel_prof_num; i++) {
        int vt = (int)vpp[2 * i];
        double amp = vpp[2 * i + 1];
        if (vt == 1) *u += amp;
        else if (vt == 2) *v += amp;
        else if (vt == 3) *w += amp;
    }
    double co[3] = {x, y, z};
    for (int i = 0; i < S->n_obj; i++) {
        double rx[3];
        obj_rm0(&S->obj[i], co, rx);
        if (orc_phi(S, i, rx) < 0.) {
            double L[3];
            const orc_obj *O = &S->obj[i];
            for (int a = 0; a < 3; a++) L[a] = rx[a] - O->c[a];
            *u += O->omega0[1] * L[2] - O->omega0[2] * L[1];
            *v -= O->omega0[0] * L[2] - O->omega0[2] * L[0];
            *w += O->omega0[0] * L[1] - O->omega0[1] * L[0];
        }
    }
    double tmp = sqrt((*u) * (*u) + (*v) * (*v) + (*w) * (*w));
    if (tmp > S->vmax) S->vmax = tmp;
}

/* sim_ftest::pressure (sim_manager.cc:148-150); zero otherwise */
static double mgmt_pressure(const orc_sim *S, double x, double y, double z)
{
    if (S->cfg.sim_type == RMT3D_SIM_FTEST || S->cfg.sim_type == RMT3D_SIM_MILLS)
        return 3 * M_PI * S->cfg.fmu * det_cos(2 * M_PI * x) * det_cos(2 * M_PI * y) * det_cos(2 * M_PI * z);
    return 0;
}

/* sim_ftest::exact (sim_manager.cc:152-167) */
static void mgmt_exact(const orc_sim *S, double x, double y, double z, double t, double h,
                       double *u, double *v, double *w, double *p)
{
    if (S->cfg.sim_type == RMT3D_SIM_FTEST || S->cfg.sim_type == RMT3D_SIM_MILLS) {
        ftest_vel(x, y, z, u, v, w);
        *u *= det_cos(2 * M_PI * t); *v *= det_cos(2 * M_PI * t); *w *= det_cos(2 * M_PI * t);
        x -= 0.5 * h; y -= 0.5 * h; z -= 0.5 * h;
        *p = mgmt_pressure(S, x, y, z) * (det_cos(2 * M_PI * t) - det_sin(2 * M_PI * t) / (6 * M_PI * S->cfg.fmu));
        return;
    }
    *u = *v = *w = *p = 0;
}

/* fluid_acceleration: sim_ftest (sim_manager.cc:169-178), sim_objects
 * (sim_manager.hh:400-404), base class zero */
static void mgmt_fluid_acc(const orc_sim *S, double x, double y, double z, double t, double f[3])
{
    int st = S->cfg.sim_type;
    if (st == RMT3D_SIM_FTEST || st == RMT3D_SIM_MILLS) {
        double mu = S->cfg.fmu, P = M_PI;
        f[0] = (12 * P * det_cos(2 * P * y) * det_cos(2 * P * z) * (-6 * P * mu * det_cos(2 * P * t) + det_sin(2 * P * t)) * det_sin(2 * P * x) +
                P * det_cos(2 * P * t) * det_cos(2 * P * t) * (2 + det_cos(4 * P * y) + det_cos(4 * P * z)) * det_sin(4 * P * x)) / 4;
        f[1] = -(P * det_cos(2 * P * t) * det_cos(2 * P * t) * (-1 + det_cos(4 * P * x) - 2 * det_cos(4 * P * z)) * det_sin(4 * P * y)) / 8;
        f[2] = -(P * det_cos(2 * P * t) * det_cos(2 * P * t) * (-1 + det_cos(4 * P * x) - 2 * det_cos(4 * P * y)) * det_sin(4 * P * z)) / 8;
        return;
    }
    f[0] = f[1] = f[2] = 0;
    if (st == RMT3D_SIM_OBJECTS) f[2] += S->gravity;
}

/* sim_objects::solid_acceleration (sim_manager.cc:212-295) for spheres
 * (no anchors: object::passive_acc_deformation, object.hh:56-58) */
static void solid_acc(const orc_sim *S, int oi, const double rval[3], double x, double y, double z, double f[3])
{
    f[0] = f[1] = f[2] = 0;
    if (S->cfg.sim_type != RMT3D_SIM_OBJECTS) return;
    const double *av = S->avg_vel + 3 * oi;
    double krx = S->wall_acc_mult * av[0] * av[0];
    double kry = S->wall_acc_mult * av[1] * av[1];
    double krz = S->wall_acc_mult * av[2] * av[2];
    if (krx < S->min_wall_acc) krx = S->min_wall_acc;
    if (kry < S->min_wall_acc) kry = S->min_wall_acc;
    if (krz < S->min_wall_acc) krz = S->min_wall_acc;
    double phiv = orc_phi(S, oi, rval);
    double d;
    if (S->walls[0]) { d = x - S->wall_pos[0]; if (d <= S->wall_dist) f[0] += delta_func_wall(d * S->winv) * krx * wall_trans_func(phiv, S->dx, S->dxsp); }
    if (S->walls[1]) { d = S->wall_pos[1] - x; if (d <= S->wall_dist) f[0] -= delta_func_wall(d * S->winv) * krx * wall_trans_func(phiv, S->dx, S->dxsp); }
    if (S->walls[2]) { d = y - S->wall_pos[2]; if (d <= S->wall_dist) f[1] += delta_func_wall(d * S->winv) * kry * wall_trans_func(phiv, S->dy, S->dysp); }
    if (S->walls[3]) { d = S->wall_pos[3] - y; if (d <= S->wall_dist) f[1] -= delta_func_wall(d * S->winv) * kry * wall_trans_func(phiv, S->dy, S->dysp); }
    if (S->walls[4]) { d = z - S->wall_pos[4]; if (d <= S->wall_dist) f[2] += delta_func_wall(d * S->winv) * krz * wall_trans_func(phiv, S->dz, S->dzsp); }
    if (S->walls[5]) { d = S->wall_pos[5] - z; if (d <= S->wall_dist) f[2] -= delta_func_wall(d * S->winv) * krz * wall_trans_func(phiv, S->dz, S->dzsp); }
    double fn = sqrt(f[0] * f[0] + f[1] * f[1] + f[2] * f[2]);
    if (fn >= S->K_stiff) {
        double nf = S->K_stiff / fn;
        f[0] *= nf; f[1] *= nf; f[2] *= nf;
    }
    /* object::passive_acc_deformation leaves the anchor displacement zero
     * for spheres, so the anchor term K*(ldx,ldy,ldz) vanishes */
    f[2] += S->gravity * orc_heaviside(S, phiv);
}

/* ------------------------------------------------------------------ */
/* construction                                                         */
/* ------------------------------------------------------------------ */
orc_sim *orc_create(const rmt3d_config *cfg, const rmt3d_object *objs)
{
    /* the multigrid colourings need an even cell count across a wrap */
    for (int d = 0; d < 3; d++)
        if (cfg->prd[d] && (cfg->sys_size[d] % 2)) return NULL;
    orc_sim *S = (orc_sim *)calloc(1, sizeof(orc_sim));
    S->cfg = *cfg;
    int st = cfg->sim_type;
    S->n_obj = (st == RMT3D_SIM_FTEST) ? 0 : cfg->n_obj;
    S->m = cfg->sys_size[0]; S->n = cfg->sys_size[1]; S->o = cfg->sys_size[2];
    S->m4 = S->m + 4; S->n4 = S->n + 4; S->o4 = S->o + 4;
    S->mn4 = S->m4 * S->n4;
    S->Ng = S->mn4 * S->o4;
    S->G0 = 2 * (1 + S->m4 * (1 + S->n4));
    for (int d = 0; d < 3; d++) S->prd[d] = cfg->prd[d];
    S->fm = S->prd[0] ? S->m : S->m + 1;
    S->fn = S->prd[1] ? S->n : S->n + 1;
    S->fo = S->prd[2] ? S->o : S->o + 1;
    const double *b = cfg->xyz_bounds;
    S->lxL = b[1] - b[0]; S->lyL = b[3] - b[2]; S->lzL = b[5] - b[4];
    S->dx = S->lxL / S->m; S->dy = S->lyL / S->n; S->dz = S->lzL / S->o;
    S->dxsp = 1.0 / S->dx; S->dysp = 1.0 / S->dy; S->dzsp = 1.0 / S->dz;
    S->lx = (double *)malloc(sizeof(double) * (size_t)S->m4);
    S->ly = (double *)malloc(sizeof(double) * (size_t)S->n4);
    S->lz = (double *)malloc(sizeof(double) * (size_t)S->o4);
    for (int i = 0; i < S->m4; i++) S->lx[i] = b[0] + (i - 2 + 0.5) * S->dx;
    for (int j = 0; j < S->n4; j++) S->ly[j] = b[2] + (j - 2 + 0.5) * S->dy;
    for (int k = 0; k < S->o4; k++) S->lz[k] = b[4] + (k - 2 + 0.5) * S->dz;

    /* sim_manager ctor + create_objs + setup_const (sim_manager.hh:164-210,
     * sim_manager.cc:19-81) */
    double dh = S->dx < S->dy ? S->dx : S->dy;
    dh = dh < S->dz ? dh : S->dz;
    S->vmax = 1;
    S->obj = (orc_obj *)calloc((size_t)(S->n_obj > 0 ? S->n_obj : 1), sizeof(orc_obj));
    S->sm = (orc_slmat *)calloc((size_t)(S->n_obj > 0 ? S->n_obj : 1), sizeof(orc_slmat));
    S->avg_vel = (double *)calloc(3 * (size_t)(S->n_obj > 0 ? S->n_obj : 1), sizeof(double));
    for (int i = 0; i < S->n_obj; i++) {
        orc_slmat *M = &S->sm[i];
        M->rho = objs[i].srho; M->invrho = 1 / M->rho; M->G = objs[i].shear_mod;
        M->c = sqrt(M->G / M->rho);
        M->ex_visc_mult = cfg->ex_visc_mult; M->ev_trans_mult = cfg->ev_trans_mult;
        M->dt_pad = cfg->sdt_pad; M->dt_ex_pad = cfg->dt_ex_pad;
        M->ex_mu = M->ex_visc_mult * cfg->fmu;  /* create_objs override (sim_manager.cc:72) */
        orc_obj *O = &S->obj[i];
        O->type = objs[i].type;
        for (int a = 0; a < 3; a++) O->c[a] = objs[i].basic[a];
        O->R = objs[i].basic[3];
        O->eig[0] = O->eig[1] = O->eig[2] = 1;
        if (O->type == RMT3D_SPHERE) {
            int sd = (int)objs[i].extra[0];
            double sl = objs[i].extra[1];
            O->eig[0] = O->eig[1] = O->eig[2] = sl;
            O->eig[sd] = 1. / (sl * sl);
            for (int a = 0; a < 3; a++) O->omega0[a] = objs[i].extra[2 + a];
        }
    }
    {
        double xxsp3 = 3.0 / (dh * dh);
        double dt_s = 100000;
        for (int i = 0; i < S->n_obj; i++) {
            orc_slmat *M = &S->sm[i];
            double x3 = 3 / (dh * dh);
            double dt1 = 0.5 * M->rho / (M->ex_mu * (M->ev_trans_mult + 1) * x3) * M->dt_ex_pad;
            double dt0 = dh / M->c * M->dt_pad;
            double odt = (dt0 < dt1) ? dt0 : dt1;
            if (odt < dt_s) dt_s = odt;
        }
        double dt_f = 0.5 * cfg->frho / (cfg->fmu * xxsp3) * cfg->fdt_pad;
        S->dt_reg = (dt_s < dt_f) ? dt_s : dt_f;
    }
    S->eps = cfg->wt_n * dh;
    S->eps_inv = 1. / S->eps;
    S->tf1 = 0.5 / S->eps;
    S->tf2 = 0.5 / M_PI;
    S->tf3 = M_PI / S->eps;
    if (cfg->dt < S->dt_reg) S->dt_reg = cfg->dt;
    S->K_stiff = 0.0025 / (S->dt_reg * S->dt_reg);

    if (st == RMT3D_SIM_OBJECTS || st == RMT3D_SIM_MILLS) {
        /* sim_objects ctor + set_wall_acc (sim_manager.hh:363-444) */
        for (int i = 0; i < 6; i++) { S->walls[i] = cfg->walls[i]; S->wall_pos[i] = cfg->wall_pos[i]; }
        S->wall_dist = cfg->wall_dist;
        S->winv = 1 / S->wall_dist;
        S->gravity = cfg->gravity;
        S->min_wall_acc = 20 * S->winv;
        double mvsq = -1, mvol = -1, rho_f = cfg->frho;
        for (int i = 0; i < S->n_obj; i++) {
            double rho_s = S->sm[i].rho;
            double vsq = 0.5 * fabs((rho_s - rho_f) * S->gravity) / rho_f;
            double R = S->obj[i].R;
            double volume = M_PI * 4 / 3 * pow(R, 3);
            double mvf = volume / R / R;
            vsq *= mvf;
            if (vsq > mvsq) mvsq = vsq;
            if (mvf > mvol) mvol = mvf;
            double sa = fabs((rho_s - rho_f) * S->gravity * S->winv * mvf / rho_s);
            if (sa > S->min_wall_acc) S->min_wall_acc = sa;
        }
        S->wall_acc_mult = 0.5 * mvol * pow(S->winv, 2);
        S->vmax = sqrt(mvsq);
    }

    /* fluid_3d ctor (fluid_3d.cc:12-77): setup_faces updates vmax
     * (fluid_3d.cc:2951-2974) */
    for (int f = 0; f < 6; f++)
        for (int c = 0; c < 3; c++)
            if (cfg->bct[c][f] == 1 && cfg->bcv[c][f] > S->vmax) S->vmax = cfg->bcv[c][f];
    orc_face_setup(S);
    S->dt = cfg->dt;
    S->time = 0;
    S->u = (orc_field *)calloc((size_t)S->Ng, sizeof(orc_field));
    S->rm = (orc_rmap *)calloc((size_t)S->Ng, sizeof(orc_rmap));
    for (int a = 0; a < S->Ng; a++)
        for (int f = 0; f < 3; f++)
            for (int c = 0; c < 3; c++) S->rm[a].gphi[f][c] = -1.e200;
    orc_band_init(&S->ext, S->Ng);
    orc_band_init(&S->tmp, S->Ng);
    S->mask = (uint32_t *)calloc((size_t)S->Ng, sizeof(uint32_t));
    S->list = (int *)malloc(sizeof(int) * (size_t)S->Ng);
    S->next = (int *)malloc(sizeof(int) * 6 * (size_t)S->Ng);
    S->nmask = 1; S->cur_mask = 1;
    int dm[3] = {S->m, S->n, S->o}, df[3] = {S->fm, S->fn, S->fo};
    orc_mg_setup(&S->mac, 0, dm, S->prd);
    orc_mg_setup(&S->fem, 1, df, S->prd);
    return S;
}

void orc_destroy(orc_sim *S)
{
    if (!S) return;
    free(S->lx); free(S->ly); free(S->lz);
    free(S->obj); free(S->sm); free(S->avg_vel); free(S->tr);
    free(S->u); free(S->rm);
    orc_band_free(&S->ext); orc_band_free(&S->tmp);
    free(S->mask); free(S->list); free(S->next);
    orc_mg_free(&S->mac); orc_mg_free(&S->fem);
    free(S);
}

/* ------------------------------------------------------------------ */
/* initialisation                                                       */
/* ------------------------------------------------------------------ */
static void field_reset(orc_field *f)   /* field::reset (fields.hh:75-92) */
{
    double p = f->p;
    memset(f, 0, sizeof *f);
    f->p = p;
}

/* sim_manager::obligatory_cfl_recheck (sim_manager.cc:56-66) */
static void cfl_recheck(orc_sim *S)
{
    double dh = S->dx < S->dy ? S->dx : S->dy;
    dh = dh < S->dz ? dh : S->dz;
    double dt_min = S->cfg.fdt_pad * dh / S->vmax;
    if (S->dt_reg > dt_min) {
        S->dt_reg = dt_min;
        if (S->cfg.dt < S->dt_reg) S->dt_reg = S->cfg.dt;
    }
    S->K_stiff = 0.0025 / (S->dt_reg * S->dt_reg);
}

/* fluid_3d::init_fluid (fluid_3d.cc:292-321) */
static void init_fluid(orc_sim *S)
{
    for (int k = 0; k < S->o4; k++)
        for (int j = 0; j < S->n4; j++)
            for (int i = 0; i < S->m4; i++) {
                orc_field *f = S->u + IX(S, i, j, k);
                field_reset(f);
                double u, v, w, p;
                if (S->cfg.has_exact) mgmt_exact(S, S->lx[i], S->ly[j], S->lz[k], S->time, S->dx, &u, &v, &w, &p);
                else mgmt_velocity(S, S->lx[i], S->ly[j], S->lz[k], &u, &v, &w);
                f->vel[0] = u; f->vel[1] = v; f->vel[2] = w;
            }
    orc_communicate(S, ORC_COMM_VEL);
    cfl_recheck(S);
    S->dt = S->dt_reg;
}

/* fluid_3d::init_pres (fluid_3d.cc:326-343) */
static void init_pres(orc_sim *S)
{
    double hdx = S->dx * 0.5, hdy = S->dy * 0.5, hdz = S->dz * 0.5;
    for (int k = 2; k < S->o4 - 1; k++)
        for (int j = 2; j < S->n4 - 1; j++)
            for (int i = 2; i < S->m4 - 1; i++) {
                orc_field *f = S->u + IX(S, i, j, k);
                f->q = 0;
                f->p = mgmt_pressure(S, S->lx[i] - hdx, S->ly[j] - hdy, S->lz[k] - hdz);
            }
    orc_communicate(S, ORC_COMM_PRES);
}

/* fluid_3d::init_refmap (fluid_3d.cc:348-400) */
static int init_refmap(orc_sim *S)
{
    for (int a = 0; a < S->Ng; a++) {
        memset(&S->rm[a], 0, sizeof(orc_rmap));
        for (int f = 0; f < 3; f++)
            for (int c = 0; c < 3; c++) S->rm[a].gphi[f][c] = -1.e200;
    }
    for (int k = 2; k < S->o4 - 2; k++)
        for (int j = 2; j < S->n4 - 2; j++)
            for (int i = 2; i < S->m4 - 2; i++) {
                orc_rmap *r = S->rm + IX(S, i, j, k);
                double X[3] = {S->lx[i], S->ly[j], S->lz[k]};
                int found = 0;
                for (int o = 0; o < S->n_obj; o++) {
                    obj_rm0(&S->obj[o], X, r->x);
                    obj_rm0(&S->obj[o], X, r->xpred);
                    if (orc_phi(S, o, r->x) < 0) {
                        if (!found) { r->c = (uint32_t)(o + 1); found = 1; }
                        else {
                            fprintf(stderr, "oracle: init_refmap: overlapping primary points!\n");
                            return -1;
                        }
                    }
                }
            }
    orc_communicate(S, ORC_COMM_C);
    orc_communicate(S, ORC_COMM_X | ORC_COMM_XPRED);
    return 0;
}

/* fluid_3d::init_extrapolate (fluid_3d_ex.cc:16-27) */
static void init_extrapolate(orc_sim *S)
{
    orc_extrapolate(S);
    for (int a = 0; a < S->ext.n; a++)
        memcpy(S->ext.e[a].r.x, S->ext.e[a].r.xpred, sizeof(double) * 3);
}

/* fluid_3d::compute_grad_basis (fluid_3d.cc:2506-2517) */
static void compute_grad_basis(orc_sim *S)
{
    double fac = 3 * S->dx / S->dt;
    for (int i = 0; i < 8; i++) {
        S->gradphi[i][0] = (2 * (i % 2) - 1) * fac;
        S->gradphi[i][1] = (2 * ((i / 2) % 2) - 1) * fac;
        S->gradphi[i][2] = (2 * (i / 4) - 1) * fac;
    }
}

static void fill_pp_rhs(orc_sim *S, double *rp);

/* fluid_3d::initialize (fluid_3d.cc:192-233): from the checkpoint named by
 * cfg.chk_dirname when chk_num > 0, else (or when that fails) from scratch */
int orc_initialize(orc_sim *S)
{
    int init_err = 1;
    if (S->cfg.chk_num > 0 && S->cfg.chk_dirname[0]) {
        init_err = orc_init_from_chk_point(S);
        if (init_err < 0) return init_err;
        if (init_err == 0) {
            orc_set_bc_cc(S);
            if ((S->cfg.dump_code & 1) && S->cfg.ntracers > 0 && orc_tracers_init(S)) return RMT3D_ERR_FATAL;
            compute_grad_basis(S);
            return 0;
        }
    }
    S->cfg.chk_num = 0;
    S->time = 0;
    init_fluid(S);
    init_pres(S);
    if (init_refmap(S)) return -1;
    init_extrapolate(S);
    orc_set_bc_cc(S);
    /* fluid_3d::alloc_tracers (fluid_3d.cc:115-117, 217-219) */
    if ((S->cfg.dump_code & 1) && S->cfg.ntracers > 0 && orc_tracers_init(S)) return RMT3D_ERR_FATAL;
    compute_grad_basis(S);
    return init_err;
}

void orc_init_extrapolate(orc_sim *S) { init_extrapolate(S); }
void orc_fill_pp_rhs(orc_sim *S, double *rp) { fill_pp_rhs(S, rp); }

/* ------------------------------------------------------------------ */
/* advection (a)                                                        */
/* ------------------------------------------------------------------ */
/* fluid_3d::del_lim / del_f / mono_diff (fluid_3d.cc:2310-2340) */
static inline double del_lim(double u0, double u1, double u2)
{
    double t1 = 2 * (u1 - u0), t2 = 2 * (u2 - u1);
    if (t1 * t2 > 0) return dmin(fabs(t1), fabs(t2));
    return 0;
}
static inline double del_f(double u0, double u1, double u2)
{
    double dl = del_lim(u0, u1, u2);
    double cd = 0.5 * (u2 - u0);
    return dsign(cd) * dmin(fabs(cd), dl);
}
double orc_mono_diff(double u0, double u1, double u2, double u3, double u4)
{
    double dl = del_lim(u1, u2, u3);
    double cd = 0.5 * (u3 - u1);
    double nd = (del_f(u0, u1, u2) + del_f(u2, u3, u4)) / 6.0;
    return dsign(cd) * dmin(dl, fabs(4.0 / 3.0 * cd - nd));
}

/* fluid_3d::normal_derivatives_velocity (fluid_3d.cc:582-617) */
static void normal_derivatives_velocity(orc_sim *S)
{
    int str[3] = {1, S->m4, S->mn4};
    double dhsp[3] = {S->dxsp, S->dysp, S->dzsp};
#pragma omp parallel for schedule(static)
    for (int k = 0; k < S->o; k++)
        for (int j = 0; j < S->n; j++)
            for (int i = 0; i < S->m; i++) {
                int c = IX(S, i + 2, j + 2, k + 2);
                orc_field *f = S->u + c;
                for (int d = 0; d < 3; d++) {
                    int s = str[d];
                    for (int v = 0; v < 3; v++)
                        f->nvd[d][v] = dhsp[d] * orc_mono_diff(S->u[c - 2 * s].vel[v], S->u[c - s].vel[v], f->vel[v],
                                                               S->u[c + s].vel[v], S->u[c + 2 * s].vel[v]);
                }
            }
}

/* ref_map::abut_primary (fields.hh:495-506) */
static int abut_primary(const orc_sim *S, const orc_rmap *r, int ind)
{
    unsigned o = r->c & OMASK;
    int str[3] = {1, S->m4, S->mn4};
    for (int d = 0; d < 3; d++) {
        if ((S->rm[ind + str[d]].c & OMASK) == o) return 1;
        if ((S->rm[ind - str[d]].c & OMASK) == o) return 1;
    }
    return 0;
}

static int refmap_nvd(orc_sim *S, orc_rmap *sf, int ind, int i, int j, int k)
{
    int str[3] = {1, S->m4, S->mn4};
    double dhsp[3] = {S->dxsp, S->dysp, S->dzsp};
    unsigned oid = sf->c & OMASK;
    for (int d = 0; d < 3; d++) {
        /* m1/p1 positions are offset; m2 is offset by -2 then +2 and p2 is
         * never offset (fluid_3d.cc:652-660) */
        int m1p[3] = {i + 2, j + 2, k + 2}, m2p[3] = {i + 2, j + 2, k + 2};
        int p1p[3] = {i + 2, j + 2, k + 2}, p2p[3] = {i + 2, j + 2, k + 2};
        m1p[d] -= 1; p1p[d] += 1;
        int s = str[d];
        orc_rmap *m2 = orc_get_refmap_pos(S, ind - 2 * s, oid, m2p);
        orc_rmap *m1 = orc_get_refmap_pos(S, ind - s, oid, m1p);
        orc_rmap *p1 = orc_get_refmap_pos(S, ind + s, oid, p1p);
        orc_rmap *p2 = orc_get_refmap_pos(S, ind + 2 * s, oid, p2p);
        if (!m2 || !m1 || !p1 || !p2) {
            fprintf(stderr, "oracle: normal_derivatives_refmap: missing necessary cells at (%d %d %d) dir %d\n", i, j, k, d);
            return -1;
        }
        for (int v = 0; v < 3; v++)
            sf->nvd[d][v] = dhsp[d] * orc_mono_diff(m2->x[v], m1->x[v], sf->x[v], p1->x[v], p2->x[v]);
    }
    return 0;
}

/* fluid_3d::normal_derivatives_refmap (fluid_3d.cc:621-806) */
static int normal_derivatives_refmap(orc_sim *S)
{
    for (int k = 0; k < S->o; k++)
        for (int j = 0; j < S->n; j++)
            for (int i = 0; i < S->m; i++) {
                int ind = IX(S, i + 2, j + 2, k + 2);
                orc_rmap *sf = S->rm + ind;
                if ((sf->c & OMASK) > 0)
                    if (refmap_nvd(S, sf, ind, i, j, k)) return -1;
                for (int a = 0; a < S->ext.cnt[ind]; a++) {
                    orc_rmap *e = &S->ext.e[S->ext.start[ind] + a].r;
                    if (abut_primary(S, e, ind))
                        if (refmap_nvd(S, e, ind, i, j, k)) return -1;
                }
            }
    return 0;
}

/* fluid_3d::neg_pres_grad / neg_pres_dgrad (fluid_3d.cc:2347-2376) */
static void neg_pres_grad(const orc_sim *S, int c, int use_dp, double acc[3])
{
    double pr[8];
    for (int n = 0; n < 8; n++) {
        const orc_field *f = S->u + c + n % 2 + S->m4 * ((n / 2) % 2) + (n / 4) * S->mn4;
        pr[n] = use_dp ? f->dp : f->p;
    }
    double xa = 0, ya = 0, za = 0;
    for (int n = 0; n < 4; n++) {
        xa += pr[2 * n + 1] - pr[2 * n];
        int yi = (n / 2) * 4 + n % 2;
        ya += pr[yi + 2] - pr[yi];
        za += pr[n + 4] - pr[n];
    }
    xa *= -0.25 * S->dxsp; ya *= -0.25 * S->dysp; za *= -0.25 * S->dzsp;
    acc[0] = xa; acc[1] = ya; acc[2] = za;
}

/* fluid_3d::acceleration (fluid_3d.cc:980-1033) */
static void acceleration(orc_sim *S, int c, double x, double y, double z, double acc[3], int pres)
{
    const orc_field *fp = S->u + c;
    const orc_field *fx = fp + 1, *fy = fp + S->m4, *fz = fp + S->mn4;
    for (int a = 0; a < 3; a++)
        acc[a] += S->dxsp * (fx->sigma[0][a] - fp->sigma[0][a]) + S->dysp * (fy->sigma[1][a] - fp->sigma[1][a]) +
                  S->dzsp * (fz->sigma[2][a] - fp->sigma[2][a]);
    double tmp[3] = {0, 0, 0};
    if (pres) neg_pres_grad(S, c, 0, tmp);
    for (int a = 0; a < 3; a++) acc[a] += tmp[a];
    for (int a = 0; a < 3; a++) tmp[a] = 0;
    double an[3] = {0, 0, 0}, sfrac = 0;
    for (int s = 0; s < S->n_obj; s++) {
        orc_rmap *rp = orc_get_refmap(S, c, (unsigned)(s + 1));
        if (rp) {
            double phiv = rm_phi(S, rp);
            if (phiv < S->eps) {
                double lf = orc_heaviside(S, phiv);
                solid_acc(S, s, rp->x, x, y, z, an);
                for (int a = 0; a < 3; a++) tmp[a] += an[a];
                sfrac += lf;
            }
        }
    }
    if (sfrac > 1.)
        for (int a = 0; a < 3; a++) tmp[a] /= sfrac;
    for (int a = 0; a < 3; a++) acc[a] += tmp[a];
    if (sfrac < 1.) {
        mgmt_fluid_acc(S, x, y, z, S->time, an);
        for (int a = 0; a < 3; a++) acc[a] += an[a] * (1 - sfrac);
    }
}

/* compute_half_time_edge_velocities + field::godunov_extrap
 * (fluid_3d.cc:1038-1066, fields.hh:147-177, non-TANG branch) */
static void half_time_edge_velocities(orc_sim *S)
{
    double dh2[3] = {S->dx / 2, S->dy / 2, S->dz / 2};
    double dt2 = S->dt / 2.0;
#pragma omp parallel for schedule(static)
    for (int k = 0; k < S->o; k++)
        for (int j = 0; j < S->n; j++)
            for (int i = 0; i < S->m; i++) {
                int c = IX(S, i + 2, j + 2, k + 2);
                double acc[3] = {0, 0, 0};
                acceleration(S, c, S->lx[i + 2], S->ly[j + 2], S->lz[k + 2], acc, 1);
                orc_field *f = S->u + c;
                double adv[3] = {0, 0, 0};
                for (int a = 0; a < 3; a++) {
                    for (int b = 0; b < 3; b++) adv[a] -= f->vel[b] * f->nvd[b][a];
                    adv[a] *= dt2;
                }
                for (int fc = 0; fc < 6; fc++) {
                    int dim = fc / 2;
                    double d = dh2[dim] * (2 * (fc % 2) - 1);
                    for (int b = 0; b < 3; b++)
                        f->fvel[fc][b] = f->vel[b] + d * f->nvd[dim][b] + adv[b] + dt2 * acc[b];
                }
            }
}

/* ref_map::godunov_extrap (fields.hh:572-592, non-TANG) */
static void refmap_godunov_extrap(const orc_sim *S, orc_rmap *r, const orc_field *f)
{
    double dh2[3] = {S->dx / 2, S->dy / 2, S->dz / 2};
    double dt2 = S->dt / 2.0;
    for (int fc = 0; fc < 6; fc++) {
        int dim = fc / 2, a1 = (dim + 1) % 3, a2 = (dim + 2) % 3;
        double nv = f->vel[dim];
        double d = dh2[dim] * (2 * (fc % 2) - 1);
        double v1 = f->vel[a1], v2 = f->vel[a2];
        for (int b = 0; b < 3; b++)
            r->fx[fc][b] = r->x[b] + (d - dt2 * nv) * r->nvd[dim][b] - dt2 * (v1 * r->nvd[a1][b] + v2 * r->nvd[a2][b]);
    }
}

/* fluid_3d::compute_half_time_edge_refmap (fluid_3d.cc:1071-1100) */
static void half_time_edge_refmap(orc_sim *S)
{
#pragma omp parallel for schedule(static)
    for (int k = 0; k < S->o; k++)
        for (int j = 0; j < S->n; j++)
            for (int i = 0; i < S->m; i++) {
                int c = IX(S, i + 2, j + 2, k + 2);
                if ((S->rm[c].c & OMASK) > 0) refmap_godunov_extrap(S, S->rm + c, S->u + c);
                for (int a = 0; a < S->ext.cnt[c]; a++) {
                    orc_rmap *e = &S->ext.e[S->ext.start[c] + a].r;
                    if (abut_primary(S, e, c)) refmap_godunov_extrap(S, e, S->u + c);
                }
            }
}

/* godunov_set_velocity (fluid_3d.cc:2156-2174) */
void orc_godunov_set_velocity(double *nl, double *vl, double *wl, double *nr, double *vr, double *wr)
{
    if (*nl + *nr > 0 && *nl > 0) {
        *nr = *nl; *vr = *vl; *wr = *wl;
    } else if (*nl + *nr < 0 && *nr < 0) {
        *nl = *nr; *vl = *vr; *wl = *wr;
    } else {
        *nl = *nr = 0;
        double t = 0.5 * (*vl + *vr);
        *vl = *vr = t;
        t = 0.5 * (*wl + *wr);
        *wl = *wr = t;
    }
}

/* godunov_set_refmap (fluid_3d.cc:2184-2204) */
void orc_godunov_set_refmap(double nl, double nr, double *ml, double *mvl, double *mwl,
                            double *mr, double *mvr, double *mwr)
{
    if (nl + nr > 0 && nl > 0) {
        *mr = *ml; *mvr = *mvl; *mwr = *mwl;
    } else if (nl + nr < 0 && nr < 0) {
        *ml = *mr; *mvl = *mvr; *mwl = *mwr;
    } else {
        double t = 0.5 * (*ml + *mr);
        *ml = *mr = t;
        t = 0.5 * (*mvl + *mvr);
        *mvl = *mvr = t;
        t = 0.5 * (*mwl + *mwr);
        *mwl = *mwr = t;
    }
}

static int upwind_map(orc_sim *S, orc_field *f, orc_rmap *sfp, int ind, int eid_dummy)
{
    (void)eid_dummy;
    int str[3] = {1, S->m4, S->mn4};
    for (int dir = 0; dir < 3; dir++) {
        orc_field *ng = f + str[dir];
        orc_rmap *mn = orc_get_refmap(S, ind + str[dir], sfp->c & OMASK);
        if (!mn) {
            fprintf(stderr, "oracle: godunov_upwinding_set: missing neighbour map\n");
            return -1;
        }
        int ff = 2 * dir + 1, nf = ff ^ 1;
        int a1 = (dir + 1) % 3, a2 = (dir + 2) % 3;
        orc_godunov_set_refmap(f->fvel[ff][dir], ng->fvel[nf][dir], &sfp->fx[ff][dir], &sfp->fx[ff][a1], &sfp->fx[ff][a2],
                               &mn->fx[nf][dir], &mn->fx[nf][a1], &mn->fx[nf][a2]);
    }
    return 0;
}

/* fluid_3d::godunov_upwinding_set (fluid_3d.cc:811-944), non-TANG */
static int godunov_upwinding_set(orc_sim *S)
{
    orc_communicate(S, ORC_COMM_FVEL | ORC_COMM_FX);
    orc_communicate_band_fx(S);
    int str[3] = {1, S->m4, S->mn4};
    for (int k = -1; k < S->o; k++)
        for (int j = -1; j < S->n; j++)
            for (int i = -1; i < S->m; i++) {
                int ind = IX(S, i + 2, j + 2, k + 2);
                orc_field *f = S->u + ind;
                orc_rmap *sfp = S->rm + ind;
                if ((sfp->c & OMASK) > 0)
                    if (upwind_map(S, f, sfp, ind, 0)) return -1;
                for (int a = 0; a < S->ext.cnt[ind]; a++) {
                    orc_rmap *e = &S->ext.e[S->ext.start[ind] + a].r;
                    int found = 0;
                    for (int d = 0; d < 3 && !found; d++) {
                        if ((S->rm[ind + str[d]].c & OMASK) == (e->c & OMASK)) {
                            if (upwind_map(S, f, e, ind, 0)) return -1;
                            found = 1;
                        }
                    }
                }
                for (int dir = 0; dir < 3; dir++) {
                    orc_field *ng = f + str[dir];
                    int ff = 2 * dir + 1, nf = ff ^ 1;
                    int a1 = (dir + 1) % 3, a2 = (dir + 2) % 3;
                    orc_godunov_set_velocity(&f->fvel[ff][dir], &f->fvel[ff][a1], &f->fvel[ff][a2], &ng->fvel[nf][dir],
                                             &ng->fvel[nf][a1], &ng->fvel[nf][a2]);
                }
            }
    return 0;
}

/* ------------------------------------------------------------------ */
/* projections (d)                                                      */
/* ------------------------------------------------------------------ */
/* fill_mac_rhs / mac.run / extract_mac_soln / communicate<2> /
 * subtract_q_grad (fluid_3d.cc:2209-2304) */
static int mac_solve(orc_sim *S)
{
    double fac = S->dx / (2 * S->dt);
    orc_mglev *L = &S->mac.lev[0];
    for (int k = 0; k < S->o; k++)
        for (int j = 0; j < S->n; j++)
            for (int i = 0; i < S->m; i++) {
                orc_field *e = S->u + IX(S, i + 2, j + 2, k + 2);
                double r = 0;
                for (int f = 0; f < 3; f++) r += fac * (e->fvel[2 * f + 1][f] - e->fvel[2 * f][f]);
                L->b[i + S->m * (j + S->n * k)] = r;
            }
    double res = orc_mg_run(&S->mac);
    S->mac_res = res;
    S->mac_cycles = S->mac.last_cycles;
    if (isnan(res)) {
        fprintf(stderr, "oracle: mac_solve(): NaN detected in multigrid solve\n");
        return -1;
    }
    for (int k = 0; k < S->o; k++)
        for (int j = 0; j < S->n; j++)
            for (int i = 0; i < S->m; i++)
                S->u[IX(S, i + 2, j + 2, k + 2)].q = L->x[i + S->m * (j + S->n * k)];
    orc_communicate(S, ORC_COMM_PRES);
    double f2 = 2 * S->dt * S->dxsp;
    int str[3] = {1, S->m4, S->mn4};
    for (int k = 0; k < S->o; k++)
        for (int j = 0; j < S->n; j++)
            for (int i = 0; i < S->m; i++) {
                int c = IX(S, i + 2, j + 2, k + 2);
                orc_field *e = S->u + c;
                for (int f = 0; f < 3; f++) {
                    int s = str[f];
                    double down = f2 * (e->q - S->u[c - s].q);
                    double up = f2 * (S->u[c + s].q - e->q);
                    e->fvel[2 * f][f] -= down;
                    e->fvel[2 * f + 1][f] -= up;
                }
            }
    return 0;
}

/* fluid_3d::map_ref_to_local (fluid_3d.cc:2524-2548) */
static void map_ref_to_local(const orc_sim *S, int ei, int ej, int ek, int ln[8])
{
    int psm = S->fm, psmn = S->fm * S->fn;
    int eid = ei + ej * psm + ek * psmn;
    int xe = psm - 1, ye = S->fn - 1, ze = S->fo - 1;
    int wd = (S->prd[0] || (ei >= 0 && ei <= S->m - 1)) && (S->prd[1] || (ej >= 0 && ej <= S->n - 1)) &&
             (S->prd[2] || (ek >= 0 && ek <= S->o - 1));
    for (int i = 0; i < 8; i++) {
        if (!wd) { ln[i] = -1; continue; }
        ln[i] = eid + i % 2 + ((i / 2) % 2) * psm + (i / 4) * psmn;
        if (ei + i % 2 > xe) ln[i] = -1;
        if (ei < 0 && i % 2 == 0) ln[i] = -1;
        if (ej + ((i / 2) % 2) > ye) ln[i] = -1;
        if (ej < 0 && ((i / 2) % 2 == 0)) ln[i] = -1;
        if (ek + (i / 4) > ze) ln[i] = -1;
        if (ek < 0 && (i / 4) == 0) ln[i] = -1;
    }
}

/* fluid_3d::fill_pp_rhs (fluid_3d.cc:2554-2649), pres_update = true */
static void fill_pp_rhs(orc_sim *S, double *rp)
{
    memset(rp, 0, sizeof(double) * (size_t)(S->fm * S->fn * S->fo));
    int ln[8];
    for (int k = -1; k < S->fo; k++)
        for (int j = -1; j < S->fn; j++)
            for (int i = -1; i < S->fm; i++) {
                map_ref_to_local(S, i, j, k, ln);
                const orc_field *e = S->u + IX(S, i + 2, j + 2, k + 2);
                for (int n = 0; n < 8; n++) {
                    if (ln[n] < 0) continue;
                    double val = 0;
                    for (int l = 0; l < 3; l++) val += e->dvel[l] * S->gradphi[n][l];
                    rp[ln[n]] += val;
                }
            }
    for (int s = 0; s < 6; s++) {
        if (S->neigh_face[s] != -1) continue;
        const orc_face *F = &S->faces[s];
        int inds[3] = {0, 0, 0}, off[3];
        inds[F->n] = s % 2;
        off[0] = F->u0_off % S->m4;
        off[1] = (F->u0_off / S->m4) % S->n4;
        off[2] = F->u0_off / S->mn4;
        int st[3], en[3];
        for (int a = 0; a < 3; a++) { st[a] = -1; en[a] = F->len[a] + 1; }
        st[F->n] = 0; en[F->n] = 1;
        for (int k = st[2]; k < en[2]; k++)
            for (int j = st[1]; j < en[1]; j++)
                for (int i = st[0]; i < en[0]; i++) {
                    int g[8];
                    map_ref_to_local(S, i + off[0], j + off[1], k + off[2], g);
                    const orc_field *u = S->u + S->G0 + F->u0_off + i + S->m4 * j + S->mn4 * k;
                    for (inds[F->t[0]] = 0; inds[F->t[0]] < 2; inds[F->t[0]]++)
                        for (inds[F->t[1]] = 0; inds[F->t[1]] < 2; inds[F->t[1]]++) {
                            int id = inds[0] + 2 * inds[1] + 4 * inds[2];
                            if (g[id] >= 0) rp[g[id]] += -u->fvel[s][s / 2] * S->gradphi[id][s / 2];
                        }
                }
    }
}

/* pp_solve + extract_pp_soln + communicate<2> + subtract_pres_grad
 * (fluid_3d.cc:2653-2753) */
static int pp_solve(orc_sim *S)
{
    orc_mglev *L = &S->fem.lev[0];
    fill_pp_rhs(S, L->b);
    double res = orc_mg_run(&S->fem);
    S->fem_res = res;
    S->fem_cycles = S->fem.last_cycles;
    if (isnan(res)) return 1;
    int fm = S->fm, fn = S->fn, fo = S->fo;
    double mnoxsp = 1 / (double)(fm * fn * fo);
    double lm = orc_blocked_sum(L->x, fm * fn * fo, mnoxsp);
    for (int k = 0; k < fo; k++)
        for (int j = 0; j < fn; j++)
            for (int i = 0; i < fm; i++) {
                orc_field *f = S->u + IX(S, i + 2, j + 2, k + 2);
                double v = L->x[i + fm * (j + fn * k)];
                f->p += (v - lm);
                f->dp = (v - lm);
            }
    orc_communicate(S, ORC_COMM_PRES);
    for (int k = 0; k < S->o; k++)
        for (int j = 0; j < S->n; j++)
            for (int i = 0; i < S->m; i++) {
                int c = IX(S, i + 2, j + 2, k + 2);
                double npg[3];
                neg_pres_grad(S, c, 1, npg);
                for (int a = 0; a < 3; a++) {
                    npg[a] *= S->dt;
                    S->u[c].vel[a] += npg[a];
                    S->u[c].dvel[a] += npg[a];
                }
            }
    return 0;
}

/* ------------------------------------------------------------------ */
/* reference-map update (c)                                             */
/* ------------------------------------------------------------------ */
/* fluid_3d::update_reference_map (fluid_3d.cc:1149-1255); returns the
 * number of "new extrap can't find old" problems */
static int update_reference_map(orc_sim *S)
{
    double cdt = S->dt;
    for (int k = 0; k < S->o; k++)
        for (int j = 0; j < S->n; j++)
            for (int i = 0; i < S->m; i++) {
                int c = IX(S, i + 2, j + 2, k + 2);
                orc_field *f = S->u + c;
                orc_rmap *sf = S->rm + c;
                if ((sf->c & OMASK) > 0) {
                    double av[3];
                    av[0] = -0.5 * cdt * (f->fvel[0][0] + f->fvel[1][0]) * S->dxsp;
                    av[1] = -0.5 * cdt * (f->fvel[2][1] + f->fvel[3][1]) * S->dysp;
                    av[2] = -0.5 * cdt * (f->fvel[4][2] + f->fvel[5][2]) * S->dzsp;
                    for (int d = 0; d < 3; d++) {
                        double dxt = av[0] * (sf->fx[1][d] - sf->fx[0][d]) + av[1] * (sf->fx[3][d] - sf->fx[2][d]) +
                                     av[2] * (sf->fx[5][d] - sf->fx[4][d]);
                        sf->xpred[d] = sf->x[d] + dxt;
                    }
                }
            }
    orc_communicate(S, ORC_COMM_C);
    orc_communicate(S, ORC_COMM_X | ORC_COMM_XPRED);
    orc_band_copy(&S->tmp, &S->ext, S->Ng);
    orc_band_reset(&S->ext, S->Ng);
    orc_extrapolate(S);
    int prob = 0;
    for (int ind = 0; ind < S->Ng; ind++) {
        orc_rmap *sf = S->rm + ind;
        if (sf->c & OMASK)
            for (int a = 0; a < 3; a++) sf->x[a] = 0.5 * (sf->x[a] + sf->xpred[a]);
        for (int a = 0; a < S->ext.cnt[ind]; a++) {
            orc_rmap *e = &S->ext.e[S->ext.start[ind] + a].r;
            unsigned oid = e->c & OMASK;
            int found = 0;
            for (int b = 0; b < S->tmp.cnt[ind] && !found; b++) {
                orc_rmap *old = &S->tmp.e[S->tmp.start[ind] + b].r;
                if ((old->c & OMASK) == oid) {
                    for (int q = 0; q < 3; q++) e->x[q] = 0.5 * (old->x[q] + e->xpred[q]);
                    found = 1;
                }
            }
            if (!found) {
                double dist = orc_phi(S, (int)oid - 1, e->xpred);
                if (dist < S->eps && (e->c >> LDIG) <= 3) prob++;
                else memcpy(e->x, e->xpred, sizeof(double) * 3);
            }
        }
    }
    return prob;
}

/* fluid_3d::revert_reference_map_update (fluid_3d.cc:1257-1282) */
static void revert_reference_map_update(orc_sim *S)
{
    for (int ind = 0; ind < S->Ng; ind++) {
        orc_rmap *sf = S->rm + ind;
        if (sf->c & OMASK)
            for (int a = 0; a < 3; a++) sf->x[a] = 2 * sf->x[a] - sf->xpred[a];
        for (int a = 0; a < S->ext.cnt[ind]; a++) {
            orc_rmap *e = &S->ext.e[S->ext.start[ind] + a].r;
            unsigned oid = e->c & OMASK;
            int found = 0;
            for (int b = 0; b < S->tmp.cnt[ind] && !found; b++) {
                orc_rmap *old = &S->tmp.e[S->tmp.start[ind] + b].r;
                if ((old->c & OMASK) == oid) {
                    found = 1;
                    memcpy(e->x, old->x, sizeof(double) * 3);
                }
            }
        }
    }
}

/* fluid_3d::update_reference_map_full (fluid_3d.cc:1285-1306) */
static void update_reference_map_full(orc_sim *S)
{
    for (int ind = 0; ind < S->Ng; ind++)
        if ((S->rm[ind].c & OMASK) > 0) memcpy(S->rm[ind].x, S->rm[ind].xpred, sizeof(double) * 3);
    for (int a = 0; a < S->ext.n; a++) memcpy(S->ext.e[a].r.x, S->ext.e[a].r.xpred, sizeof(double) * 3);
}

/* ------------------------------------------------------------------ */
/* stresses (b)                                                         */
/* ------------------------------------------------------------------ */
/* matrix::det / inverse / aat / ata (Utilities/matlib.hh:227-232,303-329,366-370) */
double orc_mat_det(const double v[9])
{
    return v[0] * (v[4] * v[8] - v[5] * v[7]) + v[3] * (v[2] * v[7] - v[1] * v[8]) + v[6] * (v[1] * v[5] - v[2] * v[4]);
}
void orc_mat_inverse(const double v[9], double m2[9])
{
    double d = orc_mat_det(v);
    for (int j = 0; j < 3; j++)
        for (int i = 0; i < 3; i++) {
            m2[3 * j + i] = v[3 * ((i + 1) % 3) + (j + 1) % 3] * v[3 * ((i + 2) % 3) + (j + 2) % 3] -
                            v[3 * ((i + 1) % 3) + (j + 2) % 3] * v[3 * ((i + 2) % 3) + (j + 1) % 3];
            m2[3 * j + i] /= d;
        }
}
static inline int symind(int i, int j)
{
    if (i > j) { int t = j; j = i; i = t; }
    return j + i * (5 - i) / 2;
}
/* aat: s(i,j) = sum_k F(i,k) F(j,k) for j >= i; ata: sum_k F(k,i) F(k,j) */
void orc_mat_aat(const double F[9], double s[6])
{
    for (int i = 0; i < 3; i++)
        for (int j = i; j < 3; j++) {
            double v = 0;
            for (int k = 0; k < 3; k++) v += F[3 * i + k] * F[3 * j + k];
            s[symind(i, j)] = v;
        }
}
void orc_mat_ata(const double F[9], double s[6])
{
    for (int i = 0; i < 3; i++)
        for (int j = i; j < 3; j++) {
            double v = 0;
            for (int k = 0; k < 3; k++) v += F[3 * k + i] * F[3 * k + j];
            s[symind(i, j)] = v;
        }
}

static inline void unpack_eid(const orc_sim *S, int e, int *i, int *j, int *k)
{
    *k = e / S->mn4;
    *j = (e % S->mn4) / S->m4;
    *i = e % S->m4;
}

static void nbr_set(const orc_sim *S, int ind, unsigned oid, const int lp[3], int F, int ad1, int ad2,
                    orc_rmap **r1p, orc_rmap **r1p1, orc_rmap **r1m, orc_rmap **r1m1,
                    orc_rmap **r2p, orc_rmap **r2p1, orc_rmap **r2m, orc_rmap **r2m1)
{
    int str[3] = {1, S->m4, S->mn4};
    int strd = str[F], as1 = str[ad1], as2 = str[ad2];
    int q[3];
    orc_sim *SS = (orc_sim *)S;
#define POS(da, va, db, vb) do { q[0] = lp[0]; q[1] = lp[1]; q[2] = lp[2]; q[da] += va; if (db >= 0) q[db] += vb; } while (0)
    POS(ad1, 1, -1, 0);  *r1p = orc_get_refmap_pos(SS, ind + as1, oid, q);
    POS(ad1, 1, F, -1);  *r1p1 = orc_get_refmap_pos(SS, ind + as1 - strd, oid, q);
    POS(ad1, -1, -1, 0); *r1m = orc_get_refmap_pos(SS, ind - as1, oid, q);
    POS(ad1, -1, F, -1); *r1m1 = orc_get_refmap_pos(SS, ind - as1 - strd, oid, q);
    POS(ad2, 1, -1, 0);  *r2p = orc_get_refmap_pos(SS, ind + as2, oid, q);
    POS(ad2, 1, F, -1);  *r2p1 = orc_get_refmap_pos(SS, ind + as2 - strd, oid, q);
    POS(ad2, -1, -1, 0); *r2m = orc_get_refmap_pos(SS, ind - as2, oid, q);
    POS(ad2, -1, F, -1); *r2m1 = orc_get_refmap_pos(SS, ind - as2 - strd, oid, q);
#undef POS
}

/* fluid_3d::solid_stress<F> (fluid_3d.cc:1777-1997), spheres: no active
 * stress, actuate is a no-op */
static void solid_stress(orc_sim *S, int eid, int F, double ss[3], double *sfrac)
{
    int str[3] = {1, S->m4, S->mn4};
    double facs[3] = {S->dxsp, S->dysp, S->dzsp};
    int strd = str[F];
    int ad1 = (F + 1) % 3, ad2 = (F + 2) % 3;
    int ind = S->G0 + eid;
    int lp[3];
    unpack_eid(S, ind, &lp[0], &lp[1], &lp[2]);
    const orc_field *fp = S->u + ind;
    int ne = S->ext.cnt[ind];
    for (int ii = -1; ii < ne; ii++) {
        orc_rmap *rf = (ii == -1) ? S->rm + ind : &S->ext.e[S->ext.start[ind] + ii].r;
        unsigned oid = rf->c & OMASK;
        for (int a = 0; a < 3; a++) rf->gphi[F][a] = -1.e200;
        rf->energy[F] = -1.e200;
        rf->detF[F] = -1.e200;
        if (oid == 0) continue;
        int q[3] = {lp[0], lp[1], lp[2]};
        q[F] -= 1;
        orc_rmap *rn = orc_get_refmap_pos(S, ind - strd, oid, q);
        if (!rn) continue;
        double phr = rm_phi(S, rf), phn = rm_phi(S, rn);
        double phiv = 0.5 * (phr + phn);
        if (phiv > S->eps) continue;
        double tf = orc_heaviside(S, phiv);
        *sfrac += tf;
        const orc_slmat *M = &S->sm[oid - 1];
        double G = M->G * tf;
        orc_rmap *r1p, *r1p1, *r1m, *r1m1, *r2p, *r2p1, *r2m, *r2m1;
        nbr_set(S, ind, oid, lp, F, ad1, ad2, &r1p, &r1p1, &r1m, &r1m1, &r2p, &r2p1, &r2m, &r2m1);
        if (!r1p || !r1p1 || !r1m || !r1m1 || !r2p || !r2p1 || !r2m || !r2m1) continue;
        rf->gphi[F][F] = facs[F] * (phr - phn);
        rf->gphi[F][ad1] = 0.25 * facs[ad1] * (rm_phi(S, r1p) + rm_phi(S, r1p1) - rm_phi(S, r1m) - rm_phi(S, r1m1));
        rf->gphi[F][ad2] = 0.25 * facs[ad2] * (rm_phi(S, r2p) + rm_phi(S, r2p1) - rm_phi(S, r2m) - rm_phi(S, r2m1));
        {   /* ref_map::normalize_grad_phi (fields.hh:383-395) */
            double nr = 0;
            for (int a = 0; a < 3; a++) nr += rf->gphi[F][a] * rf->gphi[F][a];
            nr = sqrt(nr);
            if (nr != 0.) {
                double ni = 1. / nr;
                for (int a = 0; a < 3; a++) rf->gphi[F][a] = ni * rf->gphi[F][a];
            }
        }
        double dR[3][3];
        for (int a = 0; a < 3; a++) {
            dR[F][a] = facs[F] * (rf->x[a] - rn->x[a]);
            dR[ad1][a] = 0.25 * facs[ad1] * (r1p->x[a] + r1p1->x[a] - r1m->x[a] - r1m1->x[a]);
            dR[ad2][a] = 0.25 * facs[ad2] * (r2p->x[a] + r2p1->x[a] - r2m->x[a] - r2m1->x[a]);
        }
        double gx[9] = {dR[0][0], dR[1][0], dR[2][0], dR[0][1], dR[1][1], dR[2][1], dR[0][2], dR[1][2], dR[2][2]};
        double Fd[9];
        orc_mat_inverse(gx, Fd);
        rf->detF[F] = orc_mat_det(Fd);
        double sg[6], cg[6];
        orc_mat_aat(Fd, sg);
        orc_mat_ata(Fd, cg);
        rf->energy[F] = 0.5 * G * ((cg[0] + cg[3] + cg[5]) - 3 - 2 * det_log(rf->detF[F]));
        double savg = (sg[0] + sg[3] + sg[5]) / 3.0;
        for (int a = 0; a < 3; a++) sg[symind(a, a)] -= savg;
        double exmu = M->ex_mu * tf * (1 + M->ev_trans_mult * tderiv_func_in(S, phiv));
        for (int a = 0; a < 6; a++) sg[a] *= G;
        for (int a = 0; a < 3; a++) {
            ss[a] += sg[symind(F, a)];
            double dn = fp->vel[a] - fp[-strd].vel[a];
            double tr = (a == F) ? dn
                                 : 0.25 * (fp[str[a]].vel[F] + fp[str[a] - strd].vel[F] - fp[-str[a]].vel[F] -
                                           fp[-str[a] - strd].vel[F]);
            ss[a] += exmu * facs[F] * (dn + tr);
        }
    }
}

/* fluid_3d::collision_stress<F> (fluid_3d.cc:1465-1773); the neighbour
 * positions are unpacked from eid, not G0+eid, as in the reference (:1478) */
static void collision_stress(orc_sim *S, int eid, int F, double fs[3])
{
    int str[3] = {1, S->m4, S->mn4};
    double facs[3] = {S->dxsp, S->dysp, S->dzsp};
    int strd = str[F];
    int ad1 = (F + 1) % 3, ad2 = (F + 2) % 3;
    int ind = S->G0 + eid;
    int lp[3];
    unpack_eid(S, eid, &lp[0], &lp[1], &lp[2]);
    int pobj = (S->rm[ind].c & OMASK) != 0;
    int eobj = S->ext.cnt[ind];
    int nobj = pobj + eobj;
    if (nobj <= 1) return;
    orc_rmap *co[80];
    int nco = 0;
    if (pobj) co[nco++] = S->rm + ind;
    for (int a = 0; a < eobj && nco < 80; a++) co[nco++] = &S->ext.e[S->ext.start[ind] + a].r;
    for (int ii = 0; ii < nobj - 1; ii++) {
        orc_rmap *rf0 = co[ii];
        unsigned o0 = rf0->c & OMASK;
        double G0 = S->sm[o0 - 1].G;
        int q[3] = {lp[0], lp[1], lp[2]};
        q[F] -= 1;
        orc_rmap *rn0 = orc_get_refmap_pos(S, ind - strd, o0, q);
        if (!rn0) continue;
        double phiv0 = 0.5 * (rm_phi(S, rf0) + rm_phi(S, rn0));
        if (phiv0 > S->eps) continue;
        orc_rmap *r1p, *r1p1, *r1m, *r1m1, *r2p, *r2p1, *r2m, *r2m1;
        nbr_set(S, ind, o0, lp, F, ad1, ad2, &r1p, &r1p1, &r1m, &r1m1, &r2p, &r2p1, &r2m, &r2m1);
        if (!r1p || !r1p1 || !r1m || !r1m1 || !r2p || !r2p1 || !r2m || !r2m1) continue;
        double d0[3];
        d0[F] = facs[F] * (rm_phi(S, rf0) - rm_phi(S, rn0));
        d0[ad1] = 0.25 * facs[ad1] * (rm_phi(S, r1p) + rm_phi(S, r1p1) - rm_phi(S, r1m) - rm_phi(S, r1m1));
        d0[ad2] = 0.25 * facs[ad2] * (rm_phi(S, r2p) + rm_phi(S, r2p1) - rm_phi(S, r2m) - rm_phi(S, r2m1));
        double nr = 0;
        for (int a = 0; a < 3; a++) nr += d0[a] * d0[a];
        if (nr > 0) {
            nr = sqrt(nr);
            for (int a = 0; a < 3; a++) d0[a] /= nr;
        }
        for (int jj = ii + 1; jj < nobj; jj++) {
            orc_rmap *rf1 = co[jj];
            unsigned o1 = rf1->c & OMASK;
            orc_rmap *rn1 = orc_get_refmap(S, ind - strd, o1);
            if (!rn1) continue;
            double phiv1 = 0.5 * (rm_phi(S, rf1) + rm_phi(S, rn1));
            if (phiv1 > S->eps || phiv0 + phiv1 > S->eps) continue;
            orc_rmap *s1p, *s1p1, *s1m, *s1m1, *s2p, *s2p1, *s2m, *s2m1;
            nbr_set(S, ind, o1, lp, F, ad1, ad2, &s1p, &s1p1, &s1m, &s1m1, &s2p, &s2p1, &s2m, &s2m1);
            if (!s1p || !s1p1 || !s1m || !s1m1 || !s2p || !s2p1 || !s2m || !s2m1) continue;
            double d1[3];
            d1[F] = facs[F] * (rm_phi(S, rf1) - rm_phi(S, rn1));
            d1[ad1] = 0.25 * facs[ad1] * (rm_phi(S, s1p) + rm_phi(S, s1p1) - rm_phi(S, s1m) - rm_phi(S, s1m1));
            d1[ad2] = 0.25 * facs[ad2] * (rm_phi(S, s2p) + rm_phi(S, s2p1) - rm_phi(S, s2m) - rm_phi(S, s2m1));
            nr = 0;
            for (int a = 0; a < 3; a++) nr += d1[a] * d1[a];
            if (nr > 0) {
                nr = sqrt(nr);
                for (int a = 0; a < 3; a++) d1[a] /= nr;
            }
            double nv[3], den = 0;
            for (int a = 0; a < 3; a++) { nv[a] = d1[a] - d0[a]; den += nv[a] * nv[a]; }
            den = sqrt(den);
            if (den < SMALL_NUMBER) continue;
            for (int a = 0; a < 3; a++) nv[a] /= den;
            double G1 = S->sm[o1 - 1].G;
            double cf = G0 * coll_func(S, phiv0) + G1 * coll_func(S, phiv1);
            double cs[6];
            for (int a = 0; a < 3; a++)
                for (int b = 0; b < 3; b++) cs[symind(a, b)] = -cf * nv[a] * nv[b];
            double avg = (cs[0] + cs[3] + cs[5]) / 3.;
            for (int a = 0; a < 3; a++) cs[symind(a, a)] -= avg;
            for (int a = 0; a < 3; a++) fs[a] += cs[symind(F, a)];
#pragma omp atomic
            S->n_coll++;
        }
    }
}

/* collision_stress of one lower face alone (a test hook: fs is added to) */
void orc_collision_face(orc_sim *S, int eid, int F, double fs[3]) { collision_stress(S, eid, F, fs); }

/* fluid_3d::compute_stress (fluid_3d.cc:1360-1461); fluid_stress<F> is the
 * within_local_domain branch (fluid_3d.cc:2001-2031), always taken here */
void orc_compute_stress(orc_sim *S)
{
    int str[3] = {1, S->m4, S->mn4};
    double facs[3] = {S->dxsp, S->dysp, S->dzsp};
#pragma omp parallel for schedule(static)
    for (int k = -1; k <= S->o + 1; k++)
        for (int j = -1; j <= S->n + 1; j++)
            for (int i = -1; i <= S->m + 1; i++) {
                int eid = i + S->m4 * j + S->mn4 * k;
                orc_field *fp = S->u + S->G0 + eid;
                for (int F = 0; F < 3; F++) {
                    double ss[3] = {0, 0, 0}, fs[3] = {0, 0, 0}, sfrac = 0;
                    solid_stress(S, eid, F, ss, &sfrac);
                    if (sfrac >= 1) {
                        for (int s = 0; s < 3; s++) fp->sigma[F][s] = ss[s] / sfrac;
                    } else {
                        for (int s = 0; s < 3; s++) fs[s] = S->cfg.fmu * facs[F] * (fp->vel[s] - fp[-str[F]].vel[s]);
                        for (int s = 0; s < 3; s++) fp->sigma[F][s] = (1 - sfrac) * fs[s] + ss[s];
                    }
                    collision_stress(S, eid, F, fp->sigma[F]);
                }
            }
}

/* fluid_3d::compute_ustar (fluid_3d.cc:1309-1352), explicit viscosity */
static void compute_ustar(orc_sim *S)
{
    double cdt = S->dt;
#pragma omp parallel for schedule(static)
    for (int k = 0; k < S->o; k++)
        for (int j = 0; j < S->n; j++)
            for (int i = 0; i < S->m; i++) {
                int c = IX(S, i + 2, j + 2, k + 2);
                orc_field *f = S->u + c;
                double acc[3] = {0, 0, 0};
                acceleration(S, c, S->lx[i + 2], S->ly[j + 2], S->lz[k + 2], acc, 1);
                double av[3];
                av[0] = -0.5 * cdt * (f->fvel[0][0] + f->fvel[1][0]) * S->dxsp;
                av[1] = -0.5 * cdt * (f->fvel[2][1] + f->fvel[3][1]) * S->dysp;
                av[2] = -0.5 * cdt * (f->fvel[4][2] + f->fvel[5][2]) * S->dzsp;
                double dv[3];
                for (int d = 0; d < 3; d++) {
                    dv[d] = av[0] * (f->fvel[1][d] - f->fvel[0][d]) + av[1] * (f->fvel[3][d] - f->fvel[2][d]) +
                            av[2] * (f->fvel[5][d] - f->fvel[4][d]);
                    acc[d] *= cdt;
                }
                for (int d = 0; d < 3; d++) dv[d] += acc[d];
                for (int d = 0; d < 3; d++) {
                    f->dvel[d] = 0;
                    f->dvel[d] += dv[d];
                }
                for (int d = 0; d < 3; d++) f->vel[d] += f->dvel[d];
            }
}

/* ------------------------------------------------------------------ */
/* step                                                                 */
/* ------------------------------------------------------------------ */
/* fluid_3d::step_forward (fluid_3d.cc:429-577), no tracers, non-TANG */
int orc_step(orc_sim *S, int debug)
{
    (void)debug;
    S->nt++;
    S->n_coll = S->n_redraw_out = S->n_redraw_in = 0;
    if (S->trace) orc_tracers_predict(S);   /* fluid_3d.cc:458-461 */
    orc_compute_stress(S);
    normal_derivatives_velocity(S);
    if (normal_derivatives_refmap(S)) return RMT3D_ERR_FATAL;
    half_time_edge_velocities(S);
    half_time_edge_refmap(S);
    S->time += 0.5 * S->dt;
    orc_set_bc_cf(S);
    if (godunov_upwinding_set(S)) return RMT3D_ERR_FATAL;
    orc_set_bc_cf(S);
    if (mac_solve(S)) return RMT3D_ERR_FATAL;
    orc_set_bc_cf(S);
    int err = update_reference_map(S);
    if (err) {
        revert_reference_map_update(S);
        return RMT3D_REFMAP_UPDATE_ERR;
    }
    orc_compute_stress(S);
    compute_ustar(S);
    orc_set_bc_cc(S);
    orc_communicate(S, ORC_COMM_DVEL);
    S->time += 0.5 * S->dt;
    if (pp_solve(S)) return RMT3D_PROJECTION_ERR;
    update_reference_map_full(S);
    orc_redraw_boundary(S);
    orc_set_bc_cc(S);
    orc_communicate(S, ORC_COMM_DVEL);
    if (S->trace) orc_tracers_advect(S);    /* fluid_3d.cc:569-572 */
    return RMT3D_NORMAL_EXIT;
}

/* fluid_3d::div_u, L2 part (fluid_3d_utils.cc:507-544) */
double orc_div_u(orc_sim *S)
{
    int nn = S->fm * S->fn * S->fo;
    double *r = (double *)malloc(sizeof(double) * (size_t)nn);
    fill_pp_rhs(S, r);
    int id = 0;
    for (int k = 0; k < S->fo; k++) {
        double nk = (k == 0 || k == S->fo - 1) ? 0.5 : 1.0;
        if (S->prd[2]) nk = 1.0;
        for (int j = 0; j < S->fn; j++) {
            double nj = (j == 0 || j == S->fn - 1) ? 0.5 : 1.0;
            if (S->prd[1]) nj = 1.0;
            for (int i = 0; i < S->fm; i++) {
                double ni = (i == 0 || i == S->fm - 1) ? 0.5 : 1.0;
                if (S->prd[0]) ni = 1.0;
                double wt = ni * nj * nk;
                r[id] = wt * r[id] * r[id];
                id++;
            }
        }
    }
    double l2 = orc_blocked_sum(r, nn, 1.0);
    l2 /= nn;
    free(r);
    return sqrt(l2);
}

/* fluid_3d::init_iter (fluid_3d.cc:235-287) */
int orc_init_iter(orc_sim *S, int init_err)
{
    if (init_err > 0) {
        double tol = 1e-4;
        for (int it = 0; it < S->cfg.num_iters; it++) {
            S->time = 0;
            orc_set_bc_cc(S);
            int e = orc_step(S, S->cfg.debug_flag);
            if (e < 0) return e;
            double d = orc_div_u(S);
            init_fluid(S);
            if (init_refmap(S)) return RMT3D_ERR_FATAL;
            orc_band_reset(&S->ext, S->Ng);
            init_extrapolate(S);
            if (d < tol) break;
        }
        S->nt = 0;
        S->time = 0.0;
    }
    orc_compute_stress(S);
    return 0;
}

/* run_sim.cc:98-106 + stepper ctor (stepper.hh:211-225) + stepper::run's
 * set_dt (stepper.hh:336) */
int orc_prepare_run(orc_sim *S, int init_err, double *dt_out, int *total)
{
    double sdt = S->cfg.dt;
    if (S->dt_reg < sdt) sdt = S->dt_reg;
    S->cfg.dt = sdt;
    double df = S->cfg.T / (S->cfg.frames - 1);
    int steps = (int)(df / sdt);
    double rem = df - steps * sdt;
    if (rem > 0) sdt -= (sdt - rem) / (++steps);
    /* a resumed run has its first chk_num frames behind it (stepper.hh:
     * 214-216, 358-361) */
    int tot = (S->cfg.frames - 1 - S->cfg.chk_num) * steps;
    if (tot < 0) tot = 0;
    S->dt = sdt;
    S->cfg.dt = sdt;
    int e = orc_init_iter(S, init_err);
    if (e) return e;
    S->dt = sdt;
    if (dt_out) *dt_out = sdt;
    if (total) *total = (S->cfg.stop_short > 0 && S->cfg.stop_short < tot) ? S->cfg.stop_short : tot;
    return 0;
}
