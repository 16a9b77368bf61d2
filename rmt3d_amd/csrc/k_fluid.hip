// Fluid-stage kernels of the FSI step (FP64 parity path).
//
// One thread per cell, x fastest, so every field access of a wavefront is a
// unit-stride run of the SoA arrays.  Each kernel restates one reference
// stage; the citations give the reference lines whose arithmetic order it
// keeps.
#include <string.h>
#include "dev.h"
#include "launch.h"

namespace rmt {

__device__ __forceinline__ int IX(const Geo &g, int i, int j, int k) { return i + g.m4 * (j + g.n4 * k); }

// ---------------------------------------------------------------------------
// stresses (b): fluid_3d::compute_stress, solid_stress<F>, collision_stress<F>
// (fluid_3d.cc:1360-1461, 1465-1773, 1777-1997, 2001-2031)
// ---------------------------------------------------------------------------
__device__ __forceinline__ int get_pos(const Geo &g, const Maps &M, int ind, unsigned oid, int p0, int p1, int p2) {
    if (!(p0 >= 0 && p0 <= g.m4 - 1 && p1 >= 0 && p1 <= g.n4 - 1 && p2 >= 0 && p2 <= g.o4 - 1)) return -1;
    return M.get(ind, oid);
}

__device__ __forceinline__ double map_phi(const Maps &M, const ObjC *O, int nobj, int ref) {
    unsigned o = M.c(ref) & OMASK;
    if (o == 0) return __builtin_nan("");
    return obj_phi(O, nobj, (int)o - 1, M.x(ref, 0), M.x(ref, 1), M.x(ref, 2));
}

struct Nb8 { int r1p, r1p1, r1m, r1m1, r2p, r2p1, r2m, r2m1; };

// The face direction F is a template parameter throughout the stress path:
// every small array is then indexed by constants and stays in registers.
// get_pos for N cells at once: the same result as N calls, with the
// loads of each stage (primary ids, then band ranges) issued together so
// that the lookups overlap instead of running one dependent chain after
// another
template <int N>
__device__ __forceinline__ void get_pos_n(const Geo &g, const Maps &M, const int (&ind)[N], const bool (&inb)[N],
                                         unsigned oid, int (&out)[N]) {
    unsigned c[N];
#pragma unroll
    for (int k = 0; k < N; k++) {
        bool in = inb[k] && ind[k] >= 0 && ind[k] < M.Ng;
        out[k] = in ? -2 : -1;
        c[k] = in ? M.r.c[ind[k]] : 0u;
    }
    int s[N], e[N];
#pragma unroll
    for (int k = 0; k < N; k++) {
        if (out[k] == -2 && c[k] == oid) out[k] = ind[k];
        s[k] = e[k] = 0;
        if (out[k] == -2) {
            s[k] = M.b.start[ind[k]];
            e[k] = s[k] + M.b.cnt[ind[k]];
        }
    }
#pragma unroll
    for (int k = 0; k < N; k++) {
        if (out[k] != -2) continue;
        int r = -1;
        for (int a = s[k]; a < e[k]; a++)
            if ((M.b.c[a] & OMASK) == oid) { r = M.Ng + a; break; }
        out[k] = r;
    }
}

template <int F>
__device__ __forceinline__ bool nbr_set(const Geo &g, const Maps &M, int ind, unsigned oid, const int lp[3], Nb8 &n) {
    constexpr int ad1 = (F + 1) % 3, ad2 = (F + 2) % 3;
    const int str[3] = {1, g.m4, g.mn4};
    const int lim[3] = {g.m4 - 1, g.n4 - 1, g.o4 - 1};
    int strd = str[F], as1 = str[ad1], as2 = str[ad2];
    // the eight cells: +-1 along ad1 and ad2, each also one back along F
    const int da[8] = {ad1, ad1, ad1, ad1, ad2, ad2, ad2, ad2};
    const int va[8] = {1, 1, -1, -1, 1, 1, -1, -1};
    const int back[8] = {0, 1, 0, 1, 0, 1, 0, 1};
    int cell[8];
    bool inb[8];
#pragma unroll
    for (int k = 0; k < 8; k++) {
        int q[3] = {lp[0], lp[1], lp[2]};
        q[da[k]] += va[k];
        q[F] -= back[k];
        inb[k] = q[0] >= 0 && q[0] <= lim[0] && q[1] >= 0 && q[1] <= lim[1] && q[2] >= 0 && q[2] <= lim[2];
        cell[k] = ind + va[k] * (da[k] == ad1 ? as1 : as2) - back[k] * strd;
    }
    int r[8];
    get_pos_n<8>(g, M, cell, inb, oid, r);
    n.r1p = r[0]; n.r1p1 = r[1]; n.r1m = r[2]; n.r1m1 = r[3];
    n.r2p = r[4]; n.r2p1 = r[5]; n.r2m = r[6]; n.r2m1 = r[7];
    return n.r1p >= 0 && n.r1p1 >= 0 && n.r1m >= 0 && n.r1m1 >= 0 && n.r2p >= 0 && n.r2p1 >= 0 && n.r2m >= 0 &&
           n.r2m1 >= 0;
}

template <int F>
__device__ __forceinline__ void solid_stress(const Geo &g, const Phys &P, const ObjC *O, const Fluid &U, const Maps &M,
                                             int ind, double ss[3], double &sfrac) {
    const int str[3] = {1, g.m4, g.mn4};
    const double facs[3] = {g.dxsp, g.dysp, g.dzsp};
    constexpr int ad1 = (F + 1) % 3, ad2 = (F + 2) % 3;
    int strd = str[F];
    int lp[3];
    lp[2] = ind / g.mn4; lp[1] = (ind % g.mn4) / g.m4; lp[0] = ind % g.m4;
    int ne = M.b.cnt[ind], es = M.b.start[ind];
    for (int ii = -1; ii < ne; ii++) {
        int rf = (ii == -1) ? ind : g.Ng + es + ii;
        unsigned oid = M.c(rf) & OMASK;
        for (int a = 0; a < 3; a++) M.gphir(rf, 3 * F + a) = -1.e200;
        M.enr(rf, F) = -1.e200;
        M.detFr(rf, F) = -1.e200;
        if (oid == 0) continue;
        int q[3] = {lp[0], lp[1], lp[2]};
        q[F] -= 1;
        int rn = get_pos(g, M, ind - strd, oid, q[0], q[1], q[2]);
        if (rn < 0) continue;
        double phr = M.phi(rf), phn = M.phi(rn);
        double phiv = 0.5 * (phr + phn);
        if (phiv > P.eps) continue;
        double tf = heaviside(P, phiv);
        sfrac += tf;
        const ObjC &oc = O[oid - 1];
        double G = oc.G * tf;
        Nb8 n;
        if (!nbr_set<F>(g, M, ind, oid, lp, n)) continue;
        double gp[3];
        gp[F] = facs[F] * (phr - phn);
        gp[ad1] = 0.25 * facs[ad1] *
                  (M.phi(n.r1p) + M.phi(n.r1p1) - M.phi(n.r1m) -
                   M.phi(n.r1m1));
        gp[ad2] = 0.25 * facs[ad2] *
                  (M.phi(n.r2p) + M.phi(n.r2p1) - M.phi(n.r2m) -
                   M.phi(n.r2m1));
        {
            double nr = 0;
            for (int a = 0; a < 3; a++) nr += gp[a] * gp[a];
            nr = sqrt(nr);
            if (nr != 0.) {
                double ni = 1. / nr;
                for (int a = 0; a < 3; a++) gp[a] = ni * gp[a];
            }
        }
        for (int a = 0; a < 3; a++) M.gphir(rf, 3 * F + a) = gp[a];
        double dR[3][3];
        for (int a = 0; a < 3; a++) {
            dR[F][a] = facs[F] * (M.x(rf, a) - M.x(rn, a));
            dR[ad1][a] = 0.25 * facs[ad1] * (M.x(n.r1p, a) + M.x(n.r1p1, a) - M.x(n.r1m, a) - M.x(n.r1m1, a));
            dR[ad2][a] = 0.25 * facs[ad2] * (M.x(n.r2p, a) + M.x(n.r2p1, a) - M.x(n.r2m, a) - M.x(n.r2m1, a));
        }
        double v[9] = {dR[0][0], dR[1][0], dR[2][0], dR[0][1], dR[1][1], dR[2][1], dR[0][2], dR[1][2], dR[2][2]};
        // matrix::inverse / det (matlib.hh:303-329, 366-370)
        double d = v[0] * (v[4] * v[8] - v[5] * v[7]) + v[3] * (v[2] * v[7] - v[1] * v[8]) +
                   v[6] * (v[1] * v[5] - v[2] * v[4]);
        double Fd[9];
        for (int j = 0; j < 3; j++)
            for (int i = 0; i < 3; i++) {
                Fd[3 * j + i] = v[3 * ((i + 1) % 3) + (j + 1) % 3] * v[3 * ((i + 2) % 3) + (j + 2) % 3] -
                                v[3 * ((i + 1) % 3) + (j + 2) % 3] * v[3 * ((i + 2) % 3) + (j + 1) % 3];
                Fd[3 * j + i] /= d;
            }
        double detF = Fd[0] * (Fd[4] * Fd[8] - Fd[5] * Fd[7]) + Fd[3] * (Fd[2] * Fd[7] - Fd[1] * Fd[8]) +
                      Fd[6] * (Fd[1] * Fd[5] - Fd[2] * Fd[4]);
        M.detFr(rf, F) = detF;
        double sg[6], cg[6];
        for (int i = 0; i < 3; i++)
            for (int j = i; j < 3; j++) {
                double a = 0, b = 0;
                for (int k = 0; k < 3; k++) {
                    a += Fd[3 * i + k] * Fd[3 * j + k];
                    b += Fd[3 * k + i] * Fd[3 * k + j];
                }
                sg[symind(i, j)] = a;
                cg[symind(i, j)] = b;
            }
        M.enr(rf, F) = 0.5 * G * ((cg[0] + cg[3] + cg[5]) - 3 - 2 * det_log(detF));
        double savg = (sg[0] + sg[3] + sg[5]) / 3.0;
        for (int a = 0; a < 3; a++) sg[symind(a, a)] -= savg;
        double exmu = oc.ex_mu * tf * (1 + oc.ev_trans_mult * tderiv_func_in(P, phiv));
        for (int a = 0; a < 6; a++) sg[a] *= G;
        for (int a = 0; a < 3; a++) {
            ss[a] += sg[symind(F, a)];
            const double *va = U.vel[a];
            double dn = va[ind] - va[ind - strd];
            double tr;
            if (a == F) tr = dn;
            else {
                const double *vF = U.vel[F];
                tr = 0.25 * (vF[ind + str[a]] + vF[ind + str[a] - strd] - vF[ind - str[a]] - vF[ind - str[a] - strd]);
            }
            ss[a] += exmu * facs[F] * (dn + tr);
        }
    }
}

template <int F>
__device__ __forceinline__ void collision_stress(const Geo &g, const Phys &P, const ObjC *O, const Maps &M, int ind,
                                                 int eid, double fs[3]) {
    const int str[3] = {1, g.m4, g.mn4};
    const double facs[3] = {g.dxsp, g.dysp, g.dzsp};
    constexpr int ad1 = (F + 1) % 3, ad2 = (F + 2) % 3;
    int strd = str[F];
    int lp[3];
    // positions unpacked from eid, not G0+eid (fluid_3d.cc:1478)
    lp[2] = eid / g.mn4; lp[1] = (eid % g.mn4) / g.m4; lp[0] = eid % g.m4;
    int pobj = (M.r.c[ind] & OMASK) != 0;
    int eobj = M.b.cnt[ind], es = M.b.start[ind];
    int nobj = pobj + eobj;
    if (nobj <= 1) return;
    for (int ii = 0; ii < nobj - 1; ii++) {
        int rf0 = pobj ? (ii == 0 ? ind : g.Ng + es + ii - 1) : g.Ng + es + ii;
        unsigned o0 = M.c(rf0) & OMASK;
        double G0 = O[o0 - 1].G;
        int rn0 = get_pos(g, M, ind - strd, o0, lp[0] - (F == 0), lp[1] - (F == 1), lp[2] - (F == 2));
        if (rn0 < 0) continue;
        double phiv0 = 0.5 * (M.phi(rf0) + M.phi(rn0));
        if (phiv0 > P.eps) continue;
        Nb8 n;
        if (!nbr_set<F>(g, M, ind, o0, lp, n)) continue;
        double d0[3];
        d0[F] = facs[F] * (M.phi(rf0) - M.phi(rn0));
        d0[ad1] = 0.25 * facs[ad1] *
                  (M.phi(n.r1p) + M.phi(n.r1p1) - M.phi(n.r1m) -
                   M.phi(n.r1m1));
        d0[ad2] = 0.25 * facs[ad2] *
                  (M.phi(n.r2p) + M.phi(n.r2p1) - M.phi(n.r2m) -
                   M.phi(n.r2m1));
        double nr = 0;
        for (int a = 0; a < 3; a++) nr += d0[a] * d0[a];
        if (nr > 0) {
            nr = sqrt(nr);
            for (int a = 0; a < 3; a++) d0[a] /= nr;
        }
        for (int jj = ii + 1; jj < nobj; jj++) {
            int rf1 = pobj ? (jj == 0 ? ind : g.Ng + es + jj - 1) : g.Ng + es + jj;
            unsigned o1 = M.c(rf1) & OMASK;
            int rn1 = M.get(ind - strd, o1);
            if (rn1 < 0) continue;
            double phiv1 = 0.5 * (M.phi(rf1) + M.phi(rn1));
            if (phiv1 > P.eps || phiv0 + phiv1 > P.eps) continue;
            Nb8 s;
            if (!nbr_set<F>(g, M, ind, o1, lp, s)) continue;
            double d1[3];
            d1[F] = facs[F] * (M.phi(rf1) - M.phi(rn1));
            d1[ad1] = 0.25 * facs[ad1] *
                      (M.phi(s.r1p) + M.phi(s.r1p1) -
                       M.phi(s.r1m) - M.phi(s.r1m1));
            d1[ad2] = 0.25 * facs[ad2] *
                      (M.phi(s.r2p) + M.phi(s.r2p1) -
                       M.phi(s.r2m) - M.phi(s.r2m1));
            nr = 0;
            for (int a = 0; a < 3; a++) nr += d1[a] * d1[a];
            if (nr > 0) {
                nr = sqrt(nr);
                for (int a = 0; a < 3; a++) d1[a] /= nr;
            }
            double nv[3], den = 0;
            for (int a = 0; a < 3; a++) { nv[a] = d1[a] - d0[a]; den += nv[a] * nv[a]; }
            den = sqrt(den);
            if (den < 100. * 2.220446049250313e-16) continue;
            for (int a = 0; a < 3; a++) nv[a] /= den;
            double cf = G0 * coll_func(P, phiv0) + O[o1 - 1].G * coll_func(P, phiv1);
            double cs[6];
            for (int a = 0; a < 3; a++)
                for (int b = 0; b < 3; b++) cs[symind(a, b)] = -cf * nv[a] * nv[b];
            double avg = (cs[0] + cs[3] + cs[5]) / 3.;
            for (int a = 0; a < 3; a++) cs[symind(a, a)] -= avg;
            for (int a = 0; a < 3; a++) fs[a] += cs[symind(F, a)];
            atomicAdd(&P.counters[0], 1);
        }
    }
}

// compute_stress over eid in -1..s+1 (padded 1..s+3)
// The level-set value of every map for its object, once per stress pass
// (the stencils below read each about ten times): primary maps over the
// whole padded grid, then the band entries.
__global__ void k_phi_prim(Geo g, Phys P, const ObjC *O, RMap R) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= g.Ng) return;
    unsigned o = R.c[t] & OMASK;
    R.phi[t] = (o == 0) ? __builtin_nan("") : obj_phi(O, P.n_obj, (int)o - 1, R.x[0][t], R.x[1][t], R.x[2][t]);
}
__global__ void k_phi_band(Phys P, const ObjC *O, Band B) {
    int n = *B.n;
    for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < n; e += gridDim.x * blockDim.x) {
        unsigned o = B.c[e] & OMASK;
        B.phi[e] = (o == 0) ? __builtin_nan("") : obj_phi(O, P.n_obj, (int)o - 1, B.x[0][e], B.x[1][e], B.x[2][e]);
    }
}

// A cell takes the solid path when it holds a primary map of an object or a
// band entry; every other cell's lower-face stresses are the fluid's alone,
// mu du/dn, with the reference's blend (1 - sfrac) fs + ss at sfrac = ss = 0.
__device__ __forceinline__ bool stress_solid_cell(const Maps &M, int ind) {
    return (M.r.c[ind] != 0) || (M.b.cnt[ind] > 0);
}

// lower face F of a solid-path cell: every write is to this cell's sigma row F
// and to row F of its own maps' face fields, so faces run independently
template <int F>
__device__ __forceinline__ void stress_face(const Geo &g, const Phys &P, const ObjC *O, const Fluid &U, const Maps &M,
                                            int ind) {
    int eid = ind - g.G0;
    const int str[3] = {1, g.m4, g.mn4};
    const double facs[3] = {g.dxsp, g.dysp, g.dzsp};
    double ss[3] = {0, 0, 0}, sfrac = 0;
    solid_stress<F>(g, P, O, U, M, ind, ss, sfrac);
    double sig[3];
    if (sfrac >= 1) {
        for (int s = 0; s < 3; s++) sig[s] = ss[s] / sfrac;
    } else {
        for (int s = 0; s < 3; s++) {
            double fs = P.fmu * facs[F] * (U.vel[s][ind] - U.vel[s][ind - str[F]]);
            sig[s] = (1 - sfrac) * fs + ss[s];
        }
    }
    collision_stress<F>(g, P, O, M, ind, eid, sig);
    for (int s = 0; s < 3; s++) U.sigma[3 * F + s][ind] = sig[s];
}

// fluid cells of padded 1..s+3 in place; solid cells appended to list
// (one flat thread per cell, one atomic per workgroup for the appends)
__global__ __launch_bounds__(1024) void k_stress_fluid(Geo g, Phys P, Fluid U, Maps M, int *list, int *count) {
    __shared__ int sh[17];
    int w0 = g.m + 3, w1 = g.n + 3, w2 = g.o + 3;
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    bool in = t < w0 * w1 * w2;
    int ind = 0;
    if (in) {
        int i = t % w0 + 1, j = (t / w0) % w1 + 1, k = t / (w0 * w1) + 1;
        ind = IX(g, i, j, k);
    }
    bool solid = in && stress_solid_cell(M, ind);
    int slot = block_append(solid, count, sh);
    if (solid) list[slot] = ind;
    if (!in || solid) return;
    const int str[3] = {1, g.m4, g.mn4};
    const double facs[3] = {g.dxsp, g.dysp, g.dzsp};
    for (int F = 0; F < 3; F++)
        for (int s = 0; s < 3; s++) {
            double fs = P.fmu * facs[F] * (U.vel[s][ind] - U.vel[s][ind - str[F]]);
            U.sigma[3 * F + s][ind] = 1.0 * fs + 0.0;
        }
}

// one thread per solid cell, one launch per face direction; launched in
// 64-work-item workgroups, which the bound tells the compiler so that the
// per-cell state fits in registers instead of spilling to scratch
// the three faces in one launch, blockIdx.y = F (uniform per workgroup);
// the faces write disjoint fields, so they may run concurrently
__global__ __launch_bounds__(64) void k_stress_solid3(Geo g, Phys P, const ObjC *O, Fluid U, Maps M, const int *list,
                                                      const int *count) {
    int n = *count;
    const int F = blockIdx.y;
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < n; t += gridDim.x * blockDim.x) {
        if (F == 0) stress_face<0>(g, P, O, U, M, list[t]);
        else if (F == 1) stress_face<1>(g, P, O, U, M, list[t]);
        else stress_face<2>(g, P, O, U, M, list[t]);
    }
}

// ---------------------------------------------------------------------------
// advection (a): normal derivatives, acceleration, Godunov face extrapolation
// for the fluid and the reference maps (fluid_3d.cc:582-806, 980-1100,
// fields.hh:147-177, 572-592)
// ---------------------------------------------------------------------------
__device__ __forceinline__ void neg_pres_grad(const Geo &g, const double *pp, int c, double acc[3]) {
    double pr[8];
    for (int n = 0; n < 8; n++) pr[n] = pp[c + n % 2 + g.m4 * ((n / 2) % 2) + (n / 4) * g.mn4];
    double xa = 0, ya = 0, za = 0;
    for (int n = 0; n < 4; n++) {
        xa += pr[2 * n + 1] - pr[2 * n];
        int yi = (n / 2) * 4 + n % 2;
        ya += pr[yi + 2] - pr[yi];
        za += pr[n + 4] - pr[n];
    }
    xa *= -0.25 * g.dxsp; ya *= -0.25 * g.dysp; za *= -0.25 * g.dzsp;
    acc[0] = xa; acc[1] = ya; acc[2] = za;
}

__device__ __forceinline__ double delta_func_wall(double x) { return (x <= 0) ? 100 : (1 - x) / x; }
__device__ __forceinline__ double wall_trans(double phiv, double w, double winv) {
    if (phiv < -w) return 1;
    else if (phiv < w) return 0.5 * (1 - phiv * winv);
    return 0;
}

// sim_objects::solid_acceleration (sim_manager.cc:212-295), spheres
__device__ __forceinline__ void solid_acc(const Geo &g, const Phys &P, const ObjC *O, int oi, double r0, double r1, double r2, double x,
                          double y, double z, double f[3]) {
    f[0] = f[1] = f[2] = 0;
    if (P.sim_type != 2 && P.sim_type != 3) return;
    const ObjC &q = O[oi];
    double kr[3];
    for (int a = 0; a < 3; a++) {
        kr[a] = P.wall_acc_mult * q.avg_vel[a] * q.avg_vel[a];
        if (kr[a] < P.min_wall_acc) kr[a] = P.min_wall_acc;
    }
    double phiv = obj_phi(O, P.n_obj, oi, r0, r1, r2);
    double d;
    if (P.walls[0]) { d = x - P.wall_pos[0]; if (d <= P.wall_dist) f[0] += delta_func_wall(d * P.winv) * kr[0] * wall_trans(phiv, g.dx, g.dxsp); }
    if (P.walls[1]) { d = P.wall_pos[1] - x; if (d <= P.wall_dist) f[0] -= delta_func_wall(d * P.winv) * kr[0] * wall_trans(phiv, g.dx, g.dxsp); }
    if (P.walls[2]) { d = y - P.wall_pos[2]; if (d <= P.wall_dist) f[1] += delta_func_wall(d * P.winv) * kr[1] * wall_trans(phiv, g.dy, g.dysp); }
    if (P.walls[3]) { d = P.wall_pos[3] - y; if (d <= P.wall_dist) f[1] -= delta_func_wall(d * P.winv) * kr[1] * wall_trans(phiv, g.dy, g.dysp); }
    if (P.walls[4]) { d = z - P.wall_pos[4]; if (d <= P.wall_dist) f[2] += delta_func_wall(d * P.winv) * kr[2] * wall_trans(phiv, g.dz, g.dzsp); }
    if (P.walls[5]) { d = P.wall_pos[5] - z; if (d <= P.wall_dist) f[2] -= delta_func_wall(d * P.winv) * kr[2] * wall_trans(phiv, g.dz, g.dzsp); }
    double fn = sqrt(f[0] * f[0] + f[1] * f[1] + f[2] * f[2]);
    if (fn >= P.K_stiff) {
        double nf = P.K_stiff / fn;
        f[0] *= nf; f[1] *= nf; f[2] *= nf;
    }
    f[2] += P.gravity * heaviside(P, phiv);
}

__device__ __forceinline__ void fluid_acc(const Phys &P, double x, double y, double z, double t, double f[3]) {
    if (P.sim_type == 0 || P.sim_type == 3) {
        const double PI = 3.14159265358979323846, mu = P.fmu;
        f[0] = (12 * PI * det_cos(2 * PI * y) * det_cos(2 * PI * z) * (-6 * PI * mu * det_cos(2 * PI * t) + det_sin(2 * PI * t)) *
                    det_sin(2 * PI * x) +
                PI * det_cos(2 * PI * t) * det_cos(2 * PI * t) * (2 + det_cos(4 * PI * y) + det_cos(4 * PI * z)) * det_sin(4 * PI * x)) /
               4;
        f[1] = -(PI * det_cos(2 * PI * t) * det_cos(2 * PI * t) * (-1 + det_cos(4 * PI * x) - 2 * det_cos(4 * PI * z)) *
                 det_sin(4 * PI * y)) / 8;
        f[2] = -(PI * det_cos(2 * PI * t) * det_cos(2 * PI * t) * (-1 + det_cos(4 * PI * x) - 2 * det_cos(4 * PI * y)) *
                 det_sin(4 * PI * z)) / 8;
        return;
    }
    f[0] = f[1] = f[2] = 0;
    if (P.sim_type == 2) f[2] += P.gravity;
}

// fluid_3d::acceleration (fluid_3d.cc:980-1033).  FORCING: the sim type has
// the manufactured-solution forcing (ftest, mills); without it the fluid
// acceleration is (0, 0, gravity for objects), computed by the same
// operations, and the kernel does not carry the forcing's trigonometry.
// SOLID = false: the cell holds no map (no primary, no band entry), so the
// solid acceleration is zero; the same additions of zero are kept.
template <bool FORCING, bool SOLID = true>
__device__ __forceinline__ void acceleration(const Geo &g, const Phys &P, const ObjC *O, const Fluid &U, const Maps &M,
                                             int c, double x, double y, double z, double acc[3]) {
    for (int a = 0; a < 3; a++)
        acc[a] += g.dxsp * (U.sigma[a][c + 1] - U.sigma[a][c]) + g.dysp * (U.sigma[3 + a][c + g.m4] - U.sigma[3 + a][c]) +
                  g.dzsp * (U.sigma[6 + a][c + g.mn4] - U.sigma[6 + a][c]);
    double tmp[3];
    neg_pres_grad(g, U.p, c, tmp);
    for (int a = 0; a < 3; a++) acc[a] += tmp[a];
    tmp[0] = tmp[1] = tmp[2] = 0;
    double sfrac = 0;
    if (SOLID && (M.r.c[c] != 0 || M.b.cnt[c] > 0)) {
        for (int s = 0; s < P.n_obj; s++) {
            int rp = M.get(c, (unsigned)(s + 1));
            if (rp < 0) continue;
            double phiv = map_phi(M, O, P.n_obj, rp);
            if (phiv < P.eps) {
                double lf = heaviside(P, phiv);
                double an[3];
                solid_acc(g, P, O, s, M.x(rp, 0), M.x(rp, 1), M.x(rp, 2), x, y, z, an);
                for (int a = 0; a < 3; a++) tmp[a] += an[a];
                sfrac += lf;
            }
        }
    }
    if (sfrac > 1.)
        for (int a = 0; a < 3; a++) tmp[a] /= sfrac;
    for (int a = 0; a < 3; a++) acc[a] += tmp[a];
    if (sfrac < 1.) {
        double an[3];
        if (FORCING) {
            fluid_acc(P, x, y, z, P.time, an);
        } else {
            an[0] = an[1] = an[2] = 0;
            if (P.sim_type == 2) an[2] += P.gravity;
        }
        for (int a = 0; a < 3; a++) acc[a] += an[a] * (1 - sfrac);
    }
}

// normal derivatives of one map (fluid_3d.cc:650-691, 714-794); returns false
// if a neighbour map is missing
__device__ __forceinline__ bool refmap_nvd(const Geo &g, const Maps &M, int rf, int ind, int i, int j, int k, double nvd[3][3]) {
    const int str[3] = {1, g.m4, g.mn4};
    const double dhsp[3] = {g.dxsp, g.dysp, g.dzsp};
    unsigned oid = M.c(rf) & OMASK;
#pragma unroll
    for (int d = 0; d < 3; d++) {
        int base[3] = {i + 2, j + 2, k + 2};
        int m1p[3] = {base[0], base[1], base[2]}, p1p[3] = {base[0], base[1], base[2]};
        m1p[d] -= 1; p1p[d] += 1;
        int s = str[d];
        int m2 = get_pos(g, M, ind - 2 * s, oid, base[0], base[1], base[2]);
        int m1 = get_pos(g, M, ind - s, oid, m1p[0], m1p[1], m1p[2]);
        int p1 = get_pos(g, M, ind + s, oid, p1p[0], p1p[1], p1p[2]);
        int p2 = get_pos(g, M, ind + 2 * s, oid, base[0], base[1], base[2]);
        if (m2 < 0 || m1 < 0 || p1 < 0 || p2 < 0) return false;
        for (int v = 0; v < 3; v++)
            nvd[d][v] = dhsp[d] * mono_diff(M.x(m2, v), M.x(m1, v), M.x(rf, v), M.x(p1, v), M.x(p2, v));
    }
    return true;
}

// ref_map::godunov_extrap (fields.hh:572-592).  The 18 face states go to
// the primary or the band arrays through one branch on the map's kind, with
// constant component indices in each (a per-store choice between the two
// pointer tables made the compiler keep both tables in scratch memory).
template <typename FX>
__device__ __forceinline__ void refmap_faces_to(const Geo &g, const Phys &P, const Fluid &U, const FX &fx, int off,
                                                const double xv[3], int c, const double nvd[3][3]) {
    const double dh2[3] = {g.dx / 2, g.dy / 2, g.dz / 2};
    double dt2 = P.dt / 2.0;
#pragma unroll
    for (int fc = 0; fc < 6; fc++) {
        const int dim = fc / 2, a1 = (dim + 1) % 3, a2 = (dim + 2) % 3;
        double nv = U.vel[dim][c];
        double d = dh2[dim] * (2 * (fc % 2) - 1);
        double v1 = U.vel[a1][c], v2 = U.vel[a2][c];
#pragma unroll
        for (int b = 0; b < 3; b++)
            fx[3 * fc + b][off] = xv[b] + (d - dt2 * nv) * nvd[dim][b] - dt2 * (v1 * nvd[a1][b] + v2 * nvd[a2][b]);
    }
}
__device__ __forceinline__ void refmap_faces(const Geo &g, const Phys &P, const Maps &M, const Fluid &U, int rf, int c,
                             const double nvd[3][3]) {
    double xv[3] = {M.x(rf, 0), M.x(rf, 1), M.x(rf, 2)};
    if (rf < M.Ng) refmap_faces_to(g, P, U, M.r.fx, rf, xv, c, nvd);
    else refmap_faces_to(g, P, U, M.b.fx, rf - M.Ng, xv, c, nvd);
}

__device__ __forceinline__ bool abut_primary(const Geo &g, const Maps &M, unsigned o, int ind) {
    const int str[3] = {1, g.m4, g.mn4};
    for (int d = 0; d < 3; d++) {
        if ((M.r.c[ind + str[d]] & OMASK) == o) return true;
        if ((M.r.c[ind - str[d]] & OMASK) == o) return true;
    }
    return false;
}

// fluid face states of interior cell c (fluid_3d.cc:582-617, 980-1066)
template <bool FORCING, bool SOLID>
__device__ __forceinline__ void advect_cell(const Geo &g, const Phys &P, const ObjC *O, const Fluid &U, const Maps &M,
                                            int c, int i, int j, int k) {
    const int str[3] = {1, g.m4, g.mn4};
    const double dhsp[3] = {g.dxsp, g.dysp, g.dzsp};
    double vel[3] = {U.vel[0][c], U.vel[1][c], U.vel[2][c]};
    double nvd[3][3];
    for (int d = 0; d < 3; d++) {
        int s = str[d];
        for (int v = 0; v < 3; v++) {
            const double *a = U.vel[v];
            nvd[d][v] = dhsp[d] * mono_diff(a[c - 2 * s], a[c - s], vel[v], a[c + s], a[c + 2 * s]);
        }
    }
    double acc[3] = {0, 0, 0};
    acceleration<FORCING, SOLID>(g, P, O, U, M, c, g.lx[i + 2], g.ly[j + 2], g.lz[k + 2], acc);
    const double dh2[3] = {g.dx / 2, g.dy / 2, g.dz / 2};
    double dt2 = P.dt / 2.0;
    double adv[3] = {0, 0, 0};
    for (int a = 0; a < 3; a++) {
        for (int b = 0; b < 3; b++) adv[a] -= vel[b] * nvd[b][a];
        adv[a] *= dt2;
    }
    for (int fc = 0; fc < 6; fc++) {
        int dim = fc / 2;
        double d = dh2[dim] * (2 * (fc % 2) - 1);
        for (int b = 0; b < 3; b++) U.fvel[3 * fc + b][c] = vel[b] + d * nvd[dim][b] + adv[b] + dt2 * acc[b];
    }
}

// interior cells without maps (one thread per cell, flat index); the cells
// with maps are the solid-path list's (k_advect_solid)
template <bool FORCING>
__global__ void k_advect(Geo g, Phys P, const ObjC *O, Fluid U, Maps M) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= g.m * g.n * g.o) return;
    int i = t % g.m, j = (t / g.m) % g.n, k = t / (g.m * g.n);
    int c = IX(g, i + 2, j + 2, k + 2);
    if (stress_solid_cell(M, c)) return;
    advect_cell<FORCING, false>(g, P, O, U, M, c, i, j, k);
}

// interior cells of the solid-path list: fluid face states with the solid
// acceleration, then their maps' normal derivatives and face states
// (normal_derivatives_refmap + compute_half_time_edge_refmap), which read
// only the cell-centred velocity
// The solid cells' advection runs as two launches over the same list: the
// velocity face states (advect_cell) and the reference-map face states of
// the cell's primary and band maps.  They write disjoint fields and read
// none the other writes; apart, each needs about half the registers of the
// combined cell, so more wavefronts hide the latency of the map lookups.
template <bool FORCING>
__global__ __launch_bounds__(64) void k_advect_solid(Geo g, Phys P, const ObjC *O, Fluid U, Maps M, const int *list, const int *count,
                               int *status) {
    int n = *count;
    (void)status;
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < n; t += gridDim.x * blockDim.x) {
        int c = list[t];
        int i = c % g.m4 - 2, j = (c / g.m4) % g.n4 - 2, k = c / g.mn4 - 2;
        if (i < 0 || i >= g.m || j < 0 || j >= g.n || k < 0 || k >= g.o) continue;
        advect_cell<FORCING, true>(g, P, O, U, M, c, i, j, k);
    }
}

__global__ __launch_bounds__(64) void k_advect_solid_maps(Geo g, Phys P, Fluid U, Maps M, const int *list,
                                                          const int *count, int *status) {
    int n = *count;
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < n; t += gridDim.x * blockDim.x) {
        int c = list[t];
        int i = c % g.m4 - 2, j = (c / g.m4) % g.n4 - 2, k = c / g.mn4 - 2;
        if (i < 0 || i >= g.m || j < 0 || j >= g.n || k < 0 || k >= g.o) continue;
        double rn[3][3];
        if ((M.r.c[c] & OMASK) > 0) {
            if (!refmap_nvd(g, M, c, c, i, j, k, rn)) { atomicOr(status, ST_FATAL_NVD); continue; }
            refmap_faces(g, P, M, U, c, c, rn);
        }
        int es = M.b.start[c], ne = M.b.cnt[c];
        for (int a = 0; a < ne; a++) {
            int rf = g.Ng + es + a;
            if (abut_primary(g, M, M.b.c[es + a] & OMASK, c)) {
                if (!refmap_nvd(g, M, rf, c, i, j, k, rn)) { atomicOr(status, ST_FATAL_NVD); break; }
                refmap_faces(g, P, M, U, rf, c, rn);
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Godunov upwinding (fluid_3d.cc:811-944, 2156-2204): each cell owns its three
// upper faces, so one thread per cell writes disjoint face states
// ---------------------------------------------------------------------------
__device__ __forceinline__ void upwind_refmap(double nl, double nr, double &ml, double &mvl, double &mwl, double &mr,
                                              double &mvr, double &mwr) {
    if (nl + nr > 0 && nl > 0) {
        mr = ml; mvr = mvl; mwr = mwl;
    } else if (nl + nr < 0 && nr < 0) {
        ml = mr; mvl = mvr; mwl = mwr;
    } else {
        double t = 0.5 * (ml + mr);
        ml = mr = t;
        t = 0.5 * (mvl + mvr);
        mvl = mvr = t;
        t = 0.5 * (mwl + mwr);
        mwl = mwr = t;
    }
}

__device__ __forceinline__ bool upwind_map(const Geo &g, const Fluid &U, const Maps &M, int rf, int ind) {
    const int str[3] = {1, g.m4, g.mn4};
    unsigned oid = M.c(rf) & OMASK;
    for (int dir = 0; dir < 3; dir++) {
        int mn = M.get(ind + str[dir], oid);
        if (mn < 0) return false;
        int ff = 2 * dir + 1, nf = ff ^ 1;
        int a1 = (dir + 1) % 3, a2 = (dir + 2) % 3;
        double nl = U.fvel[3 * ff + dir][ind], nr = U.fvel[3 * nf + dir][ind + str[dir]];
        double ml = M.fxr(rf, 3 * ff + dir), mvl = M.fxr(rf, 3 * ff + a1), mwl = M.fxr(rf, 3 * ff + a2);
        double mr = M.fxr(mn, 3 * nf + dir), mvr = M.fxr(mn, 3 * nf + a1), mwr = M.fxr(mn, 3 * nf + a2);
        upwind_refmap(nl, nr, ml, mvl, mwl, mr, mvr, mwr);
        M.fxr(rf, 3 * ff + dir) = ml; M.fxr(rf, 3 * ff + a1) = mvl; M.fxr(rf, 3 * ff + a2) = mwl;
        M.fxr(mn, 3 * nf + dir) = mr; M.fxr(mn, 3 * nf + a1) = mvr; M.fxr(mn, 3 * nf + a2) = mwr;
    }
    return true;
}

// one work-item per cell of the (m+1) x (n+1) x (o+1) block from (1,1,1),
// x fastest (a flat grid: rows of m+1 cells would leave most of a third
// 64-wide workgroup per row idle)
__global__ __launch_bounds__(256) void k_upwind(Geo g, Fluid U, Maps M, int *status) {
    const int w = g.m + 1, h = g.n + 1;
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= w * h * (g.o + 1)) return;
    int row = t / w;
    int i = t - row * w + 1;
    int k = row / h;
    int j = row - k * h + 1;
    k += 1;
    int ind = IX(g, i, j, k);
    const int str[3] = {1, g.m4, g.mn4};
    // The fluid face states this work-item upwinds (its cell's three upper
    // faces and the matching lower faces of the neighbours) are written by
    // this work-item alone and not by the map upwinding below, which writes
    // map face states only: load all eighteen first, in one batch.
    double fl[3][3], fr[3][3];
#pragma unroll
    for (int dir = 0; dir < 3; dir++) {
        int ff = 2 * dir + 1, nf = ff ^ 1;
#pragma unroll
        for (int q = 0; q < 3; q++) {
            const int a = (dir + q) % 3;
            fl[dir][q] = U.fvel[3 * ff + a][ind];
            fr[dir][q] = U.fvel[3 * nf + a][ind + str[dir]];
        }
    }
    if ((M.r.c[ind] & OMASK) > 0)
        if (!upwind_map(g, U, M, ind, ind)) atomicOr(status, ST_FATAL_UPWIND);
    int es = M.b.start[ind], ne = M.b.cnt[ind];
    for (int a = 0; a < ne; a++) {
        unsigned o = M.b.c[es + a] & OMASK;
        for (int d = 0; d < 3; d++) {
            if ((M.r.c[ind + str[d]] & OMASK) == o) {
                if (!upwind_map(g, U, M, g.Ng + es + a, ind)) atomicOr(status, ST_FATAL_UPWIND);
                break;
            }
        }
    }
#pragma unroll
    for (int dir = 0; dir < 3; dir++) {
        int nb = ind + str[dir];
        int ff = 2 * dir + 1, nf = ff ^ 1;
        int a1 = (dir + 1) % 3, a2 = (dir + 2) % 3;
        double nl = fl[dir][0], vl = fl[dir][1], wl = fl[dir][2];
        double nr = fr[dir][0], vr = fr[dir][1], wr = fr[dir][2];
        if (nl + nr > 0 && nl > 0) {
            nr = nl; vr = vl; wr = wl;
        } else if (nl + nr < 0 && nr < 0) {
            nl = nr; vl = vr; wl = wr;
        } else {
            nl = nr = 0;
            double t = 0.5 * (vl + vr);
            vl = vr = t;
            t = 0.5 * (wl + wr);
            wl = wr = t;
        }
        U.fvel[3 * ff + dir][ind] = nl; U.fvel[3 * ff + a1][ind] = vl; U.fvel[3 * ff + a2][ind] = wl;
        U.fvel[3 * nf + dir][nb] = nr; U.fvel[3 * nf + a1][nb] = vr; U.fvel[3 * nf + a2][nb] = wr;
    }
}

// ---------------------------------------------------------------------------
// ghost fills: face_gm::set_ghost_cc / set_ghost_cf (fields.hh:765-849) for one
// wall face, and the periodic self-copy of communicate<flags>
// ---------------------------------------------------------------------------
struct FaceC {
    int f, n, sgn, bct[3], len[3], len4[3], u0_off, gcf0, gcf1;
    double bcv[3], h;
};

__device__ __forceinline__ int face_node(const Geo &g, const FaceC &F, int a, int b, int c) {
    return g.G0 + F.u0_off + a * F.len4[0] + b * F.len4[1] + c * F.len4[2];
}

// Several faces in one launch: work-items [start[q], start[q + 1]) serve
// face q.  The faces of one launch write disjoint cells and none reads what
// another writes (the caller groups them so), so the values are those of
// one launch per face in order.
constexpr int FACES_MAX = 6;
struct FaceSet {
    int nf;
    int start[FACES_MAX + 1];
    FaceC F[FACES_MAX];
};

__device__ __forceinline__ void bc_cc_at(const Geo &g, const FaceC &F, const Fluid &U, int t) {
    // normal index 1..2, tangential -2..len+1
    int ext[3];
    for (int d = 0; d < 3; d++) ext[d] = (d == F.n) ? 2 : F.len[d] + 4;
    int in[3];
    in[0] = t % ext[0]; in[1] = (t / ext[0]) % ext[1]; in[2] = t / (ext[0] * ext[1]);
    for (int d = 0; d < 3; d++) in[d] += (d == F.n) ? 1 : -2;
    int u = face_node(g, F, in[0], in[1], in[2]);
    int mi[3] = {in[0], in[1], in[2]};
    mi[F.n] = 1 - in[F.n];
    int um = face_node(g, F, mi[0], mi[1], mi[2]);
    for (int c = 0; c < 3; c++) {
        if (F.bct[c] == 1) U.vel[c][u] = 2 * F.bcv[c] - U.vel[c][um];
        else if (F.bct[c] == 2) {
            double dn = 2 * fabs(0.5 * F.h * (2 * in[F.n] - 1));
            U.vel[c][u] = F.sgn * dn * F.bcv[c] + U.vel[c][um];
        }
    }
}

__device__ __forceinline__ void bc_cf_at(const Geo &g, const FaceC &F, const Fluid &U, int t) {
    int ext[3];
    for (int d = 0; d < 3; d++) ext[d] = (d == F.n) ? 1 : F.len[d] + 4;
    int in[3];
    in[0] = t % ext[0]; in[1] = (t / ext[0]) % ext[1]; in[2] = t / (ext[0] * ext[1]);
    for (int d = 0; d < 3; d++) in[d] += (d == F.n) ? 1 : -2;
    int mi[3] = {in[0], in[1], in[2]};
    mi[F.n] = 0;
    int y0 = face_node(g, F, mi[0], mi[1], mi[2]);
    mi[F.n] = -1;
    int y1 = face_node(g, F, mi[0], mi[1], mi[2]);
    int u = face_node(g, F, in[0], in[1], in[2]);
    for (int c = 0; c < 3; c++) {
        double fb = F.bcv[c];
        if (F.bct[c] == 2) fb = (3 * F.h * F.bcv[c] * F.sgn + 9 * U.vel[c][y0] - U.vel[c][y1]) * 0.125;
        U.fvel[3 * F.gcf0 + c][y0] = fb;
        U.fvel[3 * F.gcf1 + c][u] = fb;
    }
}

template <bool CC>
__global__ void k_bc_faces(Geo g, FaceSet S, Fluid U) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= S.start[S.nf]) return;
    int q = 0;
    while (q + 1 < S.nf && t >= S.start[q + 1]) q++;
    if (CC) bc_cc_at(g, S.F[q], U, t - S.start[q]);
    else bc_cf_at(g, S.F[q], U, t - S.start[q]);
}

// periodic source (mpi_table region rule, Utilities/comm_info.hh:91-142)
__device__ __forceinline__ int periodic_src(const Geo &g, int i, int j, int k, int corner) {
    int c[3] = {i, j, k};
    int s[3];
    if (corner) { s[0] = g.fm; s[1] = g.fn; s[2] = g.fo; }
    else { s[0] = g.m; s[1] = g.n; s[2] = g.o; }
    const int per[3] = {g.m, g.n, g.o};
    bool any = false;
    for (int d = 0; d < 3; d++) {
        int t = (c[d] < 2) ? 0 : ((c[d] >= s[d] + 2) ? 2 : 1);
        if (t != 1) {
            if (!g.prd[d]) return -1;
            any = true;
            c[d] += (t == 0) ? per[d] : -per[d];
        }
    }
    if (!any) return -1;
    return c[0] + g.m4 * (c[1] + g.n4 * c[2]);
}

// a ghost cell of the slab decomposition: t indexes the six two-wide slabs
__device__ __forceinline__ bool slab_cell(const Geo &g, int t, int &i, int &j, int &k) {
    int sx = 2 * 2 * g.n4 * g.o4, sy = 2 * g.m4 * 2 * g.o4, sz = 2 * g.m4 * g.n4 * 2;
    if (t < sx) {
        int side = t / (2 * g.n4 * g.o4), r = t % (2 * g.n4 * g.o4);
        i = r % 2; j = (r / 2) % g.n4; k = r / (2 * g.n4);
        if (side) i += g.m + 2;
        return true;
    }
    t -= sx;
    if (t < sy) {
        int side = t / (g.m4 * 2 * g.o4), r = t % (g.m4 * 2 * g.o4);
        i = r % g.m4; j = (r / g.m4) % 2; k = r / (g.m4 * 2);
        if (side) j += g.n + 2;
        return true;
    }
    t -= sy;
    if (t < sz) {
        int side = t / (g.m4 * g.n4 * 2), r = t % (g.m4 * g.n4 * 2);
        i = r % g.m4; j = (r / g.m4) % g.n4; k = r / (g.m4 * g.n4);
        if (side) k += g.o + 2;
        return true;
    }
    return false;
}

struct CommArrays {
    double *d[40];
    uint32_t *u32[2];
    int nd, nu;
};

__global__ void k_periodic(Geo g, CommArrays A, int corner) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    int i, j, k;
    if (!slab_cell(g, t, i, j, k)) return;
    int src = periodic_src(g, i, j, k, corner);
    if (src < 0) return;
    int dst = IX(g, i, j, k);
    for (int a = 0; a < A.nd; a++) A.d[a][dst] = A.d[a][src];
    for (int a = 0; a < A.nu; a++) A.u32[a][dst] = A.u32[a][src];
}

// communicate_extrap_fx (fluid_3d_comm.cc:285-365): ghost entries take the
// face states of the source entry with the same c
__global__ void k_periodic_band_fx(Geo g, Band B) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    int i, j, k;
    if (!slab_cell(g, t, i, j, k)) return;
    int src = periodic_src(g, i, j, k, 0);
    if (src < 0) return;
    int dst = IX(g, i, j, k);
    for (int a = 0; a < B.cnt[src]; a++) {
        int es = B.start[src] + a;
        for (int b = 0; b < B.cnt[dst]; b++) {
            int ed = B.start[dst] + b;
            if (B.c[ed] == B.c[es])
                for (int q = 0; q < 18; q++) B.fx[q][ed] = B.fx[q][es];
        }
    }
}

// ---------------------------------------------------------------------------
// MAC projection pieces (fluid_3d.cc:2244-2304)
// ---------------------------------------------------------------------------
// interior cell (i, j, k) of work-item t of a flat grid over m x n x o,
// x fastest (one 256-wide workgroup spans two rows of 128 cells)
__device__ __forceinline__ bool flat_cell(const Geo &g, int &i, int &j, int &k) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= g.m * g.n * g.o) return false;
    const int row = t / g.m;
    i = t - row * g.m;
    k = row / g.n;
    j = row - k * g.n;
    return true;
}

__global__ void k_mac_rhs(Geo g, Fluid U, double *b, double fac) {
    int i, j, k;
    if (!flat_cell(g, i, j, k)) return;
    int c = IX(g, i + 2, j + 2, k + 2);
    double r = 0;
    for (int f = 0; f < 3; f++) r += fac * (U.fvel[3 * (2 * f + 1) + f][c] - U.fvel[3 * (2 * f) + f][c]);
    b[i + g.m * (j + g.n * k)] = r;
}

__global__ void k_mac_extract(Geo g, Fluid U, const double *x) {
    int i, j, k;
    if (!flat_cell(g, i, j, k)) return;
    U.q[IX(g, i + 2, j + 2, k + 2)] = x[i + g.m * (j + g.n * k)];
}

__global__ void k_sub_qgrad(Geo g, Fluid U, double f2) {
    int i, j, k;
    if (!flat_cell(g, i, j, k)) return;
    int c = IX(g, i + 2, j + 2, k + 2);
    const int str[3] = {1, g.m4, g.mn4};
    double q = U.q[c];
    for (int f = 0; f < 3; f++) {
        int s = str[f];
        double down = f2 * (q - U.q[c - s]);
        double up = f2 * (U.q[c + s] - q);
        U.fvel[3 * (2 * f) + f][c] -= down;
        U.fvel[3 * (2 * f + 1) + f][c] -= up;
    }
}

// ---------------------------------------------------------------------------
// intermediate velocity (fluid_3d.cc:1309-1352)
// ---------------------------------------------------------------------------
template <bool FORCING, bool SOLID>
__device__ __forceinline__ void ustar_cell(const Geo &g, const Phys &P, const ObjC *O, const Fluid &U, const Maps &M,
                                           int c, int i, int j, int k) {
    double acc[3] = {0, 0, 0};
    acceleration<FORCING, SOLID>(g, P, O, U, M, c, g.lx[i + 2], g.ly[j + 2], g.lz[k + 2], acc);
    double cdt = P.dt;
    double fv[18];
    for (int q = 0; q < 18; q++) fv[q] = U.fvel[q][c];
    double av[3];
    av[0] = -0.5 * cdt * (fv[0] + fv[3]) * g.dxsp;
    av[1] = -0.5 * cdt * (fv[7] + fv[10]) * g.dysp;
    av[2] = -0.5 * cdt * (fv[14] + fv[17]) * g.dzsp;
    double dv[3];
    for (int d = 0; d < 3; d++) {
        dv[d] = av[0] * (fv[3 + d] - fv[d]) + av[1] * (fv[9 + d] - fv[6 + d]) + av[2] * (fv[15 + d] - fv[12 + d]);
        acc[d] *= cdt;
    }
    for (int d = 0; d < 3; d++) dv[d] += acc[d];
    for (int d = 0; d < 3; d++) {
        double dd = 0;
        dd += dv[d];
        U.dvel[d][c] = dd;
        U.vel[d][c] += dd;
    }
}

// interior cells without maps; the others are the solid-path list's
template <bool FORCING>
__global__ void k_ustar(Geo g, Phys P, const ObjC *O, Fluid U, Maps M) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= g.m * g.n * g.o) return;
    int i = t % g.m, j = (t / g.m) % g.n, k = t / (g.m * g.n);
    int c = IX(g, i + 2, j + 2, k + 2);
    if (stress_solid_cell(M, c)) return;
    ustar_cell<FORCING, false>(g, P, O, U, M, c, i, j, k);
}

template <bool FORCING>
__global__ __launch_bounds__(64) void k_ustar_solid(Geo g, Phys P, const ObjC *O, Fluid U, Maps M, const int *list, const int *count) {
    int n = *count;
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < n; t += gridDim.x * blockDim.x) {
        int c = list[t];
        int i = c % g.m4 - 2, j = (c / g.m4) % g.n4 - 2, k = c / g.mn4 - 2;
        if (i < 0 || i >= g.m || j < 0 || j >= g.n || k < 0 || k >= g.o) continue;
        ustar_cell<FORCING, true>(g, P, O, U, M, c, i, j, k);
    }
}

// ---------------------------------------------------------------------------
// FEM projection pieces (fluid_3d.cc:2524-2753), gather form of the element
// scatter: each node sums its elements in the reference's element order
// ---------------------------------------------------------------------------
// (non-short-circuit forms, so that the tests compile to selects)
__device__ __forceinline__ bool within_domain(const Geo &g, int ei, int ej, int ek) {
    return (g.prd[0] | ((ei >= 0) & (ei <= g.m - 1))) & (g.prd[1] | ((ej >= 0) & (ej <= g.n - 1))) &
           (g.prd[2] | ((ek >= 0) & (ek <= g.o - 1)));
}
// validity of corner n of element (ei,ej,ek) (map_ref_to_local, fluid_3d.cc:2530-2545)
__device__ __forceinline__ bool corner_ok(const Geo &g, int ei, int ej, int ek, int n) {
    const int ci = n % 2, cj = (n / 2) % 2, ck = n / 4;
    return within_domain(g, ei, ej, ek) & (ei + ci <= g.fm - 1) & !((ei < 0) & (ci == 0)) &
           (ej + cj <= g.fn - 1) & !((ej < 0) & (cj == 0)) & (ek + ck <= g.fo - 1) & !((ek < 0) & (ck == 0));
}

struct GradPhi { double v[8][3]; };

struct Faces6 { FaceC f[6]; };

// quotient of non-negative integers from a float reciprocal, corrected by
// the exact remainder
__device__ __forceinline__ int fdivq(int x, int d, float rd) {
    int q = (int)((float)x * rd);
    int r = x - q * d;
    while (r < 0) { q--; r += d; }
    while (r >= d) { q++; r -= d; }
    return q;
}

// one thread per node (flat index over fm x fn x fo); rfm, rfn: 1/fm, 1/fn
__global__ void k_pp_rhs(Geo g, Fluid U, GradPhi gp, Faces6 faces, int nwall, double *r, float rfm, float rfn) {
    int tt = blockIdx.x * blockDim.x + threadIdx.x;
    if (tt >= g.fm * g.fn * g.fo) return;
    int row = fdivq(tt, g.fm, rfm);
    int a = tt - row * g.fm;
    int cz = fdivq(row, g.fn, rfn);
    int b = row - cz * g.fn;
    double acc = 0;
    // elements in (k, j, i) ascending order: corner n = 7..0.  Every
    // element index is inside the padded grid (ei, ej, ek >= -1 and at most
    // the last cell), so all 24 loads are issued unconditionally, in one
    // batch, and an element that is not a valid corner leaves acc as it is.
    double dv[8][3];
#pragma unroll
    for (int n = 7; n >= 0; n--) {
        int e = IX(g, a - n % 2 + 2, b - (n / 2) % 2 + 2, cz - n / 4 + 2);
#pragma unroll
        for (int l = 0; l < 3; l++) dv[n][l] = U.dvel[l][e];
    }
#pragma unroll
    for (int n = 7; n >= 0; n--) {
        int ei = a - n % 2, ej = b - (n / 2) % 2, ek = cz - n / 4;
        double val = 0;
#pragma unroll
        for (int l = 0; l < 3; l++) val += dv[n][l] * gp.v[n][l];
        acc = corner_ok(g, ei, ej, ek, n) ? acc + val : acc;
    }
    // boundary flux of the wall faces (fluid_3d.cc:2591-2645): node N on face s
    // receives from the elements E = N - corner whose normal coordinate is the
    // face's boundary cell, in the reference's element order (higher
    // tangential dimension outer).  A wall face holds only nodes of index 0
    // or the last in its normal dimension.
    bool edge = a == 0 || a == g.fm - 1 || b == 0 || b == g.fn - 1 || cz == 0 || cz == g.fo - 1;
    for (int w = 0; w < (edge ? nwall : 0); w++) {
        const FaceC &F = faces.f[w];
        int s = F.f, nd = F.n;
        int N[3] = {a, b, cz};
        int offn = (s % 2) ? ((nd == 0) ? g.m : (nd == 1) ? g.n : g.o) - 1 : 0;
        if (N[nd] != offn + s % 2) continue;
        int t0 = (nd == 0) ? 1 : 0, t1 = (nd == 2) ? 1 : 2;
        const int lens[3] = {g.m, g.n, g.o};
        for (int d1 = 1; d1 >= 0; d1--)
            for (int d0 = 1; d0 >= 0; d0--) {
                int E[3], cr[3];
                E[nd] = offn; cr[nd] = s % 2;
                E[t0] = N[t0] - d0; cr[t0] = d0;
                E[t1] = N[t1] - d1; cr[t1] = d1;
                if (E[t0] < -1 || E[t0] > lens[t0] || E[t1] < -1 || E[t1] > lens[t1]) continue;
                int id = cr[0] + 2 * cr[1] + 4 * cr[2];
                if (!corner_ok(g, E[0], E[1], E[2], id)) continue;
                int u = IX(g, E[0] + 2, E[1] + 2, E[2] + 2);
                acc += -U.fvel[3 * s + s / 2][u] * gp.v[id][s / 2];
            }
    }
    r[a + g.fm * (b + g.fn * cz)] = acc;
}

// one work-item per node on a flat grid (rows of fm = m + 1 nodes in
// 64-wide row launches left a nearly empty wavefront per row)
__global__ void k_pp_extract(Geo g, Fluid U, const double *x, const double *mean) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= g.fm * g.fn * g.fo) return;
    int row = t / g.fm, a = t - row * g.fm;
    int c = row / g.fn, b = row - c * g.fn;
    double v = x[t];
    double pm = *mean;
    int e = IX(g, a + 2, b + 2, c + 2);
    U.p[e] += (v - pm);
    U.dp[e] = (v - pm);
}

__global__ void k_sub_pgrad(Geo g, Fluid U, double dt) {
    int i, j, k;
    if (!flat_cell(g, i, j, k)) return;
    int c = IX(g, i + 2, j + 2, k + 2);
    double npg[3];
    neg_pres_grad(g, U.dp, c, npg);
    for (int a = 0; a < 3; a++) {
        npg[a] *= dt;
        U.vel[a][c] += npg[a];
        U.dvel[a][c] += npg[a];
    }
}

// weighted squares of the FEM rhs for div_u (fluid_3d_utils.cc:507-544)
__global__ void k_divu_terms(Geo g, const double *r, double *out) {
    int a = blockIdx.x * blockDim.x + threadIdx.x;
    int b = blockIdx.y, c = blockIdx.z;
    if (a >= g.fm) return;
    double nk = (c == 0 || c == g.fo - 1) ? 0.5 : 1.0;
    if (g.prd[2]) nk = 1.0;
    double nj = (b == 0 || b == g.fn - 1) ? 0.5 : 1.0;
    if (g.prd[1]) nj = 1.0;
    double ni = (a == 0 || a == g.fm - 1) ? 0.5 : 1.0;
    if (g.prd[0]) ni = 1.0;
    int id = a + g.fm * (b + g.fn * c);
    out[id] = ni * nj * nk * r[id] * r[id];
}

// ---------------------------------------------------------------------------
// initial conditions (fluid_3d.cc:292-400)
// ---------------------------------------------------------------------------
struct InitC {
    int has_exact;
    int vel_prof_num;
    double vel_profile[16];
    double eig[MAXOBJ > 64 ? 64 : MAXOBJ][3];
    double omega[MAXOBJ > 64 ? 64 : MAXOBJ][3];
    int type[MAXOBJ > 64 ? 64 : MAXOBJ];
};

__device__ __forceinline__ void ftest_vel(double x, double y, double z, double &u, double &v, double &w) {
    const double PI = 3.14159265358979323846;
    u = -det_sin(2 * PI * x) * det_cos(2 * PI * y) * det_cos(2 * PI * z);
    v = 0.5 * det_cos(2 * PI * x) * det_sin(2 * PI * y) * det_cos(2 * PI * z);
    w = 0.5 * det_cos(2 * PI * x) * det_cos(2 * PI * y) * det_sin(2 * PI * z);
}

__device__ __forceinline__ void obj_rm0(const InitC &I, const ObjC *O, int o, const double X[3], double xi[3]) {
    if (I.type[o] == 1)
        for (int a = 0; a < 3; a++) xi[a] = (X[a] - O[o].c[a]) * I.eig[o][a] + O[o].c[a];
    else
        for (int a = 0; a < 3; a++) xi[a] = X[a];
}

// init_fluid (fluid_3d.cc:292-318).  ZERO also clears every other field of
// the cell record; a context whose fields have not been written since their
// zero-filled allocation skips those stores (32 of 37 words per cell).
// vmax is the maximum of |u| over the cells
// the reference tracks.  Each wavefront reduces its lanes' values with
// shuffles (every lane reaches the reduction, out-of-range lanes with no
// value) and adds one atomicMax on the bit pattern (|u| >= 0, so the bits
// order like the values) to one of INIT_VMAX_SLOTS words 128 bytes apart,
// picked by the row; the host takes the maximum of the slots.
template <bool ZERO>
__global__ void k_init_fluid(Geo g, Phys P, const ObjC *O, InitC I, Fluid U, unsigned long long *vmax_slots) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    int j = blockIdx.y, k = blockIdx.z;
    bool has = false;
    double speed = 0;
    if (i < g.m4) {
        int c = IX(g, i, j, k);
        double x = g.lx[i], y = g.ly[j], z = g.lz[k];
        double u = 0, v = 0, w = 0;
        const double PI = 3.14159265358979323846;
        if (P.sim_type == 0 || P.sim_type == 3) {
            ftest_vel(x, y, z, u, v, w);
            speed = sqrt(u * u + v * v + w * w);
            has = true;
            if (I.has_exact) {
                u *= det_cos(2 * PI * P.time); v *= det_cos(2 * PI * P.time); w *= det_cos(2 * PI * P.time);
            }
        } else if (!I.has_exact) {
            for (int a = 0; a < I.vel_prof_num; a++) {
                int vt = (int)I.vel_profile[2 * a];
                double amp = I.vel_profile[2 * a + 1];
                if (vt == 1) u += amp;
                else if (vt == 2) v += amp;
                else if (vt == 3) w += amp;
            }
            double X[3] = {x, y, z};
            for (int o = 0; o < P.n_obj; o++) {
                double rx[3];
                obj_rm0(I, O, o, X, rx);
                if (obj_phi(O, P.n_obj, o, rx[0], rx[1], rx[2]) < 0.) {
                    double L[3];
                    for (int a = 0; a < 3; a++) L[a] = rx[a] - O[o].c[a];
                    u += I.omega[o][1] * L[2] - I.omega[o][2] * L[1];
                    v -= I.omega[o][0] * L[2] - I.omega[o][2] * L[0];
                    w += I.omega[o][0] * L[1] - I.omega[o][1] * L[0];
                }
            }
            speed = sqrt(u * u + v * v + w * w);
            has = true;
        }
        U.vel[0][c] = u; U.vel[1][c] = v; U.vel[2][c] = w;
        if (ZERO) {
            U.q[c] = 0; U.dp[c] = 0;
            for (int a = 0; a < 3; a++) U.dvel[a][c] = 0;
            for (int a = 0; a < 18; a++) U.fvel[a][c] = 0;
            for (int a = 0; a < 9; a++) U.sigma[a][c] = 0;
        }
    }
    double m = has ? speed : -1.0;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        double o = __shfl_xor(m, off, 64);
        m = o > m ? o : m;
    }
    int slot = (blockIdx.y + blockIdx.z * gridDim.y) % INIT_VMAX_SLOTS;
    if (__lane_id() == 0 && m >= 0)
        atomicMax(vmax_slots + 16 * slot, (unsigned long long)__double_as_longlong(m));
}

// init_pres (fluid_3d.cc:326-340)
__global__ void k_init_pres(Geo g, Phys P, Fluid U) {
    int i = blockIdx.x * blockDim.x + threadIdx.x + 2;
    int j = blockIdx.y + 2, k = blockIdx.z + 2;
    if (i >= g.m4 - 1) return;
    int c = IX(g, i, j, k);
    U.q[c] = 0;
    double p = 0;
    if (P.sim_type == 0 || P.sim_type == 3) {
        const double PI = 3.14159265358979323846;
        double x = g.lx[i] - g.dx * 0.5, y = g.ly[j] - g.dy * 0.5, z = g.lz[k] - g.dz * 0.5;
        p = 3 * PI * P.fmu * det_cos(2 * PI * x) * det_cos(2 * PI * y) * det_cos(2 * PI * z);
    }
    U.p[c] = p;
}

// init_refmap (fluid_3d.cc:348-394)
__global__ void k_init_refmap(Geo g, Phys P, const ObjC *O, InitC I, RMap R, int *status) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    int j = blockIdx.y, k = blockIdx.z;
    if (i >= g.m4) return;
    int c = IX(g, i, j, k);
    double xv[3] = {0, 0, 0};
    uint32_t cc = 0;
    bool interior = (i >= 2 && i < g.m4 - 2 && j >= 2 && j < g.n4 - 2 && k >= 2 && k < g.o4 - 2);
    if (interior) {
        double X[3] = {g.lx[i], g.ly[j], g.lz[k]};
        bool found = false;
        for (int o = 0; o < P.n_obj; o++) {
            obj_rm0(I, O, o, X, xv);
            if (obj_phi(O, P.n_obj, o, xv[0], xv[1], xv[2]) < 0) {
                if (!found) { cc = (uint32_t)(o + 1); found = true; }
                else atomicOr(status, ST_FATAL_INIT);
            }
        }
    }
    R.c[c] = cc;
    for (int a = 0; a < 3; a++) { R.x[a][c] = xv[a]; R.xpred[a][c] = xv[a]; }
    for (int a = 0; a < 18; a++) R.fx[a][c] = 0;
    for (int a = 0; a < 9; a++) R.gphi[a][c] = -1.e200;
    for (int a = 0; a < 3; a++) { R.detF[a][c] = 0; R.en[a][c] = 0; }
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
static inline dim3 rows(int nx, int ny, int nz, int bs) { return dim3((nx + bs - 1) / bs, ny, nz); }

__global__ void k_zero1(int *p) {
    if (threadIdx.x == 0) *p = 0;
}

void launch_stress(hipStream_t s, const Geo &g, const Phys &P, const ObjC *O, const Fluid &U, const Maps &M,
                   int *list, int *count) {
    hipLaunchKernelGGL(k_phi_prim, dim3((g.Ng + 255) / 256), dim3(256), 0, s, g, P, O, M.r);
    hipLaunchKernelGGL(k_phi_band, dim3(1024), dim3(256), 0, s, P, O, M.b);
    hipLaunchKernelGGL(k_zero1, dim3(1), dim3(64), 0, s, count);
    int ncell = (g.m + 3) * (g.n + 3) * (g.o + 3);
    hipLaunchKernelGGL(k_stress_fluid, dim3((ncell + 1023) / 1024), dim3(1024), 0, s, g, P, U, M, list, count);
    hipLaunchKernelGGL(k_stress_solid3, dim3(4096, 3), dim3(64), 0, s, g, P, O, U, M, list, count);
}
void launch_advect(hipStream_t s, const Geo &g, const Phys &P, const ObjC *O, const Fluid &U, const Maps &M,
                   const int *list, const int *count, int *status) {
    int n = g.m * g.n * g.o;
    if (P.sim_type == 0 || P.sim_type == 3) {
        hipLaunchKernelGGL(k_advect<true>, dim3((n + 255) / 256), dim3(256), 0, s, g, P, O, U, M);
        hipLaunchKernelGGL(k_advect_solid<true>, dim3(2048), dim3(64), 0, s, g, P, O, U, M, list, count, status);
    } else {
        hipLaunchKernelGGL(k_advect<false>, dim3((n + 255) / 256), dim3(256), 0, s, g, P, O, U, M);
        hipLaunchKernelGGL(k_advect_solid<false>, dim3(2048), dim3(64), 0, s, g, P, O, U, M, list, count, status);
    }
    hipLaunchKernelGGL(k_advect_solid_maps, dim3(2048), dim3(64), 0, s, g, P, U, M, list, count, status);
}
void launch_upwind(hipStream_t s, const Geo &g, const Fluid &U, const Maps &M, int *status) {
    int nup = (g.m + 1) * (g.n + 1) * (g.o + 1);
    hipLaunchKernelGGL(k_upwind, dim3((nup + 255) / 256), dim3(256), 0, s, g, U, M, status);
}
static FaceC make_face(const Geo &g, const FaceSpec &f) {
    FaceC F;
    F.f = f.f; F.n = f.f / 2; F.sgn = (f.f % 2 == 0) ? -1 : 1;
    int dims[3] = {g.m, g.n, g.o};
    for (int c = 0; c < 3; c++) { F.bct[c] = f.bct[c]; F.bcv[c] = f.bcv[c]; F.len[c] = dims[c]; }
    F.len4[0] = 1; F.len4[1] = g.m4; F.len4[2] = g.mn4;
    F.u0_off = (f.f % 2 == 0) ? 0 : (F.len[F.n] - 1) * F.len4[F.n];
    F.len[F.n] = 1;
    if (f.f % 2 == 0) F.len4[F.n] *= -1;
    F.gcf0 = f.f;
    F.gcf1 = 2 * F.n + (F.sgn < 0);
    const double hs[3] = {g.dx, g.dy, g.dz};
    F.h = hs[F.n];
    return F;
}
static int face_items(const FaceC &F, int depth) {
    int tot = 1;
    for (int d = 0; d < 3; d++) tot *= (d == F.n) ? depth : F.len[d] + 4;
    return tot;
}
static void launch_faces(hipStream_t s, const Geo &g, const FaceSpec *f, int nf, const Fluid &U, bool cc) {
    if (nf <= 0) return;
    FaceSet S{};
    S.nf = nf;
    S.start[0] = 0;
    for (int q = 0; q < nf; q++) {
        S.F[q] = make_face(g, f[q]);
        S.start[q + 1] = S.start[q] + face_items(S.F[q], cc ? 2 : 1);
    }
    int tot = S.start[nf];
    if (cc) hipLaunchKernelGGL(k_bc_faces<true>, dim3((tot + 255) / 256), dim3(256), 0, s, g, S, U);
    else hipLaunchKernelGGL(k_bc_faces<false>, dim3((tot + 255) / 256), dim3(256), 0, s, g, S, U);
}
// set_ghost_cc over the wall faces in order.  A face's mirror cells include
// the ghost rows of the other dimensions, which later faces write, so faces
// of different dimensions keep their order (one launch per dimension); the
// two faces of one dimension write disjoint ghost layers and read cells the
// other does not write (the extent is at least 4), so they share a launch.
void launch_bc_cc_faces(hipStream_t s, const Geo &g, const FaceSpec *f, int nf, const Fluid &U) {
    const int dims[3] = {g.m, g.n, g.o};
    for (int q = 0; q < nf;) {
        int r = q + 1;
        const int n = f[q].f / 2;
        if (dims[n] >= 4)
            while (r < nf && f[r].f / 2 == n) r++;
        launch_faces(s, g, f + q, r - q, U, true);
        q = r;
    }
}
// set_ghost_cf over the wall faces: each face writes its own two face-state
// components (its own and its neighbour's) in disjoint cells and reads only
// cell-centred velocities, so all faces share one launch
void launch_bc_cf_faces(hipStream_t s, const Geo &g, const FaceSpec *f, int nf, const Fluid &U) {
    for (int q = 0; q < nf; q += FACES_MAX) launch_faces(s, g, f + q, std::min(FACES_MAX, nf - q), U, false);
}
static int slab_total(const Geo &g) { return 2 * 2 * g.n4 * g.o4 + 2 * g.m4 * 2 * g.o4 + 2 * g.m4 * g.n4 * 2; }
void launch_periodic(hipStream_t s, const Geo &g, double *const *d, int nd, uint32_t *const *u, int nu, int corner) {
    if (!(g.prd[0] || g.prd[1] || g.prd[2])) return;
    CommArrays A;
    A.nd = nd; A.nu = nu;
    for (int a = 0; a < nd; a++) A.d[a] = d[a];
    for (int a = 0; a < nu; a++) A.u32[a] = u[a];
    int tot = slab_total(g);
    hipLaunchKernelGGL(k_periodic, dim3((tot + 255) / 256), dim3(256), 0, s, g, A, corner);
}
void launch_periodic_band_fx(hipStream_t s, const Geo &g, const Band &B) {
    if (!(g.prd[0] || g.prd[1] || g.prd[2])) return;
    int tot = slab_total(g);
    hipLaunchKernelGGL(k_periodic_band_fx, dim3((tot + 255) / 256), dim3(256), 0, s, g, B);
}
void launch_mac_rhs(hipStream_t s, const Geo &g, const Fluid &U, double *b, double fac) {
    hipLaunchKernelGGL(k_mac_rhs, dim3((g.m * g.n * g.o + 255) / 256), dim3(256), 0, s, g, U, b, fac);
}
void launch_mac_extract(hipStream_t s, const Geo &g, const Fluid &U, const double *x) {
    hipLaunchKernelGGL(k_mac_extract, dim3((g.m * g.n * g.o + 255) / 256), dim3(256), 0, s, g, U, x);
}
void launch_sub_qgrad(hipStream_t s, const Geo &g, const Fluid &U, double f2) {
    hipLaunchKernelGGL(k_sub_qgrad, dim3((g.m * g.n * g.o + 255) / 256), dim3(256), 0, s, g, U, f2);
}
void launch_ustar(hipStream_t s, const Geo &g, const Phys &P, const ObjC *O, const Fluid &U, const Maps &M,
                  const int *list, const int *count) {
    int n = g.m * g.n * g.o;
    if (P.sim_type == 0 || P.sim_type == 3) {
        hipLaunchKernelGGL(k_ustar<true>, dim3((n + 255) / 256), dim3(256), 0, s, g, P, O, U, M);
        hipLaunchKernelGGL(k_ustar_solid<true>, dim3(2048), dim3(64), 0, s, g, P, O, U, M, list, count);
    } else {
        hipLaunchKernelGGL(k_ustar<false>, dim3((n + 255) / 256), dim3(256), 0, s, g, P, O, U, M);
        hipLaunchKernelGGL(k_ustar_solid<false>, dim3(2048), dim3(64), 0, s, g, P, O, U, M, list, count);
    }
}
void launch_pp_rhs(hipStream_t s, const Geo &g, const Fluid &U, const double gradphi[8][3], const FaceSpec *faces,
                   int nwall, void *faces_dev, double *r) {
    GradPhi gp;
    for (int a = 0; a < 8; a++)
        for (int b = 0; b < 3; b++) gp.v[a][b] = gradphi[a][b];
    (void)faces_dev;
    Faces6 fc;
    memset(&fc, 0, sizeof fc);
    for (int w = 0; w < nwall && w < 6; w++) fc.f[w] = make_face(g, faces[w]);
    int nn = g.fm * g.fn * g.fo;
    hipLaunchKernelGGL(k_pp_rhs, dim3((nn + 255) / 256), dim3(256), 0, s, g, U, gp, fc, nwall, r, 1.0f / g.fm,
                       1.0f / g.fn);
}
size_t face_struct_size() { return sizeof(FaceC) * 6; }
void launch_pp_extract(hipStream_t s, const Geo &g, const Fluid &U, const double *x, const double *mean) {
    int n = g.fm * g.fn * g.fo;
    hipLaunchKernelGGL(k_pp_extract, dim3((n + 255) / 256), dim3(256), 0, s, g, U, x, mean);
}
void launch_sub_pgrad(hipStream_t s, const Geo &g, const Fluid &U, double dt) {
    hipLaunchKernelGGL(k_sub_pgrad, dim3((g.m * g.n * g.o + 255) / 256), dim3(256), 0, s, g, U, dt);
}
void launch_divu_terms(hipStream_t s, const Geo &g, const double *r, double *out) {
    hipLaunchKernelGGL(k_divu_terms, rows(g.fm, g.fn, g.fo, 64), dim3(64), 0, s, g, r, out);
}
void launch_init_fluid(hipStream_t s, const Geo &g, const Phys &P, const ObjC *O, const InitSpec &is, const Fluid &U,
                       unsigned long long *vmax_bits, int zero_rest) {
    InitC I;
    I.has_exact = is.has_exact;
    I.vel_prof_num = is.vel_prof_num;
    for (int a = 0; a < 16; a++) I.vel_profile[a] = is.vel_profile[a];
    int no = P.n_obj < 64 ? P.n_obj : 64;
    for (int o = 0; o < no; o++) {
        I.type[o] = is.type[o];
        for (int a = 0; a < 3; a++) { I.eig[o][a] = is.eig[o][a]; I.omega[o][a] = is.omega[o][a]; }
    }
    if (zero_rest)
        hipLaunchKernelGGL(k_init_fluid<true>, rows(g.m4, g.n4, g.o4, 64), dim3(64), 0, s, g, P, O, I, U, vmax_bits);
    else
        hipLaunchKernelGGL(k_init_fluid<false>, rows(g.m4, g.n4, g.o4, 64), dim3(64), 0, s, g, P, O, I, U, vmax_bits);
}
void launch_init_pres(hipStream_t s, const Geo &g, const Phys &P, const Fluid &U) {
    hipLaunchKernelGGL(k_init_pres, rows(g.m4 - 3, g.n4 - 3, g.o4 - 3, 64), dim3(64), 0, s, g, P, U);
}
void launch_init_refmap(hipStream_t s, const Geo &g, const Phys &P, const ObjC *O, const InitSpec &is, const RMap &R,
                        int *status) {
    InitC I;
    I.has_exact = is.has_exact;
    I.vel_prof_num = 0;
    int no = P.n_obj < 64 ? P.n_obj : 64;
    for (int o = 0; o < no; o++) {
        I.type[o] = is.type[o];
        for (int a = 0; a < 3; a++) { I.eig[o][a] = is.eig[o][a]; I.omega[o][a] = is.omega[o][a]; }
    }
    hipLaunchKernelGGL(k_init_refmap, rows(g.m4, g.n4, g.o4, 64), dim3(64), 0, s, g, P, O, I, R, status);
}

}  // namespace rmt
