// Tracer particles (tracers::predict / advect, Utilities/tracers.hh:221-300,
// tracers.cc:265-316) on the device, one work-item per tracer.  A tracer is
// its 4x4x4 node cube (lowest padded node p) and its position rp in
// [-0.5, 0.5]^3 relative to the cube centre, in units of the cell size.
// The velocity at rp is the tricubic interpolant of the cube's 64 cell
// velocities, evaluated as a tensor product of cubic Lagrange weights on the
// nodes -1.5, -0.5, 0.5, 1.5 (the same polynomial the reference builds from
// a precomputed inverse Vandermonde matrix).  On one rank the MPI hand-off
// of a tracer that leaves the cube range reduces to a periodic wrap, or to
// the tracer being dropped where no neighbour exists.
#include "launch.h"

namespace rmt {

namespace {

__device__ __forceinline__ void lagrange4(double r, double w[4]) {
    w[0] = -(r + 0.5) * (r - 0.5) * (r - 1.5) / 6.0;
    w[1] = (r + 1.5) * (r - 0.5) * (r - 1.5) / 2.0;
    w[2] = -(r + 1.5) * (r + 0.5) * (r - 1.5) / 2.0;
    w[3] = (r + 1.5) * (r + 0.5) * (r - 0.5) / 6.0;
}

// tracers::interp_vels (tracers.cc:395-455): the cube index is clamped to
// the grid as there
__device__ void interp(const TrParams &Q, const Fluid &U, const int p[3], const double r[3], double v[3]) {
    int ti = min(max(p[0], 0), Q.lim[0] - 4), tj = min(max(p[1], 0), Q.lim[1] - 4),
        tk = min(max(p[2], 0), Q.lim[2] - 4);
    double wx[4], wy[4], wz[4];
    lagrange4(r[0], wx);
    lagrange4(r[1], wy);
    lagrange4(r[2], wz);
    for (int f = 0; f < 3; f++) {
        const double *u = U.vel[f];
        double sk = 0;
        for (int k = 0; k < 4; k++) {
            double sj = 0;
            for (int j = 0; j < 4; j++) {
                const double *row = u + ti + Q.lim[0] * ((tj + j) + Q.lim[1] * (tk + k));
                double si = 0;
                for (int i = 0; i < 4; i++) si += wx[i] * row[i];
                sj += wy[j] * si;
            }
            sk += wz[k] * sj;
        }
        v[f] = sk;
    }
}

// tracers::dx (tracers.hh:331-442): the displacement dt v in direction f,
// slowed near a homogeneous Dirichlet wall so that a tracer never comes
// within 1 % of the box width of it; x is the tracer's coordinate
__device__ double disp(const TrParams &Q, int f, double v, double x, double dt) {
    double update = dt * v;
    int lr = (v < 0) ? 0 : 1;
    if (!(Q.bct[f][lr] == 1 && Q.bcv[f][lr] == 0)) return update;
    double lo = Q.lo[f], hi = Q.hi[f];
    const double min_sep = 0.01 * (hi - lo);
    double sep, vn;
    if (v == 0) {
        double v1 = fabs(x - lo), v2 = fabs(hi - x);
        sep = v2 < v1 ? (hi - x) : (x - lo);
    } else {
        sep = (v < 0) ? (x - lo) : (hi - x);
    }
    vn = fabs(v);
    if (sep <= min_sep) {
        if (v == 0) {
            double v1 = fabs(x - lo), v2 = fabs(hi - x);
            return (v2 < v1) ? (hi - min_sep - x) : (lo + min_sep - x);
        }
        return (v < 0) ? (lo + min_sep - x) : (hi - min_sep - x);
    }
    if (fabs(update) > 0.1 * sep) update = (det_exp(-dt * vn / sep) - 1) * ((v < 0) ? sep : -sep);
    if ((v < 0 && (x + update) < (lo + min_sep)) || (v > 0 && (x + update) > (hi - min_sep)))
        return (v < 0) ? (lo + min_sep - x) : (hi - min_sep - x);
    return update;
}

// tracers::proc_adjust (tracers.cc:259-297) with the one-rank hand-off of
// communicate / empty_buffer (:220-257, 327-352): st 0 tests rpf, st 1 rp0.
// Returns false when the tracer leaves through a dimension without a
// periodic neighbour.
__device__ bool adjust(const TrParams &Q, int p[3], double rp0[3], double rpf[3], int st) {
    int comm[3];
    for (int f = 0; f < 3; f++) {
        double *val = st == 0 ? &rpf[f] : &rp0[f];
        while (*val <= -0.5) { rp0[f] += 1.0; rpf[f] += 1.0; p[f]--; }
        while (*val > 0.5) { rp0[f] -= 1.0; rpf[f] -= 1.0; p[f]++; }
        bool low = (p[f] < 0) || (p[f] == 0 && *val < 0);
        bool high = (p[f] > Q.lim[f] - 4) || (p[f] == Q.lim[f] - 4 && *val >= 0);
        if (low) comm[f] = 0;
        else if (high) { comm[f] = 2; p[f] -= Q.lim[f] - 4; }
        else comm[f] = 1;
    }
    bool moved = false;
    for (int f = 0; f < 3; f++)
        if (comm[f] != 1) {
            moved = true;
            if (!Q.prd[f]) return false;
        }
    if (moved)
        for (int f = 0; f < 3; f++) {
            double val = st == 0 ? rpf[f] : rp0[f];
            if (p[f] < 0 || (p[f] == 0 && val < 0)) p[f] += Q.lim[f] - 4;
        }
    return true;
}

__device__ __forceinline__ double coord(const TrParams &Q, int f, int p, double rp) {
    return Q.low[f] + (1.5 + rp + p) * Q.h[f];
}

}  // namespace

// tracers::predict (tracers.hh:192-265): forward Euler to rpf
__global__ void k_tracer_predict(TrParams Q, TracerDev T, Fluid U, double dt) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= T.n || !T.alive[t]) return;
    int p[3];
    double rp0[3], rpf[3], v0[3];
    for (int f = 0; f < 3; f++) { p[f] = T.p[f][t]; rp0[f] = T.rp0[f][t]; rpf[f] = T.rpf[f][t]; }
    interp(Q, U, p, rp0, v0);
    for (int f = 0; f < 3; f++) {
        double dr = disp(Q, f, v0[f], coord(Q, f, p[f], rp0[f]), dt) / Q.h[f];
        rpf[f] = rp0[f] + dr;
    }
    bool ok = adjust(Q, p, rp0, rpf, 0);
    for (int f = 0; f < 3; f++) {
        T.p[f][t] = p[f]; T.rp0[f][t] = rp0[f]; T.rpf[f][t] = rpf[f]; T.v0[f][t] = v0[f];
    }
    if (!ok) T.alive[t] = 0;
}

// tracers::imp_euler / advect (tracers.hh:266-300): improved Euler from the
// predicted position's velocity
__global__ void k_tracer_advect(TrParams Q, TracerDev T, Fluid U, double dt) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= T.n || !T.alive[t]) return;
    int p[3];
    double rp0[3], rpf[3], vf[3];
    for (int f = 0; f < 3; f++) { p[f] = T.p[f][t]; rp0[f] = T.rp0[f][t]; rpf[f] = T.rpf[f][t]; }
    interp(Q, U, p, rpf, vf);
    for (int f = 0; f < 3; f++) {
        double dr0 = rpf[f] - rp0[f];
        double drf = disp(Q, f, vf[f], coord(Q, f, p[f], rpf[f]), dt) / Q.h[f];
        rp0[f] += 0.5 * (dr0 + drf);
    }
    bool ok = adjust(Q, p, rp0, rpf, 1);
    for (int f = 0; f < 3; f++) {
        T.p[f][t] = p[f]; T.rp0[f][t] = rp0[f]; T.rpf[f][t] = rpf[f]; T.vf[f][t] = vf[f];
    }
    if (!ok) T.alive[t] = 0;
}

void launch_tracer_predict(hipStream_t s, const TrParams &Q, const TracerDev &T, const Fluid &U, double dt) {
    if (T.n <= 0) return;
    hipLaunchKernelGGL(k_tracer_predict, dim3((T.n + 63) / 64), dim3(64), 0, s, Q, T, U, dt);
}
void launch_tracer_advect(hipStream_t s, const TrParams &Q, const TracerDev &T, const Fluid &U, double dt) {
    if (T.n <= 0) return;
    hipLaunchKernelGGL(k_tracer_advect, dim3((T.n + 63) / 64), dim3(64), 0, s, Q, T, U, dt);
}

}  // namespace rmt
