// Zero-contour surfaces of the objects for the ctr.<frame> files
// (fluid_3d::output_contours, fluid_3d_io.cc:595-958).  The lattice is the
// cell centres (i,j,k), 0 <= i <= m etc.; phi of object cnum at a cell is
// that of its map for the object (primary first, then the first band
// entry; 100 without one, fluid_3d::phi :459-466).  A vertex sits on each
// lattice edge whose end values change sign, at the linear root, with the
// blended, normalised central-difference gradient as its normal; boxes
// whose eight corners are not all of one sign are triangulated with the
// table of mc_table.h.  Vertices are numbered in lattice order (k, j, i,
// then edge direction x, y, z) and triangles in box order, the order the
// reference's loop finds them; two exclusive scans give every node its
// first vertex and triangle.
#include "launch.h"
#include "mc_table.h"

namespace rmt {

namespace {
__device__ __forceinline__ bool crosses(double v1, double v2) { return v1 > 0 ? v2 <= 0 : v2 > 0; }
}  // namespace

__global__ void k_ct_phi(Geo g, const ObjC *O, int nobj, Maps M, int cnum, double *phi) {
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= g.Ng) return;
    double v = 100;
    if ((int)(M.r.c[c] & OMASK) == cnum) {
        v = obj_phi(O, nobj, cnum - 1, M.r.x[0][c], M.r.x[1][c], M.r.x[2][c]);
    } else {
        int s = M.b.start[c], e = s + M.b.cnt[c];
        for (int a = s; a < e; a++)
            if ((int)(M.b.c[a] & OMASK) == cnum) {
                v = obj_phi(O, nobj, cnum - 1, M.b.x[0][a], M.b.x[1][a], M.b.x[2][a]);
                break;
            }
    }
    phi[c] = v;
}

// fluid_3d::edge_detect / box_detect counts (fluid_3d_io.cc:857-866, 924-945)
__global__ void k_ct_count(Geo g, CtBufs B) {
    int m1 = g.m + 1, n1 = g.n + 1, nn = m1 * n1 * (g.o + 1);
    int nd = blockIdx.x * blockDim.x + threadIdx.x;
    if (nd >= nn) return;
    int i = nd % m1, j = (nd / m1) % n1, k = nd / (m1 * n1);
    int c = (i + 2) + g.m4 * ((j + 2) + g.n4 * (k + 2));
    const double *p = B.phi;
    double v = p[c];
    int nv = 0;
    if (i < g.m && crosses(v, p[c + 1])) nv++;
    if (j < g.n && crosses(v, p[c + g.m4])) nv++;
    if (k < g.o && crosses(v, p[c + g.mn4])) nv++;
    B.vcnt[nd] = nv;
    int nt = 0;
    if (i < g.m && j < g.n && k < g.o) {
        int key = v > 0 ? 1 : 0;
        if (p[c + 1] > 0) key |= 2;
        if (p[c + g.m4 + 1] > 0) key |= 4;
        if (p[c + g.m4] > 0) key |= 8;
        if (p[c + g.mn4] > 0) key |= 16;
        if (p[c + g.mn4 + 1] > 0) key |= 32;
        if (p[c + g.mn4 + g.m4 + 1] > 0) key |= 64;
        if (p[c + g.mn4 + g.m4] > 0) key |= 128;
        nt = mc_ntri[key];
    }
    B.tcnt[nd] = nt;
}

// vertices: fluid_3d::edge_detect, grad_field, normalize and the master's
// position assembly (fluid_3d_io.cc:735-760, 887-958)
__global__ void k_ct_verts(Geo g, CtBufs B) {
    int m1 = g.m + 1, n1 = g.n + 1, nn = m1 * n1 * (g.o + 1);
    int nd = blockIdx.x * blockDim.x + threadIdx.x;
    if (nd >= nn) return;
    int i = nd % m1, j = (nd / m1) % n1, k = nd / (m1 * n1);
    int c = (i + 2) + g.m4 * ((j + 2) + g.n4 * (k + 2));
    const double *p = B.phi;
    const double tol = 2.220446049250313e-16;
    const int str[3] = {1, g.m4, g.mn4};
    const double hh[3] = {g.dx, g.dy, g.dz};
    const bool ok[3] = {i < g.m, j < g.n, k < g.o};
    int slot = B.voff[nd];
    for (int dir = 0; dir < 3; dir++) {
        B.vidx[3 * nd + dir] = -1;
        if (!ok[dir]) continue;
        int c2 = c + str[dir];
        double v1 = p[c], v2 = p[c2];
        if (!crosses(v1, v2)) continue;
        double fa = v2 - v1, fb = fabs(fa) < tol ? 0.5 : -v1 / fa;
        double x = B.gx[i], y = B.gy[j], z = B.gz[k];
        double off = fb * hh[dir];
        if (dir == 0) x += off;
        else if (dir == 1) y += off;
        else z += off;
        double nx = g.dxsp * (p[c + 1] - p[c - 1]), ny = g.dysp * (p[c + g.m4] - p[c - g.m4]),
               nz = g.dzsp * (p[c + g.mn4] - p[c - g.mn4]);
        double nx2 = g.dxsp * (p[c2 + 1] - p[c2 - 1]), ny2 = g.dysp * (p[c2 + g.m4] - p[c2 - g.m4]),
               nz2 = g.dzsp * (p[c2 + g.mn4] - p[c2 - g.mn4]);
        nx = (1 - fb) * nx + fb * nx2;
        ny = (1 - fb) * ny + fb * ny2;
        nz = (1 - fb) * nz + fb * nz2;
        double rsq = nx * nx + ny * ny + nz * nz;
        if (rsq > tol * tol) {
            rsq = 1. / sqrt(rsq);
            nx *= rsq;
            ny *= rsq;
            nz *= rsq;
        }
        B.pts[3 * slot] = (float)x;
        B.pts[3 * slot + 1] = (float)y;
        B.pts[3 * slot + 2] = (float)z;
        B.nrm[3 * slot] = (float)nx;
        B.nrm[3 * slot + 1] = (float)ny;
        B.nrm[3 * slot + 2] = (float)nz;
        B.vidx[3 * nd + dir] = slot;
        slot++;
    }
}

// triangles of each box, vertices looked up through the lattice edges
// (fluid_3d_io.cc:794-818 gives the edge of each table entry)
__global__ void k_ct_tris(Geo g, CtBufs B) {
    int m1 = g.m + 1, n1 = g.n + 1, nn = m1 * n1 * (g.o + 1);
    int nd = blockIdx.x * blockDim.x + threadIdx.x;
    if (nd >= nn || B.tcnt[nd] == 0) return;
    int i = nd % m1, j = (nd / m1) % n1, k = nd / (m1 * n1);
    int c = (i + 2) + g.m4 * ((j + 2) + g.n4 * (k + 2));
    const double *p = B.phi;
    int key = p[c] > 0 ? 1 : 0;
    if (p[c + 1] > 0) key |= 2;
    if (p[c + g.m4 + 1] > 0) key |= 4;
    if (p[c + g.m4] > 0) key |= 8;
    if (p[c + g.mn4] > 0) key |= 16;
    if (p[c + g.mn4 + 1] > 0) key |= 32;
    if (p[c + g.mn4 + g.m4 + 1] > 0) key |= 64;
    if (p[c + g.mn4 + g.m4] > 0) key |= 128;
    // edge e: node offset (di, dj, dk) and direction
    const signed char eo[12][4] = {{0, 0, 0, 0}, {1, 0, 0, 1}, {0, 1, 0, 0}, {0, 0, 0, 1}, {0, 0, 1, 0}, {1, 0, 1, 1},
                                   {0, 1, 1, 0}, {0, 0, 1, 1}, {0, 0, 0, 2}, {1, 0, 0, 2}, {1, 1, 0, 2}, {0, 1, 0, 2}};
    int base = B.toff[nd];
    for (int t = 0; t < mc_ntri[key]; t++)
        for (int v = 0; v < 3; v++) {
            int e = mc_tri[key][3 * t + v];
            int nb = nd + eo[e][0] + m1 * (eo[e][1] + n1 * eo[e][2]);
            B.tri[3 * (base + t) + v] = B.vidx[3 * nb + eo[e][3]];
        }
}

void launch_contour_count(hipStream_t s, const Geo &g, const ObjC *O, int nobj, const Maps &M, int cnum,
                          const CtBufs &B) {
    int nn = (g.m + 1) * (g.n + 1) * (g.o + 1);
    hipLaunchKernelGGL(k_ct_phi, dim3((g.Ng + 255) / 256), dim3(256), 0, s, g, O, nobj, M, cnum, B.phi);
    hipLaunchKernelGGL(k_ct_count, dim3((nn + 255) / 256), dim3(256), 0, s, g, B);
    launch_exclusive_scan(s, B.vcnt, B.voff, B.bsum, nn, B.tot);
    launch_exclusive_scan(s, B.tcnt, B.toff, B.bsum, nn, B.tot + 1);
}

void launch_contour_emit(hipStream_t s, const Geo &g, const CtBufs &B) {
    int nn = (g.m + 1) * (g.n + 1) * (g.o + 1);
    hipLaunchKernelGGL(k_ct_verts, dim3((nn + 255) / 256), dim3(256), 0, s, g, B);
    hipLaunchKernelGGL(k_ct_tris, dim3((nn + 255) / 256), dim3(256), 0, s, g, B);
}

}  // namespace rmt
