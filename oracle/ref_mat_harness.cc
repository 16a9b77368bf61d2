// ORACLE -- test infrastructure only.
//
// A ctypes harness around the reference's own standalone 3x3 algebra
// (Utilities/matlib.hh and Utilities/mat3.hh, compiled from where they lie
// in the reference tree by oracle/ref.mk into oracle/_ref/).  It pins the
// oracle's restatements of matrix::inverse/det/aat/ata (matlib.hh:227-232,
// 303-329,366-370) and mat3::inverse (mat3.hh:61-67), the two routines the
// stress and the extrapolation least squares depend on.  The rest of the
// reference (PGMG, GSL, MPI-coupled classes) does not build here.
#include "matlib.hh"
#include "mat3.hh"

extern "C" {

void ref_matrix_inverse(const double *in, double *out) {
    matrix m(in[0], in[1], in[2], in[3], in[4], in[5], in[6], in[7], in[8]);
    matrix r = m.inverse();
    for (int a = 0; a < 9; a++) out[a] = r.v[a];
}

double ref_matrix_det(const double *in) {
    matrix m(in[0], in[1], in[2], in[3], in[4], in[5], in[6], in[7], in[8]);
    return m.det();
}

// symmetric results in the reference's packed order v[ind(i,j)]
void ref_matrix_aat(const double *in, double *out6) {
    matrix m(in[0], in[1], in[2], in[3], in[4], in[5], in[6], in[7], in[8]);
    sym_matrix s = m.aat();
    for (int a = 0; a < 6; a++) out6[a] = s.v[a];
}

void ref_matrix_ata(const double *in, double *out6) {
    matrix m(in[0], in[1], in[2], in[3], in[4], in[5], in[6], in[7], in[8]);
    sym_matrix s = m.ata();
    for (int a = 0; a < 6; a++) out6[a] = s.v[a];
}

void ref_mat3_inverse(const double *in, double *out, double *det) {
    mat3 m(in[0], in[1], in[2], in[3], in[4], in[5], in[6], in[7], in[8]);
    double d;
    mat3 r = m.inverse(d);
    out[0] = r.a11; out[1] = r.a12; out[2] = r.a13;
    out[3] = r.a21; out[4] = r.a22; out[5] = r.a23;
    out[6] = r.a31; out[7] = r.a32; out[8] = r.a33;
    *det = d;
}

}
