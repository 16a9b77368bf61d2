"""Writes tests/golden/mat_algebra.npz: seeded 3x3 inputs and the outputs of
the reference's own matrix routines (Utilities/matlib.hh matrix::inverse,
det, aat, ata and Utilities/mat3.hh mat3::inverse), evaluated through
oracle/_ref/libref_mat.so, which oracle/ref.mk compiles from the reference
tree.  Run from the repository root:  python tests/golden/make_golden.py
"""
import ctypes as C
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import oracle  # noqa: E402


def dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def main():
    if oracle.build_ref(force=True) is None:
        raise SystemExit("the reference tree is absent; nothing to generate")
    R = oracle.ref_lib()
    rng = np.random.default_rng(20261202)
    n = 64
    A = rng.uniform(-2, 2, size=(n, 9))
    # include near-identity deformation gradients and a pre-stretch
    A[0] = np.eye(3).ravel()
    lam = 1.3
    A[1] = np.diag([lam, lam ** -0.5, lam ** -0.5]).ravel()
    A[2:10] = (np.eye(3)[None] + 1e-3 * rng.standard_normal((8, 3, 3))).reshape(8, 9)
    inv = np.zeros((n, 9)); det = np.zeros(n); aat = np.zeros((n, 6)); ata = np.zeros((n, 6))
    inv3 = np.zeros((n, 9)); det3 = np.zeros(n)
    for r in range(n):
        a = np.ascontiguousarray(A[r])
        o = np.zeros(9); R.ref_matrix_inverse(dp(a), dp(o)); inv[r] = o
        det[r] = R.ref_matrix_det(dp(a))
        s = np.zeros(6); R.ref_matrix_aat(dp(a), dp(s)); aat[r] = s
        s = np.zeros(6); R.ref_matrix_ata(dp(a), dp(s)); ata[r] = s
        o = np.zeros(9); d = C.c_double(); R.ref_mat3_inverse(dp(a), dp(o), C.byref(d))
        inv3[r] = o; det3[r] = d.value
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "mat_algebra.npz")
    np.savez(out, A=A, inv=inv, det=det, aat=aat, ata=ata, inv3=inv3, det3=det3)
    print("wrote", out)


if __name__ == "__main__":
    main()
