"""Known-answer tests that pin the FP64 oracle (CPU only).

Each test checks one oracle function against something that does not come
from the oracle itself: the reference's own 3x3 algebra (compiled from its
headers into oracle/_ref, and its committed outputs in tests/golden/), a
closed form, an invariant, a brute force, or a textbook discretisation
assembled independently here.
"""
import collections
import ctypes as C
import math
import os

import numpy as np
import pytest

from oracle import oracle
from oracle.oracle import OracleSim, lib
from rmt3d_amd import configs

HERE = os.path.dirname(os.path.abspath(__file__))


def dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


# ---------------------------------------------------------------- slopes
def mono(u):
    return lib().oracle_mono_diff(*[float(x) for x in u])


def test_mono_diff_constant_and_linear():
    # fluid_3d::mono_diff (fluid_3d.cc:2310-2340): a fourth-order limited
    # slope (Colella 1985; Yu 2003 eqs 3.25-3.26) is exact on linear data
    assert mono([3.0] * 5) == 0.0
    for a, b in ((0.0, 1.0), (2.0, -0.5), (-1.0, 3.25)):
        u = [a + b * k for k in range(5)]
        assert mono(u) == pytest.approx(b, rel=1e-14)


def test_mono_diff_extremum_and_bound():
    # a local extremum at the centre gives zero slope
    assert mono([0, 1, 2, 1, 0]) == 0.0
    assert mono([5, 3, 1, 2, 4]) == 0.0
    # the limited slope never exceeds twice the smaller one-sided difference
    rng = np.random.default_rng(7)
    for _ in range(2000):
        u = rng.standard_normal(5)
        s = mono(u)
        d1, d2 = u[2] - u[1], u[3] - u[2]
        if d1 * d2 <= 0:
            assert s == 0.0
        else:
            assert abs(s) <= 2 * min(abs(d1), abs(d2)) * (1 + 1e-15)
            assert np.sign(s) in (0.0, np.sign(u[3] - u[1]))


def test_mono_diff_textbook_form():
    # the published limiter: delta_lim = 2 min(|d-|,|d+|) (same sign, else 0),
    # delta_f = sgn(dc) min(|dc|, delta_lim), slope = sgn(dc) min(delta_lim,
    # |4/3 dc - (delta_f(i-1) + delta_f(i+1)) / 6|)
    def dlim(a, b, c):
        l, r = b - a, c - b
        return 2 * min(abs(l), abs(r)) if l * r > 0 else 0.0

    def df(a, b, c):
        cd = 0.5 * (c - a)
        return math.copysign(1.0, cd if cd != 0 else 1.0) * min(abs(cd), dlim(a, b, c))

    rng = np.random.default_rng(11)
    for _ in range(2000):
        u = rng.standard_normal(5)
        cd = 0.5 * (u[3] - u[1])
        want = math.copysign(1.0, cd if cd != 0 else 1.0) * min(
            dlim(u[1], u[2], u[3]), abs(4.0 / 3.0 * cd - (df(*u[0:3]) + df(*u[2:5])) / 6.0))
        assert mono(u) == pytest.approx(want, rel=1e-13, abs=1e-15)


# ---------------------------------------------------------------- upwinding
def burgers_state(ul, ur):
    """Entropy solution of u_t + (u^2/2)_x = 0 at x/t = 0.  Returns the state
    and which side it came from ('L', 'R', or 'S' for the sonic/stagnation
    value 0)."""
    if ul > ur:                       # shock, speed (ul+ur)/2
        s = 0.5 * (ul + ur)
        return (ul, "L") if s > 0 else (ur, "R") if s < 0 else (0.0, "S")
    if ul >= 0 and ul > 0:            # rarefaction entirely to the right
        return ul, "L"
    if ur < 0:                        # rarefaction entirely to the left
        return ur, "R"
    return 0.0, "S"                   # sonic point inside the fan


@pytest.mark.parametrize("ul", [-2.0, -0.5, 0.0, 0.5, 2.0])
@pytest.mark.parametrize("ur", [-2.0, -0.5, 0.0, 0.5, 2.0])
def test_godunov_velocity_sign_table(ul, ur):
    # godunov_set_velocity (fluid_3d.cc:2156-2181)
    v = np.array([ul, 1.0, 2.0, ur, 3.0, 4.0])
    lib().oracle_godunov_vel(dp(v))
    state, side = burgers_state(ul, ur)
    assert v[0] == v[3] == state
    if side == "L":
        assert (v[1], v[2], v[4], v[5]) == (1.0, 2.0, 1.0, 2.0)
    elif side == "R":
        assert (v[1], v[2], v[4], v[5]) == (3.0, 4.0, 3.0, 4.0)
    else:
        assert (v[1], v[2], v[4], v[5]) == (2.0, 3.0, 2.0, 3.0)


@pytest.mark.parametrize("ul", [-1.0, 0.0, 1.0])
@pytest.mark.parametrize("ur", [-1.0, 0.0, 1.0])
def test_godunov_refmap_follows_velocity(ul, ur):
    # godunov_set_refmap (fluid_3d.cc:2184-2204): the map is upwinded with the
    # side the velocity Riemann problem picks, averaged at stagnation
    m = np.array([1.0, 2.0, 3.0, 5.0, 7.0, 11.0])
    lib().oracle_godunov_refmap(ul, ur, dp(m))
    _, side = burgers_state(ul, ur)
    want = {"L": [1, 2, 3] * 2, "R": [5, 7, 11] * 2, "S": [3, 4.5, 7] * 2}[side]
    assert list(m) == want


# ---------------------------------------------------------------- 3x3 algebra
GOLD = np.load(os.path.join(HERE, "golden", "mat_algebra.npz"))


def _oracle_mat(A):
    L = lib()
    a = np.ascontiguousarray(A)
    inv = np.zeros(9); L.oracle_mat_inverse(dp(a), dp(inv))
    aat = np.zeros(6); L.oracle_mat_aat(dp(a), dp(aat))
    ata = np.zeros(6); L.oracle_mat_ata(dp(a), dp(ata))
    inv3 = np.zeros(9); d3 = C.c_double(); L.oracle_mat3_inverse(dp(a), dp(inv3), C.byref(d3))
    return inv, L.oracle_mat_det(dp(a)), aat, ata, inv3, d3.value


def test_matrix_algebra_matches_reference_golden():
    # golden vectors written by tests/golden/make_golden.py from the
    # reference's matlib.hh / mat3.hh; bit-exact (same FP64 operation order)
    for r in range(GOLD["A"].shape[0]):
        inv, det, aat, ata, inv3, d3 = _oracle_mat(GOLD["A"][r])
        assert np.array_equal(inv, GOLD["inv"][r])
        assert det == GOLD["det"][r]
        assert np.array_equal(aat, GOLD["aat"][r])
        assert np.array_equal(ata, GOLD["ata"][r])
        assert np.array_equal(inv3, GOLD["inv3"][r])
        assert d3 == GOLD["det3"][r]


def test_matrix_algebra_matches_reference_build():
    if oracle.build_ref() is None or oracle.ref_lib() is None:
        pytest.skip("reference tree absent: covered by the golden vectors")
    R = oracle.ref_lib()
    rng = np.random.default_rng(5)
    for _ in range(500):
        a = rng.uniform(-3, 3, 9)
        inv, det, aat, ata, inv3, d3 = _oracle_mat(a)
        o = np.zeros(9); R.ref_matrix_inverse(dp(a), dp(o))
        assert np.array_equal(inv, o)
        assert det == R.ref_matrix_det(dp(a))
        o = np.zeros(6); R.ref_matrix_aat(dp(a), dp(o)); assert np.array_equal(aat, o)
        o = np.zeros(6); R.ref_matrix_ata(dp(a), dp(o)); assert np.array_equal(ata, o)
        o = np.zeros(9); d = C.c_double(); R.ref_mat3_inverse(dp(a), dp(o), C.byref(d))
        assert np.array_equal(inv3, o) and d3 == d.value


def test_matrix_algebra_invariants():
    rng = np.random.default_rng(9)
    for _ in range(200):
        A = rng.uniform(-2, 2, (3, 3))
        inv, det, aat, ata, inv3, d3 = _oracle_mat(A.ravel())
        assert np.allclose(inv.reshape(3, 3) @ A, np.eye(3), atol=1e-9 * max(1, np.linalg.cond(A)))
        assert det == pytest.approx(np.linalg.det(A), rel=1e-12, abs=1e-13)
        sym = lambda s: np.array([[s[0], s[1], s[2]], [s[1], s[3], s[4]], [s[2], s[4], s[5]]])
        assert np.allclose(sym(aat), A @ A.T, atol=1e-13)
        assert np.allclose(sym(ata), A.T @ A, atol=1e-13)
        assert np.allclose(inv3, inv, rtol=1e-10, atol=1e-10)


# ---------------------------------------------------------------- trig, heaviside
def test_det_sincos_within_one_ulp_of_libm():
    L = lib()
    rng = np.random.default_rng(3)
    xs = np.concatenate([rng.uniform(-60, 60, 20000), np.linspace(-1e-3, 1e-3, 101),
                         [0.0, math.pi / 2, math.pi, -math.pi / 4]])
    for x in xs:
        for f, g in ((L.oracle_det_sin, math.sin), (L.oracle_det_cos, math.cos)):
            a, b = f(float(x)), g(float(x))
            assert abs(a - b) <= 1.01 * math.ulp(max(abs(b), 1e-300)) or abs(a - b) < 1e-17


def test_heaviside_closed_form():
    # sim_manager::heaviside (sim_manager.hh:259-268): the smoothed step
    # 1/2 - phi/(2 eps) - sin(pi phi/eps)/(2 pi) on |phi| < eps
    cfg, objs = configs.stretched_sphere(16)
    S = OracleSim(cfg, objs)
    eps = S.scalar(1)
    H = lambda p: lib().oracle_heaviside(S.h, p)
    assert H(eps) == 0.0 and H(2 * eps) == 0.0
    assert H(-eps) == 1.0 and H(-3 * eps) == 1.0
    assert H(0.0) == 0.5
    for p in np.linspace(-0.999, 0.999, 41) * eps:
        want = 0.5 - p / (2 * eps) - math.sin(math.pi * p / eps) / (2 * math.pi)
        assert H(p) == pytest.approx(want, abs=2e-16)
        assert H(p) + H(-p) == pytest.approx(1.0, abs=2e-16)


# ---------------------------------------------------------------- least squares
def lsq(pos, val, w):
    L = lib()
    n = len(w)
    davg, ravg, coeff, det = np.zeros(3), np.zeros(3), np.zeros(9), C.c_double()
    ok = L.oracle_lsq(n, dp(np.ascontiguousarray(pos.ravel())), dp(np.ascontiguousarray(val.ravel())),
                      dp(np.ascontiguousarray(w)), float(np.sum(w)), dp(davg), dp(ravg), dp(coeff), C.byref(det))
    return ok, davg, ravg, coeff.reshape(3, 3), det.value


def test_lsq_reproduces_affine_maps_for_any_weights():
    # extrap_helper::least_square (extrap_helper.cc:34-169): a weighted affine
    # fit is exact on affine data whatever the (positive) weights
    rng = np.random.default_rng(13)
    for _ in range(100):
        n = rng.integers(4, 60)
        pos = rng.uniform(-2, 2, (n, 3))
        A = rng.uniform(-1, 1, (3, 3))
        b = rng.uniform(-1, 1, 3)
        val = pos @ A.T + b
        w = rng.uniform(0.05, 1.0, n)
        ok, davg, ravg, coeff, det = lsq(pos, val, w)
        assert ok == 1
        assert np.allclose(coeff, A, atol=1e-11)
        assert np.allclose(davg, (w[:, None] * pos).sum(0) / w.sum(), atol=1e-14)
        assert np.allclose(ravg + coeff @ (np.zeros(3) - davg), b, atol=1e-11)


def test_lsq_rejects_degenerate_stencils():
    rng = np.random.default_rng(17)
    pos = rng.uniform(-1, 1, (20, 3))
    pos[:, 2] = 0.3                        # coplanar points: det = 0
    ok, *_ = lsq(pos, rng.standard_normal((20, 3)), np.ones(20))
    assert ok == 0
    ok, *_ = lsq(np.zeros((0, 3)), np.zeros((0, 3)), np.zeros(0))
    assert ok == 0


# ---------------------------------------------------------------- multigrid operators
def _mg_cfg(dims, prd):
    cfg, objs = configs.taylor_green(8)
    cfg.sys_size[:] = list(dims)
    cfg.prd[:] = list(prd)
    return cfg, []


def _dense(S, which):
    d = S.mg_levels(which)[0]
    n = int(np.prod(d))
    A = np.zeros((n, n))
    for c in range(n):
        e = np.zeros(n)
        e[c] = 1.0
        A[:, c] = S.mg_apply(which, e)
    return d, A


def _nid(i, j, k, d):
    return i + d[0] * (j + d[1] * k)


def _mac_laplacian(d, prd):
    # cell-centred 7-point Laplacian with homogeneous Neumann walls: the
    # graph Laplacian of face-adjacent cells (simp_mac, mg_probs.hh:497-530)
    n = int(np.prod(d))
    A = np.zeros((n, n))
    for k in range(d[2]):
        for j in range(d[1]):
            for i in range(d[0]):
                c = _nid(i, j, k, d)
                for ax in range(3):
                    for s in (-1, 1):
                        q = [i, j, k]
                        q[ax] += s
                        if q[ax] < 0 or q[ax] >= d[ax]:
                            if not prd[ax]:
                                continue
                            q[ax] %= d[ax]
                        A[c, _nid(*q, d)] += 1
                        A[c, c] -= 1
    return A


def _q1_element():
    # stiffness of the trilinear hexahedron on the unit cube by 2-point Gauss
    g = np.array([-1, 1]) / math.sqrt(3) * 0.5 + 0.5
    corners = [(a, b, c) for c in (0, 1) for b in (0, 1) for a in (0, 1)]
    K = np.zeros((8, 8))
    for x in g:
        for y in g:
            for z in g:
                grads = []
                for (a, b, c) in corners:
                    fx = (x if a else 1 - x)
                    fy = (y if b else 1 - y)
                    fz = (z if c else 1 - z)
                    dfx = 1 if a else -1
                    dfy = 1 if b else -1
                    dfz = 1 if c else -1
                    grads.append([dfx * fy * fz, fx * dfy * fz, fx * fy * dfz])
                G = np.array(grads)
                K += G @ G.T / 8.0
    return corners, K


def _fem_stiffness(cells, prd):
    # nodal Q1 Laplacian assembled element by element: walls give natural
    # (Neumann) boundary rows, periodic dimensions wrap (simp_fem,
    # mg_probs.hh:458-495, scaled by 12 so the interior row is 32/-2/-1/0)
    d = [cells[a] if prd[a] else cells[a] + 1 for a in range(3)]
    n = int(np.prod(d))
    corners, K = _q1_element()
    A = np.zeros((n, n))
    for k in range(cells[2]):
        for j in range(cells[1]):
            for i in range(cells[0]):
                ids = [_nid((i + a) % d[0], (j + b) % d[1], (k + c) % d[2], d) for (a, b, c) in corners]
                for p in range(8):
                    for q in range(8):
                        A[ids[p], ids[q]] += 12 * K[p, q]
    return d, A


@pytest.mark.parametrize("dims,prd", [((6, 4, 5), (0, 0, 0)), ((6, 4, 4), (1, 1, 1)), ((6, 4, 5), (1, 1, 0))])
def test_mac_operator_is_neumann_laplacian(dims, prd):
    cfg, objs = _mg_cfg(dims, prd)
    S = OracleSim(cfg, objs)
    d, A = _dense(S, 0)
    assert tuple(d) == tuple(dims)
    assert np.array_equal(A, _mac_laplacian(d, prd))


@pytest.mark.parametrize("dims,prd", [((6, 4, 5), (0, 0, 0)), ((6, 4, 4), (1, 1, 1)), ((6, 4, 5), (1, 1, 0))])
def test_fem_operator_is_assembled_q1_stiffness(dims, prd):
    cfg, objs = _mg_cfg(dims, prd)
    S = OracleSim(cfg, objs)
    d, A = _dense(S, 1)
    dq, Q = _fem_stiffness(dims, prd)
    assert tuple(d) == tuple(dq)
    assert np.allclose(A, Q, rtol=0, atol=1e-12)


@pytest.mark.parametrize("which", [0, 1])
@pytest.mark.parametrize("dims,prd", [((8, 8, 8), (0, 0, 0)), ((8, 8, 8), (1, 1, 1)), ((8, 4, 6), (1, 0, 1))])
def test_mg_solution_matches_dense_solve(which, dims, prd):
    cfg, objs = _mg_cfg(dims, prd)
    S = OracleSim(cfg, objs)
    d, A = _dense(S, which)
    n = A.shape[0]
    rng = np.random.default_rng(which + 3)
    b = rng.standard_normal(n)
    b -= b.mean()                       # compatible with the constant null space
    x, ncyc, res = S.mg_solve(which, b, np.zeros(n))
    xd = np.linalg.lstsq(A, b, rcond=None)[0]
    assert ncyc >= 1
    assert np.max(np.abs(A @ x - b)) <= (6 if which == 0 else 32) * 2.220446049250313e-14 * 1.0001
    dx = (x - x.mean()) - (xd - xd.mean())
    assert np.linalg.norm(dx) <= 1e-9 * np.linalg.norm(xd)


# ------------------------------------------------ exact coarsest-level solve
@pytest.mark.parametrize("which", [0, 1])
@pytest.mark.parametrize("dims,prd", [((4, 4, 4), (0, 0, 0)), ((4, 4, 4), (1, 1, 1)), ((6, 4, 4), (1, 0, 1)),
                                      ((2, 2, 2), (0, 0, 0)), ((5, 3, 2), (0, 0, 0))])
def test_pinv_of_operator_is_moore_penrose(which, dims, prd):
    # the four Moore-Penrose conditions and numpy's SVD pseudo-inverse, on the
    # (singular: constant null space) Neumann / periodic operators
    from oracle.oracle import sym_pinv
    cfg, objs = _mg_cfg(dims, prd)
    S = OracleSim(cfg, objs)
    d, A = _dense(S, which)
    P = sym_pinv(A)
    scale = np.abs(A).max()
    assert np.array_equal(P, P.T)
    assert np.abs(A @ P @ A - A).max() <= 1e-12 * scale
    assert np.abs(P @ A @ P - P).max() <= 1e-12 * np.abs(P).max()
    assert np.abs((A @ P).T - A @ P).max() <= 1e-12
    assert np.abs(P @ np.ones(len(A))).max() <= 1e-12 * np.abs(P).max()     # minimum norm: P 1 = 0
    assert np.allclose(P, np.linalg.pinv(A), rtol=0, atol=1e-11 * np.abs(P).max())


def test_pinv_of_spd_matrix_is_inverse():
    from oracle.oracle import sym_pinv
    rng = np.random.default_rng(7)
    M = rng.standard_normal((40, 40))
    A = M @ M.T + 40 * np.eye(40)
    P = sym_pinv(A)
    assert np.abs(P @ A - np.eye(40)).max() <= 1e-13


@pytest.mark.parametrize("which,dims,prd,nlev", [(0, (4, 4, 4), (0, 0, 0), 1), (1, (4, 4, 4), (0, 0, 0), 1),
                                                 (0, (16, 16, 16), (0, 0, 0), 3), (1, (16, 16, 16), (0, 0, 0), 3),
                                                 (0, (8, 8, 8), (1, 1, 1), 2)])
def test_hierarchy_ends_at_direct_level(which, dims, prd, nlev):
    # coarsening stops at the first level of at most 160 unknowns, which the
    # pseudo-inverse solves exactly; a one-level hierarchy converges in the
    # first cycle
    cfg, objs = _mg_cfg(dims, prd)
    S = OracleSim(cfg, objs)
    levels = S.mg_levels(which)
    assert len(levels) == nlev
    assert int(np.prod(levels[-1])) <= 160 < (int(np.prod(levels[-2])) if nlev > 1 else 161)
    assert S.mg_direct(which)
    if nlev == 1:
        d, A = _dense(S, which)
        b = np.random.default_rng(1).standard_normal(len(A))
        b -= b.mean()
        x, ncyc, res = S.mg_solve(which, b, np.zeros(len(A)))
        assert np.abs(A @ x - b).max() <= 1e-12


# ------------------------------------------------- fast-diagonalisation level
FDM_SHAPES = [((6, 5, 4), (0, 0, 0)), ((8, 4, 6), (1, 0, 0)), ((4, 8, 4), (1, 1, 1)), ((7, 2, 3), (0, 0, 0)),
              ((2, 6, 4), (1, 0, 1)), ((9, 9, 9), (0, 0, 0)), ((3, 5, 2), (1, 1, 0)), ((12, 3, 5), (0, 1, 1))]


def _level_dense(fem, d, prd):
    from oracle.oracle import level_apply
    n = int(np.prod(d))
    return np.stack([level_apply(fem, d, prd, e) for e in np.eye(n)], 1)


@pytest.mark.parametrize("fem", [0, 1])
@pytest.mark.parametrize("d,prd", FDM_SHAPES)
def test_fdm_solve_inverts_the_level_operator(fem, d, prd):
    # x = S D^+ S^T b solves A x = b for every b in the range of A (walls,
    # periodic and mixed dimensions, odd periodic extents, extent 2), and
    # differs from the minimum-norm solution pinv(A) b only by a constant
    from oracle.oracle import fdm_solve, level_apply
    rng = np.random.default_rng(sum(d) + 7 * fem)
    A = _level_dense(fem, d, prd)
    y = rng.standard_normal(len(A))
    b = A @ y
    x = fdm_solve(fem, d, prd, b)
    assert np.abs(level_apply(fem, d, prd, x) - b).max() <= 1e-13 * np.abs(b).max()
    xp = np.linalg.pinv(A) @ b
    assert np.abs((x - x.mean()) - (xp - xp.mean())).max() <= 1e-13 * np.abs(xp).max()
    # the null mode (the constants) is dropped by D^+: the right-hand side
    # M 1 along it (FEM: the Q1 mass row sums, 1/2 on a wall node per wall
    # dimension; MAC: the constants) solves to zero
    w = []
    for a in range(3):
        v = np.ones(d[a])
        if fem and not prd[a]:
            v[0] = v[-1] = 0.5
        w.append(v)
    bn = np.einsum("k,j,i->kji", w[2], w[1], w[0]).ravel()
    assert np.abs(A @ np.ones(len(A))).max() <= 1e-13
    assert np.abs(fdm_solve(fem, d, prd, bn)).max() <= 1e-13


@pytest.mark.parametrize("fem", [0, 1])
def test_fdm_solve_of_one_mode(fem):
    # a single separable mode is an eigenvector: for MAC on walls
    # b = cos(pi p (i+1/2)/d) ... has x = b / (sum of the 1D eigenvalues)
    from oracle.oracle import fdm_solve
    d, prd = (8, 6, 5), (0, 0, 0)
    p = (3, 1, 2)
    if fem:
        g = [np.cos(np.pi * p[a] * np.arange(d[a]) / (d[a] - 1)) for a in range(3)]
        c = [np.cos(np.pi * p[a] / (d[a] - 1)) for a in range(3)]
        lam = [6 * (1 - ci) / (2 + ci) for ci in c]
        # K v = lam M v per dimension: A (gx gy gz) = 12 sum lam_a M-products
        mv = []
        for a in range(3):
            v = g[a]
            m = 4 * v
            m[0], m[-1] = 2 * v[0], 2 * v[-1]
            m[1:] += v[:-1]
            m[:-1] += v[1:]
            mv.append(m / 6)
        b = 12 * (lam[0] * np.einsum("k,j,i->kji", mv[2], mv[1], mv[0]) +
                  lam[1] * np.einsum("k,j,i->kji", mv[2], mv[1], mv[0]) +
                  lam[2] * np.einsum("k,j,i->kji", mv[2], mv[1], mv[0])).ravel()
        want = np.einsum("k,j,i->kji", g[2], g[1], g[0]).ravel()
    else:
        g = [np.cos(np.pi * p[a] * (np.arange(d[a]) + 0.5) / d[a]) for a in range(3)]
        lam = [-2 * (1 - np.cos(np.pi * p[a] / d[a])) for a in range(3)]
        want = np.einsum("k,j,i->kji", g[2], g[1], g[0]).ravel()
        b = sum(lam) * want
    x = fdm_solve(fem, d, prd, b)
    assert np.abs(x - want).max() <= 1e-13


@pytest.mark.parametrize("which,dims,prd,levels,fdm", [
    (0, (64, 64, 64), (0, 0, 0), [(64, 64, 64), (32, 32, 32)], True),
    (1, (64, 64, 64), (0, 0, 0), [(65, 65, 65), (33, 33, 33)], True),
    (0, (96, 64, 56), (0, 1, 0), [(96, 64, 56), (48, 32, 28)], True),
    (1, (160, 160, 160), (0, 0, 0), [(161, 161, 161), (81, 81, 81), (41, 41, 41)], True),
    (0, (32, 32, 32), (0, 0, 0), [(32, 32, 32), (16, 16, 16), (8, 8, 8), (4, 4, 4)], False),
    (0, (1024, 16, 16), (0, 0, 0), [(1024, 16, 16), (512, 8, 8)], True),
    (0, (2048, 8, 8), (0, 0, 0), [(2048, 8, 8), (1024, 4, 4), (512, 2, 2)], False),
])
def test_hierarchy_ends_at_fdm_level(which, dims, prd, levels, fdm):
    # the first coarse level with 6144 < unknowns <= 300000 and at most 512
    # per dimension ends the hierarchy; smaller problems keep the
    # pseudo-inverse level, and a level too long in one dimension is passed
    cfg, objs = _mg_cfg(dims, prd)
    S = OracleSim(cfg, objs)
    assert [tuple(int(v) for v in l) for l in S.mg_levels(which)] == levels
    assert S.mg_fdm(which) == fdm


# ---------------------------------------------------------------- band layers
def _bfs_layers(rm_c, oid, pad_dims, nlayers):
    """6-neighbour BFS distance from the object's primary cells over the
    padded grid; cells on the outermost padded layer are reached but do not
    seed the next layer (extrap_helper: within the inner ghost region)."""
    m4, n4, o4 = pad_dims
    dist = {}
    q = collections.deque()
    for c in np.nonzero(rm_c == oid)[0]:
        dist[int(c)] = 0
        q.append(int(c))
    while q:
        c = q.popleft()
        L = dist[c]
        if L == nlayers:
            continue
        i, j, k = c % m4, (c // m4) % n4, c // (m4 * n4)
        if L > 0 and not (1 <= i <= m4 - 2 and 1 <= j <= n4 - 2 and 1 <= k <= o4 - 2):
            continue
        for ax, s in ((0, -1), (0, 1), (1, -1), (1, 1), (2, -1), (2, 1)):
            p = [i, j, k]
            p[ax] += s
            if not (0 <= p[0] < m4 and 0 <= p[1] < n4 and 0 <= p[2] < o4):
                continue
            e = p[0] + m4 * (p[1] + n4 * p[2])
            if e not in dist:
                dist[e] = L + 1
                q.append(e)
    return {c: L for c, L in dist.items() if L > 0}


def _centres(cfg):
    m, n, o = cfg.sys_size
    b = cfg.xyz_bounds
    h = [(b[1] - b[0]) / m, (b[3] - b[2]) / n, (b[5] - b[4]) / o]
    lx = b[0] + (np.arange(m + 4) - 1.5) * h[0]
    ly = b[2] + (np.arange(n + 4) - 1.5) * h[1]
    lz = b[4] + (np.arange(o + 4) - 1.5) * h[2]
    return lx, ly, lz, h


@pytest.mark.parametrize("maker", ["stretched", "cavity"])
def test_band_layers_are_bfs_distance_and_affine_exact(maker):
    if maker == "stretched":
        cfg, objs = configs.stretched_sphere(16)
        lam = 1.3
        eig = np.array([1 / lam, math.sqrt(lam), math.sqrt(lam)])
    else:
        cfg, objs = configs.cavity_sphere(16)
        eig = np.ones(3)
    S = OracleSim(cfg, objs)
    assert S.initialize() == 1
    f = S.download()
    m, n, o = cfg.sys_size
    pad = (m + 4, n + 4, o + 4)
    want = _bfs_layers(f["rm_c"], 1, pad, cfg.nlayers)
    got = {}
    for a in range(len(f["band_cell"])):
        c = int(f["band_c"][a])
        assert c & 1023 == 1
        cell = int(f["band_cell"][a])
        assert cell not in got
        got[cell] = c >> 10
    assert got == want
    # the initial map is affine, so every extrapolated value is exact
    lx, ly, lz, _ = _centres(cfg)
    ctr = np.array(objs[0].basic[:3])
    nb = len(f["band_cell"])
    for a in range(nb):
        cell = int(f["band_cell"][a])
        i, j, k = cell % pad[0], (cell // pad[0]) % pad[1], cell // (pad[0] * pad[1])
        x = np.array([lx[i], ly[j], lz[k]])
        xi = ctr + eig * (x - ctr)
        for d in range(3):
            assert abs(f["band_x"][d * nb + a] - xi[d]) < 1e-12


# ---------------------------------------------------------------- stress
@pytest.mark.parametrize("maker", ["stretched", "cavity"])
def test_solid_stress_closed_form(maker):
    # compute_stress/solid_stress (fluid_3d.cc:1360-1800): deep in an
    # incompressible neo-Hookean solid at rest with an affine map,
    # sigma = G dev(F F^T), F = (grad xi)^-1; identity map gives zero
    n = 32
    if maker == "stretched":
        cfg, objs = configs.stretched_sphere(n)
        lam = 1.3
        eig = np.array([1 / lam, math.sqrt(lam), math.sqrt(lam)])
    else:
        cfg, objs = configs.cavity_sphere(n)
        eig = np.ones(3)
    S = OracleSim(cfg, objs)
    assert S.initialize() == 1
    lib().oracle_compute_stress(S.h)
    f = S.download(band=False)
    F = np.diag(1 / eig)
    B = F @ F.T
    G = objs[0].shear_mod
    want = G * (B - np.trace(B) / 3 * np.eye(3))
    Ng = S.Ng
    lx, ly, lz, h = _centres(cfg)
    ctr = np.array(objs[0].basic[:3])
    R = objs[0].basic[3]
    eps = S.scalar(1)
    checked = 0
    pad = (n + 4,) * 3
    for k in range(2, n + 2):
        for j in range(2, n + 2):
            for i in range(2, n + 2):
                x = np.array([lx[i], ly[j], lz[k]])
                phi = np.linalg.norm(eig * (x - ctr)) - R
                if phi > -(eps + 4 * max(h) * eig.max()):
                    continue
                c = i + pad[0] * (j + pad[1] * k)
                sig = np.array([[f["sigma"][(3 * Fd + s) * Ng + c] for s in range(3)] for Fd in range(3)])
                assert np.allclose(sig, want, atol=1e-12), (i, j, k, sig)
                checked += 1
    assert checked > 0
