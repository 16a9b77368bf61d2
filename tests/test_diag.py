"""The macro.dat / centroid.dat reductions of the oracle (CPU only), pinned
by closed forms: uniform fields, an affine pre-stretched solid, a
symmetric sphere, and a deterministic logarithm checked against libm."""
import math

import numpy as np
import pytest

from oracle.oracle import OracleSim, lib
from rmt3d_amd import _abi, configs


def test_det_log_within_two_ulp_of_libm():
    L = lib()
    rng = np.random.default_rng(11)
    xs = np.concatenate([np.exp(rng.uniform(-700, 700, 20000)), 1 + rng.uniform(-1e-3, 1e-3, 5000),
                         rng.uniform(0.5, 2.0, 5000), [1.0, 2.0, 0.5, math.e, 1e-300, 5e-324, 1.7e308]])
    for x in xs:
        a, b = L.oracle_det_log(float(x)), math.log(float(x))
        assert abs(a - b) <= 2 * math.ulp(abs(b)) or abs(a - b) < 1e-300, (x, a, b)
    assert L.oracle_det_log(1.0) == 0.0
    assert L.oracle_det_log(0.0) == -math.inf
    assert math.isnan(L.oracle_det_log(-1.0))


def _uniform(S, u):
    Ng = S.Ng
    S.upload(u=np.full(Ng, u[0]), v=np.full(Ng, u[1]), w=np.full(Ng, u[2]))


def test_momentum_and_kinetic_energy_of_uniform_flow():
    # total_momentum = U V, kinetic = |U|^2 V / 2 (fluid_3d_utils.cc:783-870)
    cfg, objs = configs.taylor_green(16)
    S = OracleSim(cfg, objs)
    U = (1.0, 2.0, -0.5)
    _uniform(S, U)
    d = S.macro_data()
    for a, k in enumerate(("momentum_x", "momentum_y", "momentum_z")):
        assert d[k] == pytest.approx(U[a], rel=1e-14)
    assert d["kinetic"] == pytest.approx(0.5 * sum(x * x for x in U), rel=1e-14)
    assert d["energy"] == d["kinetic"]
    assert d["elastic"] == 0 and d["potential"] == 0 and d["power"] == 0
    assert d["max_dev_detF"] == 0 and d["dev_detF_l2"] == 0


def test_mac_divergence_extremes():
    # the MAC divergence sum_f (fvel[2f+1][f] - fvel[2f][f]) / h per cell
    cfg, objs = configs.taylor_green(8)
    S = OracleSim(cfg, objs)
    Ng, m = S.Ng, 8
    fv = np.zeros((18, Ng))
    a = (3 + 2) + (m + 4) * ((2 + 2) + (m + 4) * (5 + 2))
    b = (1 + 2) + (m + 4) * ((6 + 2) + (m + 4) * (0 + 2))
    fv[3, a] = 0.25          # face 1, x component: +0.25 * 8 = 2
    fv[10, b] = -0.5         # face 3, y component: -0.5 * 8 = -4
    S.upload(fvel=fv.ravel())
    d = S.macro_data()
    assert d["divu_max"] == 2.0 and d["divu_min"] == -4.0


def test_affine_prestretched_sphere_energy_and_det():
    # F = diag(l, l^-1/2, l^-1/2) inside the sphere: det F = 1 and the
    # strain energy density 0.5 G (tr(F^T F) - 3 - 2 ln det F)
    # (fluid_3d.cc:1933-1938); the face-averaged total approaches the density
    # times the solid volume
    n, lam = 32, 1.3
    cfg, objs = configs.stretched_sphere(n, lam=lam)
    S = OracleSim(cfg, objs)
    S.initialize()
    lib().oracle_compute_stress(S.h)
    d = S.macro_data()
    # the map scales by (sl^-2, sl, sl), sl = sqrt(lam) (object.hh:95-114),
    # so F = (grad xi)^-1 = diag(lam, lam^-1/2, lam^-1/2)
    sl = math.sqrt(lam)
    eig = [sl, sl, sl]
    eig[0] = 1 / (sl * sl)
    F = [1 / e for e in eig]
    W = 0.5 * (sum(f * f for f in F) - 3 - 2 * math.log(F[0] * F[1] * F[2]))
    vol = 4 / 3 * math.pi * 0.25 ** 3
    assert d["elastic"] == pytest.approx(W * vol, rel=0.15)
    assert d["max_dev_detF"] < 1e-12 and d["dev_detF_l2"] < 1e-12
    assert d["potential"] == 0.0


def test_identity_sphere_has_no_strain():
    cfg, objs = configs.cavity_sphere(24)
    S = OracleSim(cfg, objs)
    S.initialize()
    lib().oracle_compute_stress(S.h)
    d = S.macro_data()
    assert abs(d["elastic"]) < 1e-14 and d["max_dev_detF"] < 1e-14


def test_potential_energy_of_solid_cells():
    # g (z - az) summed over solid cells times the cell volume
    cfg, objs = configs.cavity_sphere(16)
    cfg.gravity = 2.0
    S = OracleSim(cfg, objs)
    S.initialize()
    st = S.download(band=False)
    m, n, o = cfg.sys_size
    c = st["rm_c"].reshape(o + 4, n + 4, m + 4)[2:o + 2, 2:n + 2, 2:m + 2]
    z = cfg.xyz_bounds[4] + (np.arange(o) + 0.5) / o
    want = 2.0 * sum(np.count_nonzero(c[k] & 1023) * (z[k] - cfg.xyz_bounds[4]) for k in range(o)) / (m * n * o)
    assert S.macro_data()["potential"] == pytest.approx(want, rel=1e-13)


def test_centroid_of_centred_sphere_and_average_velocity():
    # compute_v_centroid (fluid_3d_utils.cc:1290-1346): H-weighted mean of
    # the cell centres and of the velocity over the primary cells
    cfg, objs = configs.stretched_sphere(24)
    S = OracleSim(cfg, objs)
    S.initialize()
    _uniform(S, (0.3, -0.2, 0.1))
    c = S.centroids()
    assert c.shape == (1, 6)
    assert np.allclose(c[0, :3], 0.5, rtol=0, atol=1e-14)
    assert np.allclose(c[0, 3:], (0.3, -0.2, 0.1), rtol=1e-14, atol=0)


def test_ftest_has_no_centroids():
    cfg, objs = configs.taylor_green(8)
    S = OracleSim(cfg, objs)
    assert S.centroids().shape == (0, 6)
