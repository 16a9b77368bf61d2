"""Tracer particles of the oracle (CPU only), pinned by closed forms:
polynomial reproduction of the tricubic interpolant, exact transport in a
uniform flow with periodic wrap, the no-slip wall guard and the loss of
tracers that leave through an open boundary."""
import numpy as np
import pytest

from oracle.oracle import OracleSim, taus
from rmt3d_amd import _abi, configs


def test_taus_stream_is_uniform():
    # gsl_rng_taus is absent: its bit stream is parity unpinned; the
    # restated generator must at least be a good uniform source
    x = taus(1, 200000).astype(np.float64) / 2 ** 32
    assert abs(x.mean() - 0.5) < 3e-3
    h, _ = np.histogram(x, bins=64, range=(0, 1))
    chi2 = ((h - len(x) / 64) ** 2 / (len(x) / 64)).sum()
    assert chi2 < 120                      # 63 degrees of freedom
    assert len(np.unique(taus(1, 100000))) > 99990
    assert not np.array_equal(taus(1, 8), taus(2, 8))


def _box(n=8, periodic=True, ntr=0):
    c, o = configs.taylor_green(n)
    if not periodic:
        configs._walls(c)
        _abi.finalize_config(c)
    c.dump_code = 1 if ntr else 0
    c.ntracers = ntr
    c.dt = 0.01
    return c, o


def _cell_field(S, fn):
    m, n, o = S.dims
    k, j, i = np.meshgrid(np.arange(o + 4), np.arange(n + 4), np.arange(m + 4), indexing="ij")
    return fn(i.ravel().astype(float), j.ravel().astype(float), k.ravel().astype(float))


def test_interpolant_reproduces_tricubic_polynomials():
    # tracers::interp_vels fits the tricubic polynomial through the 4^3
    # cube: any polynomial of degree <= 3 in each index is reproduced
    c, o = _box(8)
    S = OracleSim(c, o)
    f = lambda i, j, k: 1 + 0.5 * i - 0.25 * i ** 3 + 0.125 * j ** 2 * k - k ** 3 * i ** 2 * 0.01
    u = _cell_field(S, f)
    S.upload(u=u, v=2 * u, w=-u)
    rng = np.random.default_rng(5)
    for _ in range(50):
        p = rng.integers(0, 8, 3)
        r = rng.uniform(-0.5, 0.5, 3)
        x = p + 1.5 + r                   # padded index coordinate of the point
        want = f(*x)
        v = S.tracer_interp(p, r)
        assert v[0] == pytest.approx(want, rel=1e-12, abs=1e-12)
        assert v[1] == pytest.approx(2 * want, rel=1e-12, abs=1e-12)
        assert v[2] == pytest.approx(-want, rel=1e-12, abs=1e-12)


def test_uniform_flow_transport_with_periodic_wrap():
    # improved Euler is exact in a uniform flow; tracers leaving a
    # periodic side come back through the other
    c, o = _box(8, ntr=50)
    S = OracleSim(c, o)
    assert S.tracers_init() == 0
    U = np.array([0.3, -0.45, 0.2])
    Ng = S.Ng
    S.upload(u=np.full(Ng, U[0]), v=np.full(Ng, U[1]), w=np.full(Ng, U[2]))
    S.set_dt(0.05)
    x0 = S.tracer_positions().astype(np.float64)
    for _ in range(40):
        S.tracers_move()
    x1 = S.tracer_positions().astype(np.float64)
    want = np.mod(x0 + 40 * 0.05 * U, 1.0)
    d = np.abs(x1 - want)
    d = np.minimum(d, 1 - d)
    assert d.max() < 1e-5


def test_no_slip_wall_guard():
    # tracers::dx: heading into a homogeneous Dirichlet wall a tracer stops
    # at 1 % of the box width from it
    c, o = _box(8, periodic=False, ntr=40)
    S = OracleSim(c, o)
    S.tracers_init()
    Ng = S.Ng
    S.upload(u=np.full(Ng, -2.0), v=np.zeros(Ng), w=np.zeros(Ng))
    S.set_dt(0.05)
    for _ in range(30):
        S.tracers_move()
    x = S.tracer_positions()
    assert np.all(x[:, 0] >= 0.01 - 1e-6) and np.all(x[:, 0] < 0.0101)


def test_tracers_leave_through_open_boundaries():
    # a wall with a prescribed non-zero velocity is no guard: a tracer
    # carried past it has no neighbour to go to and is dropped (written 0)
    c, o = _box(8, periodic=False, ntr=40)
    for comp in range(3):
        c.bcv[comp][0] = 1.0
    S = OracleSim(c, o)
    S.tracers_init()
    Ng = S.Ng
    S.upload(u=np.full(Ng, -2.0), v=np.zeros(Ng), w=np.zeros(Ng))
    S.set_dt(0.05)
    for _ in range(30):
        S.tracers_move()
    assert np.all(S.tracer_positions() == 0)


def test_initial_positions_fill_the_box():
    c, o = _box(16, ntr=4000)
    S = OracleSim(c, o)
    S.initialize()
    x = S.tracer_positions()
    assert x.shape == (4000, 3)
    assert np.all(x >= 0) and np.all(x < 1)
    assert np.all(np.abs(x.mean(axis=0) - 0.5) < 0.02)
