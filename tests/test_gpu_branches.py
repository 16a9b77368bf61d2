"""GPU parity of the step's rare branches and error paths, and at the
workloads' own sizes.

Each case asserts that the branch it names is taken (the step's branch
counters, or the returned error code) on the HIP path and on the oracle alike,
then compares the states: FP64 fields within 1e-10 relative L2 (bit-exact in
practice), object ids and band codes bit-exact.
"""
import math

import numpy as np
import pytest

from rmt3d_amd import _abi, configs

from .helpers import compare_states, pair
from .test_oracle_pins import _coords, corner_sphere, remove_edge_ghost_entry

pytestmark = pytest.mark.gpu

TOL = 1e-10


def _init(cfg, objs, precision=_abi.PREC_FP64):
    g, o = pair(cfg, objs, precision)
    eg, eo = g.initialize(), o.initialize()
    assert eg == eo
    return g, o, eg


def _assert_match(g, o, cfg, tol=TOL, what=""):
    r = compare_states(g.download(), o.download(), cfg)
    print(what, r, flush=True)
    assert r["rm_c_equal"] and r.get("band_equal", True)
    for k in ("vel", "p"):
        assert r[k] < tol, (k, r[k])
    return r


def test_collision_stress_parity():
    """Two spheres 0.03 apart in a lid-driven box, closer than the transition
    width eps = 1.5 h: collision_stress contributes on both paths in every
    step, with the same count, and the states agree."""
    cfg, objs = configs.cavity_sphere(32, steps=10)
    cfg.n_obj = 2
    objs = [configs.sphere((0.31, 0.5, 0.45), 0.17, G=1.5), configs.sphere((0.69, 0.5, 0.45), 0.17, G=0.7)]
    g, o, e = _init(cfg, objs)
    g.prepare_run(e)
    o.prepare_run(e)
    total = 0
    for s in range(4):
        assert g.step() == 0 and o.step() == 0
        cg, co = g.step_counters(), o.step_counters()
        print("step", s, "counters gpu", cg, "oracle", co, flush=True)
        assert cg == co
        assert cg[0] > 0
        total += cg[0]
    _assert_match(g, o, cfg, what="collision")
    assert total > 0


def test_redraw_parity_and_affine_transport():
    """A stress-free stretched sphere carried by a uniform periodic flow: the
    redraw moves maps on and off the primary grid on both paths with the same
    counts; the HIP path's maps equal the exactly transported affine map
    xi0(x - U t) and the states agree with the oracle."""
    n = 24
    c = _abi.default_config()
    for i in range(3):
        c.sys_size[i] = n
    c.fmu, c.fdt_pad, c.dt = 0.01, 0.4, 10.0
    c.sim_type = _abi.SIM_OBJECTS
    c.nlayers, c.wt_n, c.dump_code, c.n_obj = 4, 1.0, 0, 1
    _abi.finalize_config(c)
    lam = 1.3
    ob = configs.sphere((0.5, 0.5, 0.5), 0.15, G=0.0, stretch_dir=0, sqrt_lam=math.sqrt(lam))
    g, o, _ = _init(c, [ob])
    U = np.array([0.4, -0.3, 0.2])
    Ng = g.Ng
    vel = {k: np.full(Ng, U[a]) for a, k in enumerate(("u", "v", "w"))}
    g.upload(**vel)
    o.upload(**vel)
    g.prepare_run(0)
    o.prepare_run(0)
    X, _ = _coords(c)
    e = np.array([1 / lam, math.sqrt(lam), math.sqrt(lam)])
    flips = 0
    for s in range(8):
        assert g.step() == 0 and o.step() == 0
        cg, co = g.step_counters(), o.step_counters()
        assert cg == co, (s, cg, co)
        flips += cg[1] + cg[2]
    assert flips > 0
    f = g.download()
    want = 0.5 + e[:, None] * (X - U[:, None] * g.time - 0.5)
    prim = f["rm_c"] != 0
    assert np.abs(f["rm_x"].reshape(3, Ng)[:, prim] - want[:, prim]).max() < 1e-12
    nb = len(f["band_cell"])
    assert np.abs(f["band_x"].reshape(3, nb) - want[:, f["band_cell"]]).max() < 1e-12
    _assert_match(g, o, c, what="redraw")


def test_refmap_update_error_and_revert():
    """REFMAP_UPDATE_ERR (fluid_3d.cc:518-522, 1226-1243): the old band
    lacks one layer-3 map within eps of the surface; the fresh extrapolation
    creates it, both paths return 1 and revert the maps, and the reverted
    states are identical."""
    cfg, objs = corner_sphere(4.0)
    g, o, e = _init(cfg, objs)
    g.prepare_run(e)
    o.prepare_run(e)
    assert g.step() == 0 and o.step() == 0
    f0 = o.download()
    band, _ = remove_edge_ghost_entry(f0, cfg, 3)
    g.set_band(*band)
    o.set_band(*band)
    rg, ro = g.step(), o.step()
    assert rg == ro == _abi.REFMAP_UPDATE_ERR, (rg, ro)
    fg, fo = g.download(), o.download()
    for k in ("rm_x", "rm_xpred", "band_x", "band_c", "band_cell", "rm_c"):
        assert np.array_equal(fg[k], fo[k]), k
    _assert_match(g, o, cfg, what="refmap error")


@pytest.mark.parametrize("prec", [_abi.PREC_FP64, _abi.PREC_MIXED])
def test_projection_error_on_nan(prec):
    """PROJECTION_ERR (fluid_3d.cc:543-552): a NaN left in the pressure
    solver's initial guess gives a NaN residual in the next step's pressure
    projection; both paths return 2 with the same state."""
    cfg, objs = configs.stretched_sphere(24, steps=10)
    g, o, e = _init(cfg, objs, prec)
    g.prepare_run(e)
    o.prepare_run(e)
    assert g.step() == 0 and o.step() == 0
    nn = int(np.prod(g.mg_dims(1)))
    bad = np.full(nn, np.nan)
    rhs = np.zeros(nn)
    _, _, rg = g.mg_solve(1, rhs, bad)
    _, _, ro = o.mg_solve(1, rhs, bad)
    assert math.isnan(rg) and math.isnan(ro)
    sg, so = g.step(), o.step()
    assert sg == so == _abi.PROJECTION_ERR, (sg, so)
    fg, fo = g.download(), o.download()
    for k in ("u", "v", "w", "dvel", "rm_x", "rm_c"):
        if prec == _abi.PREC_FP64:
            assert np.array_equal(fg[k], fo[k]), k
        else:
            assert np.allclose(fg[k], fo[k], rtol=1e-9, atol=1e-12), k


def _full_run(maker, n, nsteps, prec, **kw):
    cfg, objs = maker(n, **kw)
    g, o, e = _init(cfg, objs, prec)
    _, dtg, _ = g.prepare_run(e)
    _, dto, _ = o.prepare_run(e)
    assert dtg == dto
    flips = 0
    for s in range(nsteps):
        rg, ro = g.step(), o.step()
        assert rg == ro == 0, (s, rg, ro)
        cg, co = g.step_counters(), o.step_counters()
        if prec == _abi.PREC_FP64:
            assert cg == co, (s, cg, co)
        flips += cg[1] + cg[2]
    return cfg, g, o, flips


def test_cavity_sphere_64_twenty_steps():
    """The 64^3 lid-driven cavity with one sphere (the C2 workload) for 20
    steps, FP64, against the oracle."""
    cfg, g, o, flips = _full_run(configs.cavity_sphere, 64, 20, _abi.PREC_FP64, steps=20)
    _assert_match(g, o, cfg, what="cavity 64^3 x 20")


@pytest.mark.parametrize("prec,tol", [(_abi.PREC_FP64, TOL), (_abi.PREC_MIXED, 1e-6)])
def test_stretched_sphere_128_three_steps(prec, tol):
    """The headline 128^3 stretched-sphere workload for 3 steps against the
    FP64 oracle: FP64 within 1e-10, mixed within 1e-6."""
    cfg, g, o, _ = _full_run(configs.stretched_sphere, 128, 3, prec, steps=10)
    _assert_match(g, o, cfg, tol=tol, what="stretched 128^3 x 3 prec %d" % prec)
