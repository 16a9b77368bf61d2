"""Parity of the HIP path against the FP64 oracle on the same inputs.

FP64 fields must agree to 1e-10 relative L2 (the bar of the FP64 GPU path);
primary object ids, band cells and band codes (integer indexing) must agree
bit-exactly."""
import numpy as np
import pytest

from rmt3d_amd import configs

from .helpers import compare_states, pair, rel_l2

pytestmark = pytest.mark.gpu

TOL = 1e-10


@pytest.mark.parametrize("which", [0, 1])
@pytest.mark.parametrize("maker", ["walls", "periodic"])
def test_mg_solve_matches_oracle(which, maker):
    cfg, objs = (configs.stretched_sphere(16) if maker == "walls" else configs.taylor_green(16))
    g, o = pair(cfg, objs)
    dims = g.mg_dims(which)
    n = int(np.prod(dims))
    rng = np.random.default_rng(1 + which)
    b = rng.standard_normal(n)
    b -= b.mean()
    x0 = np.zeros(n)
    xg, ncg, rg = g.mg_solve(which, b, x0)
    xo, nco, ro = o.mg_solve(which, b, x0)
    print("MG", maker, which, dims, "cycles", ncg, nco, "res", rg, ro)
    assert ncg == nco
    assert rel_l2(xg - xg.mean(), xo - xo.mean()) < 1e-12
    assert rg <= (6 if which == 0 else 32) * 2.220446049250313e-16 * 100


def _mg_pair_solve(which, dims, prd, prec, seed):
    from rmt3d_amd import _abi
    cfg, objs = configs.taylor_green(8)
    cfg.sys_size[:] = list(dims)
    cfg.prd[:] = list(prd)
    for a in range(3):                      # equal cell size in every dimension
        cfg.xyz_bounds[2 * a] = 0.0
        cfg.xyz_bounds[2 * a + 1] = dims[a] / 64.0
    g, o = pair(cfg, [], precision=_abi.PREC_MIXED if prec else _abi.PREC_FP64)
    mg_dims = g.mg_dims(which)
    n = int(np.prod(mg_dims))
    rng = np.random.default_rng(seed + which)
    b = rng.standard_normal(n)
    b -= b.mean()
    x0 = np.zeros(n)
    xg, ncg, rg = g.mg_solve(which, b, x0)
    xo, nco, ro = o.mg_solve(which, b, x0)
    print("MG", dims, prd, which, prec, "levels", g.diag(10 + which), "tail", g.diag(12 + which),
          "fdm", g.diag(18 + which), "cycles", ncg, nco, "res", rg, ro, flush=True)
    assert g.diag(18 + which) == float(o.mg_fdm(which))
    assert rg <= (6 if which == 0 else 32) * 2.220446049250313e-16 * 100
    d = rel_l2(xg - xg.mean(), xo - xo.mean())
    if prec == 0:
        assert ncg == nco
        assert d == 0.0
    else:
        assert d < 1e-9
    return g


@pytest.mark.parametrize("which", [0, 1])
@pytest.mark.parametrize("dims,prd", [((64, 64, 64), (0, 0, 0)), ((64, 64, 64), (1, 1, 1)),
                                      ((64, 48, 80), (1, 0, 1)), ((96, 64, 56), (0, 1, 0)),
                                      ((192, 96, 4), (0, 0, 0)), ((4, 96, 192), (0, 0, 0)),
                                      ((128, 128, 128), (0, 0, 0))])
@pytest.mark.parametrize("prec", [0, 1])
def test_mg_solve_fdm_level_matches_oracle(which, dims, prd, prec):
    # the first coarse level ends the hierarchy and is solved by fast
    # diagonalisation (walls, periodic, mixed and anisotropic extents); FP64
    # must match the oracle bit for bit, the mixed path converges to the same
    # FP64 tolerance (the autotuner picks the solve's launch form)
    g = _mg_pair_solve(which, dims, prd, prec, 3)
    assert g.diag(18 + which) == 1.0 and g.diag(10 + which) == 2.0


@pytest.mark.parametrize("which", [0, 1])
@pytest.mark.parametrize("direct", [1, 0])
@pytest.mark.parametrize("dims,prd,fits", [((128, 128, 128), (0, 0, 0), (1, 1)), ((64, 64, 64), (1, 1, 1), (1, 1)),
                                           ((64, 48, 80), (1, 0, 1), (1, 1)), ((192, 96, 4), (0, 0, 0), (1, 1)),
                                           ((4, 96, 192), (0, 0, 0), (1, 0))])
def test_mg_solve_fdm_mfma_form_matches_oracle(which, direct, dims, prd, fits, monkeypatch):
    # with the built-in defaults the mixed path's FP32 fast-diagonalisation
    # solve runs as three MFMA launches with direct operands (form 2: every
    # extent of these levels is at most 128); with RMT3D_FDM_DIRECT=0, as
    # the three launches staged in LDS wherever planes, pencil blocks and
    # bases fit a workgroup (fits: MAC, FEM level), and as six
    # single-contraction launches where they do not (the 3 x 49 x 97 FEM
    # level: padded 112 x 112 z-bases beside 64 pencils); the FP64 defect
    # correction around it converges to the oracle's solution
    monkeypatch.setenv("RMT3D_TUNE", "0")
    if not direct:
        monkeypatch.setenv("RMT3D_FDM_DIRECT", "0")
    g = _mg_pair_solve(which, dims, prd, 1, 7)
    assert g.diag(18 + which) == 1.0
    assert g.diag(31 + which) == (2.0 if direct else float(fits[which]))


@pytest.mark.parametrize("which", [0, 1])
@pytest.mark.parametrize("dims,prd", [((128, 128, 128), (0, 0, 0)), ((64, 48, 80), (1, 0, 1)),
                                      ((192, 96, 4), (0, 0, 0)), ((96, 64, 56), (0, 1, 0))])
def test_fdm_direct_form_is_bitwise_the_staged_form(which, dims, prd, monkeypatch):
    # the direct-operand MFMA kernels take the staged kernels' operands in
    # the same k-slot order: the mixed solves' iterates, and so their
    # solutions and cycle counts, are identical bit for bit
    from rmt3d_amd import _abi
    from rmt3d_amd.lib import Simulation
    cfg, _ = configs.taylor_green(8)
    cfg.sys_size[:] = list(dims)
    cfg.prd[:] = list(prd)
    for a in range(3):
        cfg.xyz_bounds[2 * a] = 0.0
        cfg.xyz_bounds[2 * a + 1] = dims[a] / 64.0
    monkeypatch.setenv("RMT3D_TUNE", "0")
    monkeypatch.setenv("RMT3D_FDM_DIRECT", "0")
    a = Simulation(cfg, [], _abi.PREC_MIXED)
    monkeypatch.delenv("RMT3D_FDM_DIRECT")
    b = Simulation(cfg, [], _abi.PREC_MIXED)
    assert a.diag(31 + which) == 1.0 and b.diag(31 + which) == 2.0
    n = int(np.prod(a.mg_dims(which)))
    rng = np.random.default_rng(17 + which)
    rhs = rng.standard_normal(n)
    rhs -= rhs.mean()
    xa, na, ra = a.mg_solve(which, rhs, np.zeros(n))
    xb, nb, rb = b.mg_solve(which, rhs, np.zeros(n))
    print("direct vs staged", dims, prd, which, "cycles", na, nb, "res", ra, rb, flush=True)
    assert na == nb and ra == rb
    assert np.array_equal(xa, xb)
    a.close()
    b.close()


def test_fdm_mfma_form_gives_the_same_states(monkeypatch):
    """The mixed path's FP32 fast-diagonalisation solve on MFMA (three
    launches, direct operands) and as six single-contraction launches (RMT3D_FDM_FUSED=0)
    precondition the same FP64 defect correction: a 64^3 cavity with a
    sphere steps to velocities and pressures 1e-10 apart (relative L2; the
    velocity as one field, since v alone is nearly zero in the lid-driven
    box) either way."""
    from rmt3d_amd import _abi
    from rmt3d_amd.lib import Simulation
    cfg, objs = configs.cavity_sphere(64, steps=10)
    monkeypatch.setenv("RMT3D_TUNE", "0")
    monkeypatch.setenv("RMT3D_FDM_FUSED", "0")
    a = Simulation(cfg, objs, _abi.PREC_MIXED)
    monkeypatch.delenv("RMT3D_FDM_FUSED")
    b = Simulation(cfg, objs, _abi.PREC_MIXED)
    assert a.diag(18) == 1.0 and a.diag(19) == 1.0
    assert a.diag(31) == 0.0 and a.diag(32) == 0.0
    assert b.diag(31) == 2.0 and b.diag(32) == 2.0
    for s in (a, b):
        e = s.initialize()
        s.prepare_run(e)
        for _ in range(3):
            assert s.step() == 0
    fa, fb = a.download(), b.download()
    va = np.concatenate([fa[k] for k in ("u", "v", "w")])
    vb = np.concatenate([fb[k] for k in ("u", "v", "w")])
    for k, d in (("vel", rel_l2(va, vb)), ("p", rel_l2(fa["p"], fb["p"]))):
        print(k, d, flush=True)
        assert d < TOL, (k, d)
    a.close()
    b.close()


@pytest.mark.parametrize("dims,prd", [((64, 64, 64), (0, 0, 0)), ((64, 48, 80), (1, 0, 1)),
                                      ((160, 128, 192), (1, 1, 1))])
def test_fem_solve_repeats_bit_for_bit(dims, prd):
    # every smoothing launch form the autotuner can pick reads only values
    # that no other workgroup of the same launch writes, so the FP64 FEM solve
    # is the oracle's bit for bit and a second solve of the same right-hand
    # side repeats it bit for bit (walls, mixed periodic, all-periodic)
    g = _mg_pair_solve(1, dims, prd, 0, 11)
    assert g.diag(25) >= 0.0
    n = int(np.prod(g.mg_dims(1)))
    rng = np.random.default_rng(12)
    b = rng.standard_normal(n)
    b -= b.mean()
    x1, n1, r1 = g.mg_solve(1, b, np.zeros(n))
    x2, n2, r2 = g.mg_solve(1, b, np.zeros(n))
    assert n1 == n2 and r1 == r2
    assert np.array_equal(x1, x2)


@pytest.mark.parametrize("which", [0, 1])
@pytest.mark.parametrize("dims,prd", [((160, 160, 160), (0, 0, 0)), ((160, 128, 192), (1, 1, 1))])
@pytest.mark.parametrize("prec", [0, 1])
def test_mg_solve_fused_levels_match_oracle(which, dims, prd, prec):
    # a level between the finest and the fast-diagonalisation level runs the
    # fused multi-colour sweeps (LDS tiles with halos, wrapped on periodic
    # dimensions)
    g = _mg_pair_solve(which, dims, prd, prec, 5)
    assert g.diag(18 + which) == 1.0 and g.diag(10 + which) == 3.0


def _init_pair(cfg, objs):
    g, o = pair(cfg, objs)
    eg = g.initialize()
    eo = o.initialize()
    assert eg == eo == 1
    return g, o


@pytest.mark.parametrize("n", [16, 24])
def test_initialize_matches_oracle(n):
    cfg, objs = configs.stretched_sphere(n)
    g, o = _init_pair(cfg, objs)
    fg, fo = g.download(), o.download()
    r = compare_states(fg, fo, cfg)
    print("init", n, r)
    assert r["rm_c_equal"]
    assert r["band_equal"]
    assert r["band_x"] < 1e-13
    for k in ("u", "v", "w"):
        assert r[k] < TOL


def _run_steps(maker, n, steps, **kw):
    cfg, objs = maker(n, **kw)
    g, o = _init_pair(cfg, objs)
    init_err = 1
    if maker is configs.taylor_green:
        # the Taylor-Green field enters as a checkpoint: initialize_from_chk_point
        # succeeds, so init_iter skips its warm-up (fluid_3d.cc:194-213, 238)
        u, v, w = configs.taylor_green_field(cfg)
        g.upload(u=u, v=v, w=w)
        o.upload(u=u, v=v, w=w)
        init_err = 0
    _, dtg, _ = g.prepare_run(init_err)
    _, dto, _ = o.prepare_run(init_err)
    assert dtg == dto
    for s in range(steps):
        rg, ro = g.step(), o.step()
        assert rg == ro == 0, (rg, ro)
    return cfg, g, o


@pytest.mark.parametrize("case", [("stretched", 24, 3), ("cavity", 24, 3), ("tg", 16, 3),
                                  ("stretched", 32, 6), ("cavity", 32, 6), ("tg", 32, 10)])
def test_steps_match_oracle(case):
    name, n, steps = case
    maker = {"stretched": configs.stretched_sphere, "cavity": configs.cavity_sphere,
             "tg": configs.taylor_green}[name]
    cfg, g, o = _run_steps(maker, n, steps)
    fg, fo = g.download(), o.download()
    r = compare_states(fg, fo, cfg)
    print(name, n, steps, r, "time", g.time, o.time)
    assert abs(g.time - o.time) < 1e-14
    assert r["rm_c_equal"]
    assert r.get("band_equal", True)
    for k in ("vel", "p"):
        assert r[k] < TOL, (k, r[k])


@pytest.mark.parametrize("seed", [5, 8])
def test_multi_sphere_steps_match_oracle(seed):
    """Several spheres of different radii and moduli in a lid-driven box:
    overlapping bands, per-cell band order by object, collision stress."""
    from rmt3d_amd import serve
    rng = np.random.default_rng(seed)
    cfg, objs, steps = serve.mixed_request(rng, steps=4, grids=(32,), max_spheres=6)
    g, o = _init_pair(cfg, objs)
    _, dtg, _ = g.prepare_run(1)
    _, dto, _ = o.prepare_run(1)
    assert dtg == dto
    for s in range(steps):
        rg, ro = g.step(), o.step()
        assert rg == ro == 0, (rg, ro)
    fg, fo = g.download(), o.download()
    r = compare_states(fg, fo, cfg)
    print("multi-sphere", seed, len(objs), r)
    assert r["rm_c_equal"]
    assert r.get("band_equal", True)
    for k in ("vel", "p"):
        assert r[k] < TOL, (k, r[k])


@pytest.mark.parametrize("prec", [0, 1])
def test_graph_replay_matches_eager_launches(prec):
    """V-cycles replayed from their captured graph (default) and launched one
    kernel at a time (per-kernel profiling) give identical states, replay
    after replay, with the step's own eager work in between."""
    from rmt3d_amd import _abi
    from rmt3d_amd.lib import Simulation
    cfg, objs = configs.stretched_sphere(32)
    out = []
    for level in (0, 2):
        s = Simulation(cfg, objs, prec)
        e = s.initialize()
        s.prepare_run(e)
        s.profile(level)
        for _ in range(4):
            assert s.step() == 0
        out.append(s.download())
        s.close()
    for k in ("u", "v", "w", "p", "rm_x", "band_x"):
        assert np.array_equal(out[0][k], out[1][k]), k


def test_shear_spheres_match_oracle():
    """Simple shear (moving walls in z, periodic in x and y) with several
    spheres of different moduli."""
    cfg, objs = configs.shear_spheres(32, 5, rmin=0.09, rmax=0.14, seed=4)
    g, o = _init_pair(cfg, objs)
    _, dtg, _ = g.prepare_run(1)
    _, dto, _ = o.prepare_run(1)
    assert dtg == dto
    for s in range(4):
        rg, ro = g.step(), o.step()
        assert rg == ro == 0, (rg, ro)
    r = compare_states(g.download(), o.download(), cfg)
    print("shear spheres", len(objs), r)
    assert r["rm_c_equal"] and r.get("band_equal", True)
    for k in ("vel", "p"):
        assert r[k] < TOL, (k, r[k])


@pytest.mark.parametrize("prec", [0, 1])
def test_autotuned_shapes_give_identical_states(prec, monkeypatch):
    """The multigrid autotuner changes launch shapes only: a 64^3 cavity
    with a sphere (fused sweeps and a fast-diagonalisation level) steps to
    the same bits with the tuned shapes and with the built-in defaults.  On
    the mixed path it also picks the FP32 V-cycle's smoothing sweeps and
    the MFMA or six-launch fast-diagonalisation form; those runs are held
    to the same bits where the choices equal the defaults, else to 1e-10
    (relative L2)."""
    from rmt3d_amd import _abi
    from rmt3d_amd.lib import Simulation
    cfg, objs = configs.cavity_sphere(64, steps=10)
    p = _abi.PREC_MIXED if prec else _abi.PREC_FP64
    monkeypatch.setenv("RMT3D_TUNE", "0")
    a = Simulation(cfg, objs, p)
    monkeypatch.delenv("RMT3D_TUNE")
    b = Simulation(cfg, objs, p)
    keys = (16, 17, 24, 25, 26, 27, 28, 29, 31, 32, 33, 34)
    print("tuned:", [b.diag(k) for k in keys], "defaults:", [a.diag(k) for k in keys], flush=True)
    assert a.diag(28) == 0 and a.diag(29) == 0
    assert b.diag(28) in (1, 2) and b.diag(29) in (1, 2)
    assert a.diag(33) == a.diag(34) == 22.0
    if prec:
        assert b.diag(33) in (22.0, 21.0, 12.0) and b.diag(34) in (22.0, 21.0, 12.0, 3.0)
    else:
        assert b.diag(33) == b.diag(34) == 22.0
    for s in (a, b):
        e = s.initialize()
        s.prepare_run(e)
        for _ in range(3):
            assert s.step() == 0
    fa, fb = a.download(), b.download()
    same = all(a.diag(k) == b.diag(k) for k in (31, 32, 33, 34))
    if not prec or same:
        for k in ("u", "v", "w", "p", "rm_x", "band_x"):
            assert np.array_equal(fa[k], fb[k]), k
    else:
        va = np.concatenate([fa[k] for k in ("u", "v", "w")])
        vb = np.concatenate([fb[k] for k in ("u", "v", "w")])
        assert rel_l2(va, vb) < TOL and rel_l2(fa["p"], fb["p"]) < TOL
    a.close()
    b.close()


@pytest.mark.parametrize("sched", ["2,1", "1,2", "0,3"])
def test_mixed_smoothing_schedules_give_the_same_states(sched, monkeypatch):
    """The mixed path's FP32 V-cycle only preconditions the FP64 defect
    correction: with 2 + 1, 1 + 2 or 0 + 3 smoothing sweeps (the last
    restricting the right-hand side without a residual pass) instead of
    2 + 2 a 64^3 cavity with a sphere steps to velocities and pressures
    1e-10 apart."""
    from rmt3d_amd import _abi
    from rmt3d_amd.lib import Simulation
    cfg, objs = configs.cavity_sphere(64, steps=10)
    monkeypatch.setenv("RMT3D_MIXED_VCYCLE", "2,2")
    a = Simulation(cfg, objs, _abi.PREC_MIXED)
    monkeypatch.setenv("RMT3D_MIXED_VCYCLE", sched)
    b = Simulation(cfg, objs, _abi.PREC_MIXED)
    d, u = (int(v) for v in sched.split(","))
    assert a.diag(33) == a.diag(34) == 22.0
    assert b.diag(33) == b.diag(34) == 10.0 * d + u
    for s in (a, b):
        e = s.initialize()
        s.prepare_run(e)
        for _ in range(3):
            assert s.step() == 0
    fa, fb = a.download(), b.download()
    va = np.concatenate([fa[k] for k in ("u", "v", "w")])
    vb = np.concatenate([fb[k] for k in ("u", "v", "w")])
    for k, dv in (("vel", rel_l2(va, vb)), ("p", rel_l2(fa["p"], fb["p"]))):
        print(sched, k, dv, flush=True)
        assert dv < TOL, (k, dv)
    a.close()
    b.close()


def test_mixed_fused_passes_give_identical_states(monkeypatch):
    """The mixed path's defect-correction step writes the first colour pass
    with the residual and adds the correction during the last sweep; with
    RMT3D_MIXED_FUSE=0 the same values come from separate launches.  A 64^3
    cavity with a sphere steps to the same bits either way."""
    from rmt3d_amd import _abi
    from rmt3d_amd.lib import Simulation
    cfg, objs = configs.cavity_sphere(64, steps=10)
    # built-in launch shapes and schedule, so that only the fusion differs
    monkeypatch.setenv("RMT3D_TUNE", "0")
    monkeypatch.setenv("RMT3D_MIXED_FUSE", "0")
    a = Simulation(cfg, objs, _abi.PREC_MIXED)
    monkeypatch.delenv("RMT3D_MIXED_FUSE")
    b = Simulation(cfg, objs, _abi.PREC_MIXED)
    for s in (a, b):
        e = s.initialize()
        s.prepare_run(e)
        for _ in range(3):
            assert s.step() == 0
    fa, fb = a.download(), b.download()
    for k in ("u", "v", "w", "p", "rm_x", "band_x"):
        assert np.array_equal(fa[k], fb[k]), k
    assert a.diag(1) == b.diag(1) and a.diag(2) == b.diag(2)
    a.close()
    b.close()
