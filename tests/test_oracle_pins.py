"""Known-answer tests for the oracle functions whose behaviour the step
parity tests alone do not pin (CPU only).

Each test states what the mathematics or the geometry of the configuration
fixes, independently of the oracle's code:

  collision_stress       closed form on reference maps whose level sets are
                         linear (exact finite-difference normals)
  redraw_boundary        the set of cells a translated sphere covers
                         (brute force over the grid)
  corrector error rule   "a new band map without an old one, within eps of
                         the surface and in layers 1-3, is an error"
                         (constructed bands with one entry removed); revert
                         restores the maps
  wall ghost rules       linear / quadratic fields reproduced exactly by the
                         Dirichlet and Neumann ghost and face formulas
  reference-map transport an affine map in uniform flow is advected exactly,
                         through upwinding, extrapolation and redraws
  grad-phi weighted LSQ  numpy weighted affine fits over the candidate set
                         the gathering rule defines
"""
import math

import numpy as np
import pytest

from oracle.oracle import OracleSim
from rmt3d_amd import configs
from rmt3d_amd._abi import SIM_OBJECTS, default_config, finalize_config

OM, LD = 1023, 10


def _coords(cfg):
    m, n, o = cfg.sys_size
    b = cfg.xyz_bounds
    h = [(b[1] - b[0]) / m, (b[3] - b[2]) / n, (b[5] - b[4]) / o]
    ax = [b[2 * a] + (np.arange(d + 4) - 2 + 0.5) * h[a] for a, d in enumerate((m, n, o))]
    Z, Y, X = np.meshgrid(ax[2], ax[1], ax[0], indexing="ij")
    return np.stack([X.ravel(), Y.ravel(), Z.ravel()]), h


def _unpack(cfg, ind):
    m4, n4 = cfg.sys_size[0] + 4, cfg.sys_size[1] + 4
    return ind % m4, (ind // m4) % n4, ind // (m4 * n4)


def _walled(n, objs, nlayers=4, wt_n=1.0):
    c = default_config()
    for i in range(3):
        c.sys_size[i] = n
    for f in range(6):
        for comp in range(3):
            c.bct[comp][f] = 1
            c.bcv[comp][f] = 0.0
    c.fmu = 0.01
    c.dt = 10.0
    c.T = 8.0
    c.frames = 500
    c.sim_type = SIM_OBJECTS
    c.nlayers = nlayers
    c.wt_n = wt_n
    c.ex_visc_mult = 1.0
    c.ev_trans_mult = 0.0
    c.dump_code = 0
    c.n_obj = len(objs)
    finalize_config(c)
    return c, objs


# ------------------------------------------------------------ collision stress
B0 = 0.00137          # keeps every face's level-set values off the eps thresholds
def _collision_state(a0, a1, b1, G=(1.5, 0.7), n=16):
    """Two objects whose maps cover every cell: object 1 on the primary grid,
    object 2 as a layer-1 band entry.  Map k is xi_k(x) = c_k + (R_k + l_k(x)) e_k
    with l_k linear, so phi_k(xi_k(x)) = l_k(x) exactly."""
    objs = [configs.sphere((0.3, 0.5, 0.5), 0.2, G=G[0]), configs.sphere((0.7, 0.5, 0.5), 0.2, G=G[1])]
    cfg, objs = _walled(n, objs)
    o = OracleSim(cfg, objs)
    X, h = _coords(cfg)
    x0 = np.array([0.5, 0.5, 0.5])[:, None]
    l0 = (np.asarray(a0)[:, None] * (X - x0)).sum(0) + B0
    l1 = (np.asarray(a1)[:, None] * (X - x0)).sum(0) + b1
    e0 = np.array([0.6, 0.0, 0.8])
    e1 = np.array([0.0, -0.6, 0.8])
    xi0 = np.array(objs[0].basic[:3])[:, None] + (0.2 + l0)[None, :] * e0[:, None]
    xi1 = np.array(objs[1].basic[:3])[:, None] + (0.2 + l1)[None, :] * e1[:, None]
    assert (0.2 + l0).min() > 0 and (0.2 + l1).min() > 0
    Ng = X.shape[1]
    o.upload(rm_c=np.ones(Ng, dtype=np.uint32), rm_x=xi0.ravel(), rm_xpred=xi0.ravel())
    cells = np.arange(Ng, dtype=np.int32)
    o.set_band(cells, np.full(Ng, 2 | (1 << LD), dtype=np.uint32), xi1, xi1)
    return cfg, o, X, h


def _collision_closed_form(cfg, o, X, h, a0, a1, b1, G, ind, F):
    eps = o.scalar(1)
    x0 = np.array([0.5, 0.5, 0.5])
    xf = X[:, ind].copy()
    xf[F] -= 0.5 * h[F]                     # lower face centre
    p0 = float(np.dot(a0, xf - x0) + B0)    # phi averaged over the face's two cells
    p1 = float(np.dot(a1, xf - x0) + b1)
    assert min(abs(p0 - eps), abs(p1 - eps), abs(p0 + p1 - eps)) > 1e-9
    if p0 > eps or p1 > eps or p0 + p1 > eps:
        return np.zeros(3)
    d0 = np.asarray(a0) / np.linalg.norm(a0)
    d1 = np.asarray(a1) / np.linalg.norm(a1)
    nv = d1 - d0
    if np.linalg.norm(nv) < 100 * np.finfo(float).eps:
        return np.zeros(3)
    nv /= np.linalg.norm(nv)

    def cfun(x):
        return 0.0 if x > eps else 0.5 - x / (2 * eps)

    cf = G[0] * cfun(p0) + G[1] * cfun(p1)
    s = -cf * np.outer(nv, nv)
    s -= np.trace(s) / 3 * np.eye(3)
    return s[F]


@pytest.mark.parametrize("case", ["approaching", "parallel"])
def test_collision_stress_closed_form(case):
    """-(G0 c(phi0) + G1 c(phi1)) dev(n n^T), n = unit(grad phi1 - grad phi0),
    for faces within the transition zone of both objects and of their sum;
    zero elsewhere and for parallel normals (fluid_3d.cc:1465-1773)."""
    G = (1.5, 0.7)
    if case == "approaching":
        a0, a1, b1 = (0.2, 0.04, 0.02), (-0.2, 0.06, -0.04), 0.01291
    else:
        a0, a1, b1 = (0.2, 0.04, 0.02), (0.2, 0.04, 0.02), -0.03117
    cfg, o, X, h = _collision_state(a0, a1, b1, G)
    m = cfg.sys_size[0]
    m4 = m + 4
    G0 = 2 * (1 + m4 * (1 + m4))
    hit = skip = 0
    for k in range(4, m + 1):
        for j in range(4, m + 1):
            for i in range(4, m + 1):
                ind = i + m4 * (j + m4 * k)
                for F in range(3):
                    got = o.collision_face(ind - G0, F)
                    want = _collision_closed_form(cfg, o, X, h, a0, a1, b1, G, ind, F)
                    assert np.allclose(got, want, rtol=1e-12, atol=1e-12), (i, j, k, F, got, want)
                    if np.any(want != 0):
                        hit += 1
                    else:
                        skip += 1
    o.close()
    if case == "approaching":
        assert hit > 100 and skip > 100, (hit, skip)
    else:
        assert hit == 0


# --------------------------------------------------------------- redraw
def _translate(o, cfg, which, delta):
    """Shift the maps of the objects in `which` by -delta (the objects move by
    +delta); the band keeps its cells and codes."""
    f = o.download()
    Ng = o.Ng
    x = f["rm_x"].reshape(3, Ng).copy()
    sel = np.isin(f["rm_c"] & OM, which) & (f["rm_c"] != 0)
    x[:, sel] -= np.asarray(delta)[:, None]
    o.upload(rm_x=x.ravel())
    nb = len(f["band_cell"])
    bx = f["band_x"].reshape(3, nb).copy()
    bsel = np.isin(f["band_c"] & OM, which)
    bx[:, bsel] -= np.asarray(delta)[:, None]
    o.set_band(f["band_cell"], f["band_c"], bx, f["band_xpred"].reshape(3, nb))
    return f, x, bx


def _phi(objs, oid, xi):
    ob = objs[oid - 1]
    return np.linalg.norm(xi - np.array(ob.basic[:3])[:, None], axis=0) - ob.basic[3]


def test_redraw_single_sphere_covers_translated_sphere():
    """After the maps of a sphere are translated by less than h/sqrt(3),
    redraw_boundary makes the primary set exactly the cells whose centre lies
    inside the moved sphere; each primary map is the moved identity map; each
    map that left the primary grid is kept in the band at layer nlayers+1
    (fluid_3d_ex.cc:336-394)."""
    cfg, objs = _walled(14, [configs.sphere((0.5, 0.5, 0.5), 0.3)], nlayers=4)
    o = OracleSim(cfg, objs)
    o.initialize()
    X, h = _coords(cfg)
    delta = np.array([0.3, -0.2, 0.25]) * h[0]
    f0, x, bx = _translate(o, cfg, [1], delta)
    o.redraw()
    f1 = o.download()
    n_out, n_in = int(o.scalar(10)), int(o.scalar(11))
    inside = np.linalg.norm(X - delta[:, None] - 0.5, axis=0) < 0.3
    prim = f1["rm_c"] != 0
    assert np.array_equal(prim, inside)
    assert np.all(f1["rm_c"][prim] == 1)
    want = (X - delta[:, None])[:, prim]
    assert np.allclose(f1["rm_x"].reshape(3, -1)[:, prim], want, rtol=0, atol=1e-13)
    was = f0["rm_c"] != 0
    left = was & ~inside
    came = ~was & inside
    assert n_out == int(left.sum()) > 0 and n_in == int(came.sum()) > 0, (n_out, n_in)
    # band: the cells that came in lost their layer-1 entry; the cells that
    # left gained a layer nlayers+1 entry holding their old primary map
    cell1, c1 = f1["band_cell"], f1["band_c"]
    key0 = set(zip(f0["band_cell"].tolist(), f0["band_c"].tolist()))
    key1 = set(zip(cell1.tolist(), c1.tolist()))
    gone = {(int(i), 1 | (1 << LD)) for i in np.flatnonzero(came)}
    new = {(int(i), 1 | ((cfg.nlayers + 1) << LD)) for i in np.flatnonzero(left)}
    assert key1 == (key0 - gone) | new
    for e in np.flatnonzero(c1 >> LD == cfg.nlayers + 1):
        i = cell1[e]
        assert np.array_equal(f1["band_x"].reshape(3, -1)[:, e], x[:, i])
    o.close()


def test_redraw_two_spheres_least_phi_takes_the_cell():
    """Two spheres 0.01 apart; the smaller moves by h/2 towards the larger.
    Every cell's primary map afterwards belongs to the object of least phi
    among those eligible: the resident primary (phi <= 0) and any layer-1
    band map with phi < 0.  A displaced primary map is dropped."""
    objs = [configs.sphere((0.35, 0.5, 0.5), 0.2), configs.sphere((0.68, 0.5, 0.5), 0.12)]
    cfg, objs = _walled(14, objs, nlayers=4)
    o = OracleSim(cfg, objs)
    o.initialize()
    X, h = _coords(cfg)
    delta = np.array([-0.5 * h[0], 0.0, 0.0])
    f0, x, bx = _translate(o, cfg, [2], delta)
    Ng = o.Ng
    # eligibility from the state before the redraw
    cand = {}
    for ind in np.flatnonzero(f0["rm_c"]):
        oid = int(f0["rm_c"][ind])
        ph = _phi(objs, oid, x[:, ind:ind + 1])[0]
        if ph <= 0:
            cand.setdefault(int(ind), []).append((ph, oid, x[:, ind], "prim"))
    for e, (ind, c) in enumerate(zip(f0["band_cell"], f0["band_c"])):
        if c >> LD == 1:
            ph = _phi(objs, int(c & OM), bx[:, e:e + 1])[0]
            if ph < 0:
                cand.setdefault(int(ind), []).append((ph, int(c & OM), bx[:, e], "band"))
    swaps = 0
    for ind, lst in cand.items():
        assert sum(1 for t in lst if t[3] == "band") <= 1     # no cell needs the index-skip quirk
        if len(lst) == 2:
            swaps += 1
            prim = [t for t in lst if t[3] == "prim"][0]
            assert prim[0] >= -h[0]                           # not the fatal two-deep-maps case
    assert swaps > 0
    o.redraw()
    f1 = o.download()
    want_c = np.zeros(Ng, dtype=np.uint32)
    want_x = np.zeros((3, Ng))
    for ind, lst in cand.items():
        ph, oid, xv, _ = min(lst, key=lambda t: t[0])
        want_c[ind] = oid
        want_x[:, ind] = xv
    assert np.array_equal(f1["rm_c"], want_c)
    sel = want_c != 0
    assert np.array_equal(f1["rm_x"].reshape(3, -1)[:, sel], want_x[:, sel])
    o.close()


# ---------------------------------------------------------- corrector rule
def corner_sphere(wt_n, n=16):
    """A sphere close to the x = y = 0 edge of a walled box, so that its band
    reaches the edge ghost cells (i < 2 and j < 2), which no stencil of the
    step reads (normal derivatives and upwinding step along one axis from
    interior or face-ghost cells)."""
    return _walled(n, [configs.sphere((0.1, 0.1, 0.5), 0.09, G=1.0)], nlayers=4, wt_n=wt_n)


def remove_edge_ghost_entry(f, cfg, lid):
    """The band of download f without one layer-`lid` map in an edge ghost
    cell: (cell, code, x, xpred) arrays for set_band, and the removed index."""
    nb = len(f["band_cell"])
    m4, n4 = cfg.sys_size[0] + 4, cfg.sys_size[1] + 4
    cells = f["band_cell"]
    i, j = cells % m4, (cells // m4) % n4
    ok = np.flatnonzero((i < 2) & (j < 2) & (f["band_c"] >> LD == lid))
    assert len(ok)
    e = int(ok[0])
    keep = np.arange(nb) != e
    return (cells[keep], f["band_c"][keep], f["band_x"].reshape(3, nb)[:, keep],
            f["band_xpred"].reshape(3, nb)[:, keep]), e


@pytest.mark.parametrize("wt_n,lid,expect", [(4.0, 3, 1), (4.0, 4, 0), (1.0, 3, 0)])
def test_corrector_error_rule(wt_n, lid, expect):
    """update_reference_map (fluid_3d.cc:1205-1248): a band map of the new
    extrapolation with no map of its object in the old band is an error when
    its phi < eps and its layer <= 3.  The old band lacks one map in an edge
    ghost cell at phi ~ 1.6 h.  With eps = 4h a missing layer-3 map is an
    error, a missing layer-4 map is not (layer > 3); with eps = h a missing
    layer-3 map is not (phi > eps).  After the error,
    revert_reference_map_update (fluid_3d.cc:1257-1282) returns the primary
    maps to their value before the step and each band map with an old
    counterpart to the old value."""
    cfg, objs = corner_sphere(wt_n)
    o = OracleSim(cfg, objs)
    e0 = o.initialize()
    o.prepare_run(e0)
    assert o.step() == 0
    h = 1.0 / cfg.sys_size[0]
    f0 = o.download()
    band, e = remove_edge_ghost_entry(f0, cfg, lid)
    ph = _phi(objs, 1, f0["band_x"].reshape(3, -1)[:, e:e + 1])[0]
    eps = o.scalar(1)
    assert abs(ph - eps) > 0.3 * h
    o.set_band(*band)
    rc = o.step()
    assert rc == expect, (wt_n, lid, ph / h, rc)
    if expect:
        f1 = o.download()
        Ng = o.Ng
        prim = f0["rm_c"] != 0
        x0 = f0["rm_x"].reshape(3, Ng)[:, prim]
        x1 = f1["rm_x"].reshape(3, Ng)[:, prim]
        assert np.abs(x1 - x0).max() <= 4e-16 * np.abs(x0).max()
        old = {(int(c), int(k) & OM): f0["band_x"].reshape(3, -1)[:, a]
               for a, (c, k) in enumerate(zip(f0["band_cell"], f0["band_c"])) if a != e}
        nb1 = len(f1["band_cell"])
        bx1 = f1["band_x"].reshape(3, nb1)
        hits = 0
        for a, (c, k) in enumerate(zip(f1["band_cell"], f1["band_c"])):
            key = (int(c), int(k) & OM)
            if key in old:
                assert np.array_equal(bx1[:, a], old[key])
                hits += 1
        assert hits > 0.9 * nb1
    o.close()


# ------------------------------------------------------------- wall ghosts
def _wall_cfg():
    c = default_config()
    c.sys_size[0], c.sys_size[1], c.sys_size[2] = 8, 6, 5
    for i, v in enumerate([0.0, 1.0, -0.5, 0.25, 0.0, 0.5]):
        c.xyz_bounds[i] = v
    rng = np.random.default_rng(3)
    for f in range(6):
        c.bct[0][f], c.bct[1][f], c.bct[2][f] = 1, 2, (1 if f % 2 else 2)
        for comp in range(3):
            c.bcv[comp][f] = float(rng.uniform(-2, 2))
    c.fmu = 0.01
    c.sim_type = SIM_OBJECTS
    c.n_obj = 0
    c.dump_code = 0
    finalize_config(c)
    assert not any(c.prd)
    return c


def _face_geometry(cfg, f):
    X, h = _coords(cfg)
    nd = f // 2
    b = cfg.xyz_bounds
    s = X[nd] - b[2 * nd + (f % 2)]         # signed distance along +axis from the wall
    m = list(cfg.sys_size)
    I = np.stack(_unpack(cfg, np.arange(X.shape[1])))
    inner = np.ones(I.shape[1], dtype=bool)  # tangential indices of real cells
    for a in range(3):
        if a != nd:
            inner &= (I[a] >= 2) & (I[a] < m[a] + 2)
    return s, h[nd], I[nd], inner, m[nd]


@pytest.mark.parametrize("f", range(6))
def test_wall_ghost_cells_extend_linear_fields(f):
    """face_gm::set_ghost_cc (fields.hh:765-800): a Dirichlet component
    u = bcv + a s and a Neumann component u = A + bcv s (s the distance from
    the wall along +axis) are continued exactly into both ghost layers."""
    cfg = _wall_cfg()
    o = OracleSim(cfg, [])
    s, h, In, inner, mn = _face_geometry(cfg, f)
    rng = np.random.default_rng(10 + f)
    u = []
    for comp in range(3):
        bcv = cfg.bcv[comp][f]
        if cfg.bct[comp][f] == 1:
            u.append(bcv + rng.uniform(-1, 1) * s)
        else:
            u.append(rng.uniform(-1, 1) + bcv * s)
    o.upload(u=u[0], v=u[1], w=u[2])
    o.set_bc_cc()
    g = o.download(band=False)
    ghost = inner & ((In < 2) if f % 2 == 0 else (In >= mn + 2))
    assert ghost.sum() == 2 * np.prod([d for a, d in enumerate(cfg.sys_size) if a != f // 2])
    for comp, k in enumerate(("u", "v", "w")):
        assert np.allclose(g[k][ghost], u[comp][ghost], rtol=0, atol=1e-13), (f, comp)
    o.close()


@pytest.mark.parametrize("f", range(6))
def test_wall_face_values(f):
    """face_gm::set_ghost_cf (fields.hh:804-849): the wall face value is bcv
    for a Dirichlet component and, for a Neumann component, the value at the
    wall of the quadratic through the two outermost cells with slope bcv
    there, exact for u = A + bcv s + C s^2; it is stored on the wall face of
    the outermost real cell and of the inner ghost cell."""
    cfg = _wall_cfg()
    o = OracleSim(cfg, [])
    s, h, In, inner, mn = _face_geometry(cfg, f)
    rng = np.random.default_rng(20 + f)
    u, wall_val = [], []
    for comp in range(3):
        bcv = cfg.bcv[comp][f]
        A, Cq = rng.uniform(-1, 1), rng.uniform(-3, 3)
        u.append(A + bcv * s + Cq * s * s)
        wall_val.append(bcv if cfg.bct[comp][f] == 1 else A)
    o.upload(u=u[0], v=u[1], w=u[2])
    o.set_bc_cf()
    fv = o.download(band=False)["fvel"].reshape(6, 3, -1)
    nd = f // 2
    real = inner & (In == (2 if f % 2 == 0 else mn + 1))
    ghost = inner & (In == (1 if f % 2 == 0 else mn + 2))
    gface = 2 * nd + (1 if f % 2 == 0 else 0)
    for comp in range(3):
        assert np.allclose(fv[f, comp][real], wall_val[comp], rtol=0, atol=1e-13), (f, comp)
        assert np.allclose(fv[gface, comp][ghost], wall_val[comp], rtol=0, atol=1e-13), (f, comp)
    o.close()


# -------------------------------------------------- affine map, uniform flow
def test_affine_map_in_uniform_flow_is_advected_exactly():
    """A stress-free solid (G = 0) with an affine reference map
    xi0 = c + diag(e)(x - c) in a uniform periodic flow U: the velocity stays U
    and every map, primary and band, equals xi0(x - U t) after each step.
    Limited slopes and upwinded face states are exact on affine data, the
    least-squares extrapolation reproduces affine maps and the redraw moves
    whole maps, so this holds to rounding through cells changing sides."""
    n = 24
    c = default_config()
    for i in range(3):
        c.sys_size[i] = n
    c.fmu = 0.01
    c.fdt_pad = 0.4
    c.dt = 10.0
    c.sim_type = SIM_OBJECTS
    c.nlayers = 4
    c.wt_n = 1.0
    c.dump_code = 0
    c.n_obj = 1
    finalize_config(c)
    lam = 1.3
    ob = configs.sphere((0.5, 0.5, 0.5), 0.15, G=0.0, stretch_dir=0, sqrt_lam=math.sqrt(lam))
    o = OracleSim(c, [ob])
    o.initialize()
    U = np.array([0.4, -0.3, 0.2])
    Ng = o.Ng
    o.upload(u=np.full(Ng, U[0]), v=np.full(Ng, U[1]), w=np.full(Ng, U[2]))
    o.prepare_run(0)
    X, h = _coords(c)
    e = np.array([1 / lam, math.sqrt(lam), math.sqrt(lam)])
    ctr = np.full(3, 0.5)[:, None]
    flips = 0
    for _ in range(8):
        assert o.step() == 0
        flips += sum(o.step_counters()[1:])
        t = o.time
        f = o.download()
        for k, comp in zip(("u", "v", "w"), U):
            assert np.abs(f[k] - comp).max() < 1e-13
        want = ctr + e[:, None] * (X - U[:, None] * t - ctr)
        prim = f["rm_c"] != 0
        got = f["rm_x"].reshape(3, Ng)
        assert np.abs(got[:, prim] - want[:, prim]).max() < 1e-12
        nb = len(f["band_cell"])
        bw = want[:, f["band_cell"]]
        assert np.abs(f["band_x"].reshape(3, nb) - bw).max() < 1e-12
    assert flips > 0
    o.close()


# ------------------------------------------------- grad-phi weighted gathering
def test_gradphi_weighted_extrapolation_matches_numpy_fit():
    """gather_and_list + transfer (fluid_3d_ex.cc:86-328): every band map is
    the weighted affine least-squares fit, evaluated at its cell centre, of the
    same object's maps of lower layers within the search cube (radius 2);
    weights weight_fac^(|p|+|q|+|r|), times the cosine between the offset and
    the average face normal of the previous maps when the layer is <= 2 and
    all six face normals are known (non-positive weights dropped).  The
    normals, candidate set and fit are recomputed here with numpy."""
    cfg, objs = configs.stretched_sphere(16)
    o = OracleSim(cfg, objs)
    e0 = o.initialize()
    o.prepare_run(e0)
    for _ in range(2):
        assert o.step() == 0
    o.compute_stress()
    gp_prim = o.gphi_primary()
    rm_c0 = o.download(band=False)["rm_c"]
    o.extrap_now()
    pcell, pc, pg = o.gphi_prev_band()
    f = o.download()
    Ng = o.Ng
    m4, n4, o4 = (d + 4 for d in cfg.sys_size)
    X, h = _coords(cfg)
    wt = cfg.weight_fac
    assert np.array_equal(rm_c0, f["rm_c"])
    prim_xp = f["rm_xpred"].reshape(3, Ng)
    nb = len(f["band_cell"])
    bxp = f["band_xpred"].reshape(3, nb)
    new_first = {}
    for a in range(nb - 1, -1, -1):
        new_first[(int(f["band_cell"][a]), int(f["band_c"][a]) & OM)] = a
    prev_first = {}
    for a in range(len(pcell) - 1, -1, -1):
        prev_first[(int(pcell[a]), int(pc[a]) & OM)] = a

    def prev_row(ind, oid, row):
        if f["rm_c"][ind] == oid:
            return gp_prim[3 * row:3 * row + 3, ind]
        a = prev_first.get((ind, oid))
        return None if a is None else pg[3 * row:3 * row + 3, a]

    def cur(ind, oid):
        if f["rm_c"][ind] == oid:
            return 0, prim_xp[:, ind]
        a = new_first.get((ind, oid))
        return (None, None) if a is None else (int(f["band_c"][a]) >> LD, bxp[:, a])

    checked = weighted = 0
    for a in range(nb):
        ind, code = int(f["band_cell"][a]), int(f["band_c"][a])
        oid, L = code & OM, code >> LD
        i, j, k = _unpack(cfg, ind)
        gsum, nf = np.zeros(3), 0
        if 2 <= i <= m4 - 3 and 2 <= j <= n4 - 3 and 2 <= k <= o4 - 3:
            rows = [(ind, 0), (ind, 1), (ind, 2), (ind + 1, 0), (ind + m4, 1), (ind + m4 * n4, 2)]
            for c2, row in rows:
                r = prev_row(c2, oid, row)
                if r is not None and r[row] != -1.e200:
                    gsum += r
                    nf += np.any(r != 0)
        g = np.zeros(3)
        if nf == 6 and np.linalg.norm(gsum / 6) != 0:
            g = (gsum / 6) / np.linalg.norm(gsum / 6)
        use = np.linalg.norm(g) != 0 and L <= 2
        pts, vals, ws = [], [], []
        for r in range(-2, 3):
            for q in range(-2, 3):
                for p in range(-2, 3):
                    if not (0 <= i + p < m4 and 0 <= j + q < n4 and 0 <= k + r < o4):
                        continue
                    c2 = ind + p + m4 * (q + n4 * r)
                    lid, val = cur(c2, oid)
                    if lid is None or lid >= L:
                        continue
                    w = wt ** (abs(p) + abs(q) + abs(r))
                    if use:
                        v = X[:, ind] - X[:, c2]
                        w *= np.dot(v, g) / np.linalg.norm(v)
                        if w <= 0:
                            continue
                    pts.append(X[:, c2])
                    vals.append(val)
                    ws.append(w)
        if len(pts) < 8:
            continue
        P = np.array(pts)
        V = np.array(vals)
        sw = np.sqrt(np.array(ws))
        xc = X[:, ind]
        D = np.hstack([np.ones((len(P), 1)), P - xc])
        coef, *_ = np.linalg.lstsq(D * sw[:, None], V * sw[:, None], rcond=None)
        fit = coef[0]
        assert np.allclose(bxp[:, a], fit, rtol=1e-9, atol=1e-11), (a, L, use, bxp[:, a], fit)
        checked += 1
        weighted += use
    assert checked > 0.8 * nb
    assert weighted > 100, weighted
    o.close()
