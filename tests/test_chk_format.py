"""Checkpoint, recover.cfg and slice formats (CPU only).

Pins the oracle's restatement of fluid_3d_chk_pt.cc and fluid_3d_io.cc on
hand-made states whose answers follow from the reference text alone, and
checks the library's host-side writers (recover.cfg, .cfg checkpoint keys),
which need no GPU.
"""
import math
import os

import numpy as np
import pytest

from oracle.oracle import OracleSim, chk_read
from rmt3d_amd import _abi, configs
from rmt3d_amd import lib as rlib

NAMES = ("vx", "vy", "vz", "p", "X", "Y", "Z", "OID")


# ------------------------------------------------------------------ states
def small_cfg(m=4, n=3, o=2, prd_x=False):
    """A walled m x n x o box with two concentric spheres (R 1 and 2 around
    the origin), so that phi(x) = |x| - R is exact for the chosen points."""
    c = configs.default_config()
    c.sys_size[0], c.sys_size[1], c.sys_size[2] = m, n, o
    configs._walls(c)
    if prd_x:
        for comp in range(3):
            c.bct[comp][0] = c.bct[comp][1] = 0
    c.sim_type = _abi.SIM_OBJECTS
    c.n_obj = 2
    c.dt = 0.01
    c.dump_code = 0
    _abi.finalize_config(c)
    objs = [configs.sphere((0.0, 0.0, 0.0), 1.0), configs.sphere((0.0, 0.0, 0.0), 2.0)]
    return c, objs


def pad(cfg, i, j, k):
    m, n, o = cfg.sys_size
    return (i + 2) + (m + 4) * ((j + 2) + (n + 4) * (k + 2))


def hand_state(cfg):
    """Fields with distinct values per cell / node, and the map cases of
    min_phi_id (fluid_3d_io.cc:420-432) and copy_slice_to_buf's solid branch
    (:341-412) on the first rows of slice k = 0."""
    m, n, o = cfg.sys_size
    Ng = (m + 4) * (n + 4) * (o + 4)
    u = np.zeros(Ng); v = np.zeros(Ng); w = np.zeros(Ng); p = np.zeros(Ng)
    fm = m if cfg.prd[0] else m + 1
    fn = n if cfg.prd[1] else n + 1
    fo = o if cfg.prd[2] else o + 1
    for k in range(o):
        for j in range(n):
            for i in range(m):
                a = pad(cfg, i, j, k)
                u[a] = 1000 + i + 10 * j + 100 * k
                v[a] = -u[a]
                w[a] = 0.5 * u[a]
    for k in range(fo):
        for j in range(fn):
            for i in range(fm):
                p[pad(cfg, i, j, k)] = 5000 + i + 10 * j + 100 * k
    c = np.zeros(Ng, dtype=np.uint32)
    x = np.zeros((3, Ng))
    xp = np.zeros((3, Ng))

    def prim(i, j, k, oid, xv, lid=0):
        a = pad(cfg, i, j, k)
        c[a] = oid | (lid << 10)
        x[:, a] = xv
        xp[:, a] = np.asarray(xv) * 0.5

    prim(0, 0, 0, 1, (0.5, 0, 0))          # A: primary only
    prim(1, 0, 0, 1, (0.5, 0, 0))          # B: band entry of object 2 has smaller phi
    prim(0, 1, 0, 1, (0, 0.25, 0))         # E: band entry with equal phi (strict <)
    prim(1, 1, 0, 1, (0, 0, 0.5), lid=3)   # F: layer bits in the primary
    band = [  # (cell, c, x)
        (pad(cfg, 1, 0, 0), 2 | (1 << 10), (1.0, 0, 0)),      # phi -1 < -0.5
        (pad(cfg, 2, 0, 0), 1 | (2 << 10), (3.0, 0, 0)),      # C: phi 2, then ...
        (pad(cfg, 2, 0, 0), 2 | (2 << 10), (3.0, 0, 0)),      # ... phi 1 wins
        (pad(cfg, 0, 1, 0), 2 | (1 << 10), (0, 1.25, 0)),     # phi -0.75 = primary's
        (pad(cfg, 2, 1, 0), 2 | (4 << 10), (0, 0, 3.0)),      # G: band only
    ]
    band.sort(key=lambda e: e[0])
    bc = np.array([e[0] for e in band], dtype=np.int32)
    bcode = np.array([e[1] for e in band], dtype=np.uint32)
    bx = np.array([e[2] for e in band], dtype=np.float64).T.copy()
    bxp = bx * 0.25
    return dict(u=u, v=v, w=w, p=p, rm_c=c, rm_x=x.ravel(), rm_xpred=xp.ravel()), (bc, bcode, bx, bxp)


def oracle_with(cfg, objs):
    S = OracleSim(cfg, objs)
    st, band = hand_state(cfg)
    S.upload(**st)
    S.set_band(*band)
    return S, st


# ------------------------------------------------------------- write side
def test_chk_images_layout_and_buffer_reuse():
    # write_chk_pt (fluid_3d_chk_pt.cc:3-72): 8 files of o slices of
    # (m+1)(n+1) doubles; cell fields at [i + j*m], the pressure at
    # [i + j*fm]; one gathering buffer zeroed once, so the cell fields written
    # after the pressure keep the pressure's last slice past m*n
    cfg, objs = small_cfg()
    S, st = oracle_with(cfg, objs)
    img = S.chk_images()
    m, n, o = cfg.sys_size
    fm, fn = m + 1, n + 1
    L = (m + 1) * (n + 1)
    assert img.shape == (8, o, L)
    for k in range(o):
        for j in range(n):
            for i in range(m):
                a = pad(cfg, i, j, k)
                assert img[0, k, i + j * m] == st["u"][a]
                assert img[1, k, i + j * m] == st["v"][a]
                assert img[2, k, i + j * m] == st["w"][a]
        for j in range(fn):
            for i in range(fm):
                assert img[3, k, i + j * fm] == 5000 + i + 10 * j + 100 * k
        assert np.all(img[0:3, k, m * n:] == 0)
        last_p = np.array([5000 + (e % fm) + 10 * (e // fm) + 100 * (o - 1) for e in range(m * n, fm * fn)])
        for f in range(4, 8):
            assert np.array_equal(img[f, k, m * n:fm * fn], last_p)
            assert np.all(img[f, k, fm * fn:] == 0)


def test_chk_images_object_selection():
    # X/Y/Z/OID of min_phi_id + the solid branch of copy_slice_to_buf
    cfg, objs = small_cfg()
    S, _ = oracle_with(cfg, objs)
    img = S.chk_images()
    m = cfg.sys_size[0]

    def at(i, j):
        e = i + j * m
        return tuple(img[4:7, 0, e]), img[7, 0, e]

    assert at(0, 0) == ((0.5, 0, 0), 1)       # A: primary
    assert at(1, 0) == ((1.0, 0, 0), 0)       # B: smaller phi in the band, OID 0
    assert at(2, 0) == ((3.0, 0, 0), 0)       # C: last of two band entries
    assert at(3, 0) == ((100, 100, 100), 0)   # D: no map
    assert at(0, 1) == ((0, 0.25, 0), 1)      # E: equal phi keeps the primary
    assert at(1, 1) == ((0, 0, 0.5), 1)       # F: OID masks the layer bits
    assert at(2, 1) == ((0, 0, 3.0), 0)       # G: band only
    # every untouched cell of both slices reads 100 / 0
    for k in range(cfg.sys_size[2]):
        for j in range(cfg.sys_size[1]):
            for i in range(m):
                if k == 0 and (i, j) in ((0, 0), (1, 0), (2, 0), (0, 1), (1, 1), (2, 1)):
                    continue
                e = i + j * m
                assert tuple(img[4:7, k, e]) == (100, 100, 100) and img[7, k, e] == 0


def test_slice_types_on_hand_state():
    cfg, objs = small_cfg()
    S, st = oracle_with(cfg, objs)
    m, n, o = cfg.sys_size
    # OID / LID / phi / pr_phi / xpred of the same cases (types 12, 13, 8, 14, 15)
    lid = S.slice(2, 0, 13)
    assert lid[0, 0] == 0 and lid[0, 1] == 1 and lid[0, 2] == 2 and lid[1, 1] == 3 and lid[1, 2] == 4
    assert lid[0, 3] == 100
    phi = S.slice(2, 0, 8)
    assert phi[0, 0] == -0.5 and phi[0, 1] == -1.0 and phi[0, 2] == 1.0 and phi[1, 0] == -0.75
    prphi = S.slice(2, 0, 14)
    assert prphi[0, 0] == 0.25 - 1 and prphi[0, 1] == 0.25 - 2
    xpx = S.slice(2, 0, 15)
    assert xpx[0, 0] == 0.25 and xpx[0, 1] == 0.25 and xpx[0, 2] == 0.75
    # an explicit object id takes that object's map or 100 (0 for OID)
    x1 = S.slice(2, 0, 9, obj_id=1)
    assert x1[0, 1] == 0.5 and x1[0, 2] == 3.0 and x1[1, 2] == 100
    oid2 = S.slice(2, 0, 12, obj_id=2)
    assert oid2[0, 1] == 0 and oid2[1, 2] == 0
    # padding outputs: rho 0, J 100
    assert np.all(S.slice(0, 1, 20) == 0) and np.all(S.slice(0, 1, 21) == 100)
    # orientation: x-slices are (z, y), y-slices (z, x), z-slices (y, x)
    ux = S.slice(0, 2, 0)
    uy = S.slice(1, 1, 0)
    uz = S.slice(2, 1, 0)
    assert ux.shape == (o, n) and uy.shape == (o, m) and uz.shape == (n, m)
    for k in range(o):
        for j in range(n):
            assert ux[k, j] == 1000 + 2 + 10 * j + 100 * k
        for i in range(m):
            assert uy[k, i] == 1000 + i + 10 + 100 * k
    # pressure is a corner field: node counts in walled dimensions
    assert S.slice(2, 0, 7).shape == (n + 1, m + 1)
    assert S.slice(0, 0, 7).shape == (o + 1, n + 1)
    # a refused point leaves the zeroed slice (copy_slice_to_buf returns 0)
    assert np.all(S.slice(2, o, 0) == 0)


def test_slice_vorticity_and_magnitude_closed_forms():
    # the 8-cell node stencils (fluid_3d_io.cc:291-318) are exact on fields
    # linear in the cell index; dx = 1/4 makes every product exact
    cfg, objs = small_cfg(4, 4, 4)
    S = OracleSim(cfg, objs)
    m, n, o = cfg.sys_size
    Ng = (m + 4) * (n + 4) * (o + 4)
    kk, jj, ii = np.meshgrid(np.arange(o + 4), np.arange(n + 4), np.arange(m + 4), indexing="ij")
    i, j, k = (ii - 2).ravel().astype(float), (jj - 2).ravel().astype(float), (kk - 2).ravel().astype(float)
    # u = 2j + 5k, v = 3k + 7i, w = 11i + 13j  ->  curl = (13-3, 5-11, 7-2) / h
    S.upload(u=2 * j + 5 * k, v=3 * k + 7 * i, w=11 * i + 13 * j)
    for t, want in ((3, 10 * 4.0), (4, -6 * 4.0), (5, 5 * 4.0)):
        sl = S.slice(2, 1, t)
        assert np.all(sl[1:, 1:] == want), (t, sl)
    S.upload(u=np.full(Ng, 3.0), v=np.full(Ng, 4.0), w=np.zeros(Ng))
    assert np.all(S.slice(1, 2, 6) == 5.0)


def test_slice_div_edges_closed_form():
    # o_type 19: dxsp (fvel[1][0]-fvel[0][0]) + dysp (...) + dzsp (...)
    cfg, objs = small_cfg(4, 4, 4)
    S = OracleSim(cfg, objs)
    Ng = S.Ng
    fv = np.zeros((18, Ng))
    fv[3] = 1.0      # face 1, x component
    fv[10] = 0.5     # face 3, y component
    fv[14] = -0.25   # face 4, z component
    S.upload(fvel=fv.ravel())
    assert np.all(S.slice(2, 0, 19) == 4 * 1.0 + 4 * 0.5 + 4 * 0.25)


# -------------------------------------------------------------- read side
def test_restore_uses_cell_stride_for_pressure():
    # copy_chk_slice_to_buf reads gp[i + j*m] for the pressure as well
    # (fluid_3d_chk_pt.cc:303-316); with walls in x the writer used m+1
    cfg, objs = small_cfg()
    m, n, o = cfg.sys_size
    L = (m + 1) * (n + 1)
    rng = np.random.default_rng(3)
    img = rng.standard_normal((8, o, L))
    img[7] = rng.integers(0, 3, size=(o, L)).astype(float)
    S = OracleSim(cfg, objs)
    S.chk_restore(img)
    f = S.download(band=False)
    x = f["rm_x"].reshape(3, -1)
    xp = f["rm_xpred"].reshape(3, -1)
    for k in range(o):
        for j in range(n + 1):
            for i in range(m + 1):
                assert f["p"][pad(cfg, i, j, k)] == img[3, k, i + j * m]
        for j in range(n):
            for i in range(m):
                a, e = pad(cfg, i, j, k), i + j * m
                assert f["u"][a] == img[0, k, e] and f["v"][a] == img[1, k, e] and f["w"][a] == img[2, k, e]
                assert tuple(x[:, a]) == tuple(img[4:7, k, e]) == tuple(xp[:, a])
                assert f["rm_c"][a] == int(img[7, k, e])
    # only o pressure slices are stored: the top node layer is not restored
    for j in range(n + 1):
        for i in range(m + 1):
            assert f["p"][pad(cfg, i, j, o)] == 0


def test_round_trip_periodic_x():
    # with x periodic fm = m and the two strides agree: write then read
    # reproduces every stored field on the interior
    cfg, objs = small_cfg(4, 3, 2, prd_x=True)
    S, st = oracle_with(cfg, objs)
    img = S.chk_images()
    T = OracleSim(cfg, objs)
    T.chk_restore(img)
    f = T.download(band=False)
    m, n, o = cfg.sys_size
    x = f["rm_x"].reshape(3, -1)
    for k in range(o):
        for j in range(n):
            for i in range(m):
                a = pad(cfg, i, j, k)
                for key in ("u", "v", "w"):
                    assert f[key][a] == st[key][a]
        for j in range(n + 1):
            for i in range(m):
                assert f["p"][pad(cfg, i, j, k)] == st["p"][pad(cfg, i, j, k)]
    assert tuple(x[:, pad(cfg, 0, 0, 0)]) == (0.5, 0, 0) and f["rm_c"][pad(cfg, 0, 0, 0)] == 1
    assert tuple(x[:, pad(cfg, 3, 0, 0)]) == (100, 100, 100) and f["rm_c"][pad(cfg, 3, 0, 0)] == 0
    # F's layer bits do not survive (OID is the object id only)
    assert f["rm_c"][pad(cfg, 1, 1, 0)] == 1


def test_reader_status_codes(tmp_path):
    # initialize_from_chk_point: a file that cannot be opened -> 1, a slice
    # read that returns nothing -> 2 (fluid_3d_chk_pt.cc:115-150); a short
    # read keeps the previous slice in the one reading buffer
    dims = (4, 3, 2)
    L = 5 * 4
    rc, _ = chk_read(str(tmp_path), dims)
    assert rc == 1
    full = np.arange(2 * L, dtype=np.float64)
    for name in NAMES:
        full.tofile(tmp_path / name)
    rc, img = chk_read(str(tmp_path), dims)
    assert rc == 0 and np.array_equal(img[3].ravel(), full)
    (tmp_path / "Y").write_bytes(b"")
    rc, _ = chk_read(str(tmp_path), dims)
    assert rc == 2
    full[:L + 3].tofile(tmp_path / "Y")
    rc, img = chk_read(str(tmp_path), dims)
    assert rc == 0
    assert np.array_equal(img[5, 1, :3], full[L:L + 3]) and np.array_equal(img[5, 1, 3:], full[3:L])


def test_oracle_restart_from_its_own_checkpoint(tmp_path):
    # a run stopped after 3 steps, written, and resumed through the .cfg
    # keys: the resumed state holds the written velocity, pressure (x is
    # periodic here, so the pressure stride is harmless) and primary maps
    c, objs = configs.cavity_sphere(16, steps=3)
    for comp in range(3):
        c.bct[comp][0] = c.bct[comp][1] = 0
    _abi.finalize_config(c)
    S = OracleSim(c, objs)
    rc, dt, tot = S.prepare_run(S.initialize())
    for _ in range(3):
        assert S.step() == 0
    d = tmp_path / "run.out" / "chk.00001"
    d.mkdir(parents=True)
    S.write_chk_pt(str(d))
    c2 = _abi.Config.from_buffer_copy(c)
    c2.chk_num, c2.chk_step, c2.current_time = 1, 3, S.time
    c2.chk_dirname = str(d).encode()
    T = OracleSim(c2, objs)
    assert T.initialize() == 0
    assert T.time == S.time and T.nt == 3
    a, b = S.download(band=False), T.download(band=False)
    m, n, o = c.sys_size
    msk = np.zeros(S.Ng, dtype=bool)
    for k in range(o):
        for j in range(n):
            for i in range(m):
                msk[pad(c, i, j, k)] = True
    for key in ("u", "v", "w", "p"):
        assert np.array_equal(a[key][msk], b[key][msk]), key
    solid = msk & (a["rm_c"] > 0)
    assert solid.any() and np.array_equal(a["rm_c"][solid], b["rm_c"][solid])
    xa, xb = a["rm_x"].reshape(3, -1), b["rm_x"].reshape(3, -1)
    assert np.array_equal(xa[:, solid], xb[:, solid])
    # the band is rebuilt by extrapolation and the run continues
    assert len(T.download()["band_cell"]) > 0
    rc, dt2, tot2 = T.prepare_run(0)
    assert rc == 0 and tot2 == 3
    assert T.step() == 0


def test_missing_checkpoint_initialises_from_scratch(tmp_path):
    c, objs = configs.cavity_sphere(8, steps=2)
    c.chk_num = 2
    c.chk_dirname = str(tmp_path / "nowhere").encode()
    T = OracleSim(c, objs)
    assert T.initialize() == 1
    assert T.chk_num == 0 and T.time == 0


# --------------------------------------------------- library host writers
@pytest.fixture(scope="module")
def library():
    if not os.path.exists(rlib.LIB_PATH):
        pytest.skip("librmt3d.so not built")
    return rlib


def expected_recover_cfg(c, objs):
    """sim_params::write_params (Sim/sim_params.cc:489-606) line by line."""
    g = lambda v: "%g" % v
    st = {0: "ftest", 1: "stest", 2: "objects", 3: "object_mills"}[c.sim_type]
    L = ["#Keyword                  Input(default)           Type", "#System parameters"]
    rows = [("suppress", int(bool(c.suppress)), "(0)", "bool", 19), ("display", int(bool(c.display)), "(0)", "bool", 19),
            ("implicit", int(bool(c.impl)), "(0)", "bool", 19), ("godunov", int(bool(c.godunov)), "(0)", "bool", 19),
            ("x_prd", int(bool(c.prd[0])), "(0)", "bool", 19), ("y_prd", int(bool(c.prd[1])), "(0)", "bool", 19),
            ("z_prd", int(bool(c.prd[2])), "(0)", "bool", 19), ("has_exact", int(bool(c.has_exact)), "(0)", "bool", 19)]
    for k, v, d, t, w in rows:
        L.append("%-26s%d#%s%s%s" % (k, v, d, " " * w, t))
    L.append("recover_dirname           %s#                      string" % c.dirname.decode())
    L.append("chk_num                   %d#(0)                   int" % c.chk_num)
    L.append("chk_step                  %d#(0)                   int" % c.chk_step)
    L.append("sys_size                  %d %d %d#({96,96,96})    int" % tuple(c.sys_size))
    L.append("frames                    %d#(96)                  int" % c.frames)
    L.append("stop_short                %d#(96)                  int" % c.stop_short)
    L.append("chkpt_freq                %d#(1)                   int" % c.chkpt_freq)
    L.append("omp_num_thr               %d#(1)                   int" % c.omp_num_thr)
    L.append("debug_flag                %d#(0)                   int" % c.debug_flag)
    L.append("debug_obj_id              %d#(-1)                   int" % c.debug_obj_id)
    L.append("obj_body                  %d#(-1)                  int" % c.obj_body)
    L.append("data_fmt                  %d#(0)                   int" % c.data_fmt)
    L.append("out_flag                  %d#(135)                 int" % c.out_flag)
    L.append("output_dim                %d#0/1/2(2)              int" % c.output_dim)
    L.append("output_ind                %d#(48)                  int" % c.output_ind)
    L.append("dump_code                 %d#(3)                   unsigned int" % c.dump_code)
    L.append("ntracers                  %d#(1000)                int" % c.ntracers)
    L.append("num_iters                 %d#(5)                   int" % c.num_iters)
    L.append("dt                        %s#(1.)                  double" % g(c.dt))
    L.append("T                         %s#(1.)                  double" % g(c.T))
    L.append("current_time              %s#(0.)                  double" % g(c.current_time))
    L.append("xyz_bounds                %s" % " ".join(g(v) for v in c.xyz_bounds))
    L.append("                          #(0 1 0 1 0 1)         doubles")
    L.append("# SIMULATION TYPE CONFIGURATIONS")
    L.append('sim_type                  %s#("ftest")           string' % st)
    L.append("# FLUID PROPERTIES")
    L.append("fmu                       %s#(0.002)               double" % g(c.fmu))
    L.append("fdt_pad                   %s#(0.5)                 double" % g(c.fdt_pad))
    L.append("# SOLID PROPERTIES")
    L.append("n_obj                     %d#(0)                   int" % c.n_obj)
    L.append("nlayers                   %d#(8)                   int" % c.nlayers)
    for k, v, d in (("wt_n", c.wt_n, "(2.5)  "), ("ex_visc_mult", c.ex_visc_mult, "(1.)   "),
                    ("ev_trans_mult", c.ev_trans_mult, "(1.)   "), ("sdt_pad", c.sdt_pad, "(0.25) "),
                    ("dt_ex_pad", c.dt_ex_pad, "(0.8)  "), ("gravity", c.gravity, "(10.)  ")):
        L.append("%-26s%s#%s               double" % (k, g(v), d))
    L.append("# SOLIDS CONFIGURATIONS")
    n = c.n_obj
    for i in range(n):
        L.append("object_list               %d %d" % (i, objs[i].type))
    for i in range(n):
        L.append("basic_specs               %d " % i + "".join(g(v) + " " for v in objs[i].basic))
    for i in range(n):
        L.append("extra_specs               %d " % i + "".join(g(v) + " " for v in objs[i].extra))
    for i in range(n):
        L.append("srho                      %d %s" % (i, g(objs[i].srho)))
    for i in range(n):
        L.append("shear_mod                 %d %s" % (i, g(objs[i].shear_mod)))
    L.append("vel_prof_num              %d#(1)                   int" % c.vel_prof_num)
    for i in range(c.vel_prof_num):
        L.append("vel_profile               %d %d %s" % (i, int(c.vel_profile[2 * i]), g(c.vel_profile[2 * i + 1])))
    L.append("# DOMAIN BOUNDARY CONDITIONS")
    for f, name in enumerate(("bc_xl", "bc_xh", "bc_yl", "bc_yh", "bc_zl", "bc_zh")):
        L.append("%s                     %d %d %d %s %s %s" % (name, c.bct[0][f], c.bct[1][f], c.bct[2][f],
                                                               g(c.bcv[0][f]), g(c.bcv[1][f]), g(c.bcv[2][f])))
    L.append("# WALL BOUNDARIES")
    L.append("walls                     " + " ".join(str(int(bool(c.walls[i]))) for i in range(6)))
    L.append("wall_pos                  " + " ".join(g(c.wall_pos[i]) for i in range(6)))
    L.append("wall_dist                 %s" % g(c.wall_dist))
    return "\n".join(L) + "\n"


def test_recover_cfg_text(library, tmp_path):
    c, objs = configs.cavity_sphere(32, steps=7)
    c.dirname = b"/data/cavity.out"
    c.chk_num, c.chk_step, c.current_time = 4, 123, 0.123456789
    c.vel_prof_num = 1
    c.vel_profile[0], c.vel_profile[1] = 2.0, 0.75
    library.write_params(c, objs, str(tmp_path))
    got = (tmp_path / "recover.cfg").read_text()
    assert got == expected_recover_cfg(c, objs)


def test_recover_cfg_reparses(library, tmp_path):
    # recover.cfg is read back by the same parser: every key survives to
    # %g precision, and the checkpoint directory is found through
    # recover_dirname (sim_params.cc:374-398)
    c, objs = configs.cavity_sphere(32, steps=7)
    out = tmp_path / "cavity.out"
    chk = out / "chk.00004"
    chk.mkdir(parents=True)
    c.dirname = str(out).encode()
    c.chk_num, c.chk_step, c.current_time = 4, 123, 0.125
    library.write_params(c, objs, str(chk))
    c2, o2 = library.parse_cfg(str(chk / "recover.cfg"))
    assert c2.chk_num == 4 and c2.chk_step == 123 and c2.current_time == 0.125
    assert c2.recover_dirname == str(out).encode()
    assert c2.chk_dirname == str(chk).encode()
    assert c2.dirname == str(chk / "recover.out").encode()
    for k in ("sys_size", "prd", "walls", "frames", "stop_short", "nlayers", "sim_type", "n_obj"):
        a, b = getattr(c, k), getattr(c2, k)
        assert (list(a) == list(b)) if hasattr(a, "__len__") else a == b, k
    for k in ("fmu", "fdt_pad", "T", "wt_n", "ex_visc_mult", "ev_trans_mult", "sdt_pad", "dt_ex_pad", "wall_dist"):
        assert getattr(c2, k) == float("%g" % getattr(c, k)), k
    assert [list(r) for r in c2.bct] == [list(r) for r in c.bct]
    assert [o.basic[3] for o in o2] == [o.basic[3] for o in objs]


def test_cfg_checkpoint_lookup(library, tmp_path):
    # chk_num > 0 without <recover_dirname>/chk.%05d/recover.cfg resets the
    # checkpoint keys; dirname follows the .cfg name
    c, objs = configs.cavity_sphere(16, steps=2)
    path = tmp_path / "case.cfg"
    configs.write_cfg(c, objs, str(path))
    with open(path, "a") as f:
        f.write("chk_num 3\nchk_step 40\ncurrent_time 0.5\n")
    c2, _ = library.parse_cfg(str(path))
    assert c2.dirname == str(tmp_path / "case.out").encode()
    assert c2.chk_num == 0 and c2.chk_step == 0 and c2.chk_dirname == b""
    assert c2.current_time == 0.5
    chk = tmp_path / "case.out" / "chk.00003"
    chk.mkdir(parents=True)
    (chk / "recover.cfg").write_text("# present\n")
    c3, _ = library.parse_cfg(str(path))
    assert c3.chk_num == 3 and c3.chk_step == 40 and c3.chk_dirname == str(chk).encode()


def test_output_keys_defaults(library, tmp_path):
    # sim_params.cc:12-20 and the output_ind fallback at :327
    c, objs = configs.cavity_sphere(16, steps=2)
    path = tmp_path / "o.cfg"
    configs.write_cfg(c, objs, str(path))
    c2, _ = library.parse_cfg(str(path))
    assert (c2.chkpt_freq, c2.omp_num_thr, c2.debug_obj_id, c2.obj_body) == (1, 1, -1, -1)
    assert (c2.out_flag, c2.output_dim, c2.output_ind) == (135, 2, 8)
    with open(path, "a") as f:
        f.write("output_dim 0\noutput_ind 99\nout_flag 7\n")
    c3, _ = library.parse_cfg(str(path))
    assert (c3.out_flag, c3.output_dim, c3.output_ind) == (7, 0, 8)


def test_xz_helpers_round_trip(library, tmp_path):
    for name in NAMES:
        (tmp_path / name).write_bytes(np.arange(10, dtype=np.float64).tobytes())
    library.compress_checkpoint(str(tmp_path))
    assert all((tmp_path / (n + ".xz")).exists() and not (tmp_path / n).exists() for n in NAMES)
    library.decompress_checkpoint(str(tmp_path))
    assert np.array_equal(np.fromfile(tmp_path / "vx"), np.arange(10.0))
