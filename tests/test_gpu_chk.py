"""Checkpoint, restart and slice output of the HIP path against the oracle.

The files written from device state must equal the oracle's byte for byte
(the FP64 step is bit-identical, and so must be the gather); a run resumed
from those files must continue bit-identically on both; every slice output
type must agree exactly.
"""
import numpy as np
import pytest

from oracle.oracle import OracleSim
from rmt3d_amd import _abi, configs
from rmt3d_amd import lib as rlib
from rmt3d_amd.lib import Simulation

pytestmark = pytest.mark.gpu

NAMES = ("vx", "vy", "vz", "p", "X", "Y", "Z", "OID")


def run_pair(cfg, objs, steps):
    g, o = Simulation(cfg, objs), OracleSim(cfg, objs)
    eg, eo = g.initialize(), o.initialize()
    assert eg == eo
    rg, dtg, _ = g.prepare_run(eg)
    ro, dto, _ = o.prepare_run(eo)
    assert rg == ro == 0 and dtg == dto
    for s in range(steps):
        assert g.step() == 0 and o.step() == 0
    return g, o


def same_state(g, o):
    fg, fo = g.download(), o.download()
    for k in ("u", "v", "w", "p", "rm_x", "rm_xpred"):
        assert np.array_equal(fg[k], fo[k]), k
    for k in ("rm_c", "band_cell", "band_c", "band_x", "band_xpred"):
        assert np.array_equal(fg[k], fo[k]), k


def sphere_case(n=24, periodic_x=False):
    c, objs = configs.cavity_sphere(n, steps=12)
    if periodic_x:
        for comp in range(3):
            c.bct[comp][0] = c.bct[comp][1] = 0
        _abi.finalize_config(c)
    return c, objs


@pytest.mark.parametrize("periodic_x", [False, True])
def test_checkpoint_files_match_oracle(tmp_path, periodic_x):
    cfg, objs = sphere_case(periodic_x=periodic_x)
    print("checkpoint files, periodic_x", periodic_x, flush=True)
    g, o = run_pair(cfg, objs, 4)
    dg, do = tmp_path / "gpu", tmp_path / "oracle"
    dg.mkdir()
    do.mkdir()
    g.write_checkpoint(4, 1, str(dg))
    o.write_chk_pt(str(do))
    m, n, k = cfg.sys_size
    for name in NAMES:
        a, b = (dg / name).read_bytes(), (do / name).read_bytes()
        assert len(a) == 8 * k * (m + 1) * (n + 1)
        assert a == b, name
    oid = np.fromfile(dg / "OID").reshape(k, -1)[:, :m * n]
    assert set(np.unique(oid)) <= {0.0, 1.0} and (oid == 1).any()
    text = (dg / "recover.cfg").read_text()
    assert "chk_num                   1#" in text and "chk_step                  4#" in text


@pytest.mark.parametrize("periodic_x", [False, True])
def test_restart_continues_like_oracle(tmp_path, periodic_x):
    cfg, objs = sphere_case(periodic_x=periodic_x)
    print("restart, periodic_x", periodic_x, flush=True)
    # the context's own config carries dirname into recover.cfg
    cfg.dirname = str(tmp_path / "run.out").encode()
    g, o = run_pair(cfg, objs, 4)
    chk = tmp_path / "run.out" / "chk.00001"
    chk.mkdir(parents=True)
    g.write_checkpoint(4, 1, str(chk))
    c2, o2 = rlib.parse_cfg(str(chk / "recover.cfg"))
    assert c2.chk_num == 1 and c2.chk_step == 4 and c2.chk_dirname == str(chk).encode()
    h, q = Simulation(c2, o2), OracleSim(c2, o2)
    assert h.initialize() == 0 and q.initialize() == 0
    assert h.time == q.time
    same_state(h, q)
    rh, dth, toth = h.prepare_run(0)
    rq, dtq, totq = q.prepare_run(0)
    assert rh == rq == 0 and dth == dtq and toth == totq
    for s in range(4):
        assert h.step() == 0 and q.step() == 0
    same_state(h, q)
    assert h.time == q.time
    # recover.cfg holds dt and the time to %g: the resumed stepper dt is
    # rebuilt from the rounded value, as the reference's is
    assert h.dt == q.dt and abs(h.time - g.time - 4 * h.dt) < 1e-6


def test_slices_match_oracle(tmp_path):
    cfg, objs = sphere_case()
    g, o = run_pair(cfg, objs, 3)
    m, n, k = cfg.sys_size
    for t in range(22):
        for dim in range(3):
            point = cfg.sys_size[dim] // 2 - 1
            a = g.slice(dim, point, t)
            b = o.slice(dim, point, t)
            assert a.shape == b.shape, (t, dim)
            assert np.array_equal(a, b, equal_nan=True), (t, dim)
    for t, oid in ((9, 1), (12, 1), (8, 0)):
        assert np.array_equal(g.slice(2, k // 2, t, obj_id=oid), o.slice(2, k // 2, t, obj_id=oid), equal_nan=True)
    # the three file formats of write_slice
    ref = o.slice(2, k // 2, 7)
    d2, d1 = ref.shape
    h = [1.0 / m, 1.0 / n]
    gx = [cfg.xyz_bounds[0] + (i + 0.5) * h[0] for i in range(d1)]
    gy = [cfg.xyz_bounds[2] + (j + 0.5) * h[1] for j in range(d2)]
    g.write_slice(2, k // 2, 7, str(tmp_path / "p0"), fmt=0)
    want = "".join("".join("%17.14g " % ref[j, i] for i in range(d1)) + "\n" for j in range(d2))
    assert (tmp_path / "p0").read_text() == want
    g.write_slice(2, k // 2, 7, str(tmp_path / "p1"), fmt=1)
    want = "".join("".join("%17.14f %17.14f %17.14f\n" % (gx[i], gy[j], ref[j, i]) for i in range(d1)) + "\n"
                   for j in range(d2))
    assert (tmp_path / "p1").read_text() == want
    g.write_slice(2, k // 2, 7, str(tmp_path / "p2"), fmt=2)
    rows = [np.array([m] + gx, dtype=np.float32)]
    rows += [np.array([gy[j]] + list(ref[j]), dtype=np.float32) for j in range(d2)]
    assert (tmp_path / "p2").read_bytes() == np.concatenate(rows).tobytes()


def test_taylor_green_injected_through_checkpoint(tmp_path):
    # the reference has no Taylor-Green sim_type: its own recipe writes the
    # field as checkpoint files and resumes from them (chk_num,
    # recover_dirname).  Files made here with numpy, read by both.
    cfg, objs = configs.taylor_green(32, steps=20)
    out = tmp_path / "tg.out"
    chk = out / "chk.00001"
    chk.mkdir(parents=True)
    m, n, k = cfg.sys_size
    L = (m + 1) * (n + 1)
    u, v, w = configs.taylor_green_field(cfg)
    cell = lambda f: f.reshape(k + 4, n + 4, m + 4)[2:k + 2, 2:n + 2, 2:m + 2].reshape(k, m * n)
    img = np.zeros((8, k, L))
    img[0, :, :m * n], img[1, :, :m * n], img[2, :, :m * n] = cell(u), cell(v), cell(w)
    img[4:7, :, :m * n] = 100.0
    for f, name in enumerate(NAMES):
        img[f].tofile(chk / name)
    cfg.dirname = str(out).encode()
    cfg.chk_num = 1
    rlib.write_params(cfg, objs, str(chk))
    c2, o2 = rlib.parse_cfg(str(chk / "recover.cfg"))
    assert c2.chk_dirname == str(chk).encode()
    g, o = Simulation(c2, o2), OracleSim(c2, o2)
    assert g.initialize() == 0 and o.initialize() == 0
    rg, dtg, totg = g.prepare_run(0)
    ro, dto, toto = o.prepare_run(0)
    assert rg == ro == 0 and dtg == dto and totg == toto == 20
    for s in range(totg):
        assert g.step() == 0 and o.step() == 0
    fg, fo = g.download(band=False), o.download(band=False)
    for key in ("u", "v", "w", "p"):
        assert np.array_equal(fg[key], fo[key]), key
    # the field decays as Taylor-Green does: kinetic energy falls, the
    # pattern stays
    msk = np.zeros(g.Ng, dtype=bool)
    msk.reshape(k + 4, n + 4, m + 4)[2:k + 2, 2:n + 2, 2:m + 2] = True
    r = np.dot(fg["u"][msk], u[msk]) / np.dot(u[msk], u[msk])
    assert 0.5 < r < 1.0
