"""Tracers on the HIP path against the oracle: the same float positions
after the same steps (bit for bit), through walls, periodic wraps, a
checkpoint restart and the driver's tr files."""
import os
import subprocess

import numpy as np
import pytest

from oracle.oracle import OracleSim
from rmt3d_amd import _abi, configs
from rmt3d_amd import lib as rlib
from rmt3d_amd.lib import Simulation

pytestmark = pytest.mark.gpu

DRIVER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "rmt3d_amd", "run_sim")


def _pair(cfg, objs, steps):
    g, o = Simulation(cfg, objs), OracleSim(cfg, objs)
    eg, eo = g.initialize(), o.initialize()
    assert eg == eo
    g.prepare_run(eg)
    o.prepare_run(eo)
    for _ in range(steps):
        assert g.step() == 0 and o.step() == 0
    return g, o


@pytest.mark.parametrize("case", ["cavity", "taylor_green"])
def test_tracers_match_oracle(case):
    if case == "cavity":
        cfg, objs = configs.cavity_sphere(24)
    else:
        cfg, objs = configs.taylor_green(32)
    cfg.dump_code = 1
    cfg.ntracers = 500
    print("tracers", case, flush=True)
    g, o = _pair(cfg, objs, 6)
    a, b = g.tracer_positions(), o.tracer_positions()
    assert a.shape == (500, 3)
    assert np.array_equal(a, b)
    assert np.count_nonzero(np.any(a != 0, axis=1)) == 500


def test_tracer_file_and_restart(tmp_path):
    cfg, objs = configs.cavity_sphere(24)
    cfg.dump_code = 1
    cfg.ntracers = 300
    out = tmp_path / "run.out"
    cfg.dirname = str(out).encode()
    g, o = _pair(cfg, objs, 4)
    out.mkdir()
    g.write_tracers(str(out / "tr.00001"))
    g.write_tracers(str(out / "tr_st.00001"), binary=False)
    raw = (out / "tr.00001").read_bytes()
    assert np.frombuffer(raw[:4], dtype=np.int32)[0] == 300
    pos = np.frombuffer(raw[4:], dtype=np.float32).reshape(300, 3)
    assert np.array_equal(pos, o.tracer_positions())
    lines = (out / "tr_st.00001").read_text().strip().split("\n")
    assert len(lines) == 300 and float(lines[0].split()[0]) == pytest.approx(float(pos[0, 0]), rel=1e-5)
    # the checkpoint takes the frame's tracer file along, and a resumed run
    # starts its tracers there
    chk = out / "chk.00001"
    chk.mkdir()
    g.write_checkpoint(4, 1, str(chk))
    assert (chk / "tr.00001").read_bytes() == raw
    c2, o2 = rlib.parse_cfg(str(chk / "recover.cfg"))
    h, q = Simulation(c2, o2), OracleSim(c2, o2)
    assert h.initialize() == 0 and q.initialize() == 0
    assert np.array_equal(h.tracer_positions(), q.tracer_positions())
    h.prepare_run(0)
    q.prepare_run(0)
    for _ in range(3):
        assert h.step() == 0 and q.step() == 0
    assert np.array_equal(h.tracer_positions(), q.tracer_positions())


def test_run_sim_tracer_files(tmp_path):
    cfg, objs = configs.cavity_sphere(24)
    cfg.dump_code = 1
    path = str(tmp_path / "t.cfg")
    configs.write_cfg(cfg, objs, path)
    r = subprocess.run([DRIVER, path, "--steps", "20", "--quiet"], capture_output=True, text=True, timeout=300)
    print(r.stdout[-2000:], r.stderr[-2000:])
    assert r.returncode == 0
    for f in range(3):
        raw = (tmp_path / "t.out" / ("tr.%05d" % f)).read_bytes()
        assert np.frombuffer(raw[:4], dtype=np.int32)[0] == 1000      # the default ntracers
        assert len(raw) == 4 + 12 * 1000
        assert (tmp_path / "t.out" / ("tr_st.%05d" % f)).exists()
