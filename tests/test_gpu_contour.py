"""Contours on the HIP path against the oracle: the same vertices
(positions and normals as floats), the same triangles, and the ctr file
layout of fluid_3d::output_contours."""
import os
import subprocess

import numpy as np
import pytest

from oracle.oracle import OracleSim
from rmt3d_amd import configs
from rmt3d_amd.lib import Simulation

pytestmark = pytest.mark.gpu

DRIVER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "rmt3d_amd", "run_sim")


@pytest.mark.parametrize("case", ["cavity", "stretched", "two"])
def test_contours_match_oracle(case, tmp_path):
    if case == "cavity":
        cfg, objs = configs.cavity_sphere(24)
        steps = 5
    elif case == "stretched":
        cfg, objs = configs.stretched_sphere(32)
        steps = 2
    else:
        cfg, objs = configs.cavity_sphere(32)
        objs = [configs.sphere((0.3, 0.5, 0.5), 0.15), configs.sphere((0.7, 0.45, 0.55), 0.12)]
        cfg.n_obj = 2
        steps = 3
    print("contours", case, flush=True)
    g, o = Simulation(cfg, objs), OracleSim(cfg, objs)
    g.prepare_run(g.initialize())
    o.prepare_run(o.initialize())
    for _ in range(steps):
        assert g.step() == 0 and o.step() == 0
    for ob in range(1, cfg.n_obj + 1):
        a, b = g.contour(ob), o.contour(ob)
        for x, y in zip(a, b):
            assert x.shape == y.shape and np.array_equal(x, y)
        assert len(a[0]) > 50
    fn = tmp_path / "ctr.00001"
    g.write_contours(str(fn))
    raw = fn.read_bytes()
    assert np.frombuffer(raw[:4], dtype=np.int32)[0] == cfg.n_obj
    off = 4
    for ob in range(1, cfg.n_obj + 1):
        nv, nt = np.frombuffer(raw[off:off + 8], dtype=np.int32)
        off += 8
        pts, nrm, tri = o.contour(ob)
        assert (nv, nt) == (len(pts), len(tri))
        body = pts.tobytes() + nrm.tobytes() + tri.astype(np.int32).tobytes()
        assert raw[off:off + len(body)] == body
        off += len(body)
    assert off == len(raw)


def test_run_sim_contour_files(tmp_path):
    cfg, objs = configs.cavity_sphere(24)
    cfg.dump_code = 2
    path = str(tmp_path / "c.cfg")
    configs.write_cfg(cfg, objs, path)
    r = subprocess.run([DRIVER, path, "--steps", "20", "--quiet"], capture_output=True, text=True, timeout=300)
    print(r.stdout[-2000:], r.stderr[-2000:])
    assert r.returncode == 0
    for f in range(3):
        raw = (tmp_path / "c.out" / ("ctr.%05d" % f)).read_bytes()
        assert np.frombuffer(raw[:4], dtype=np.int32)[0] == 1
        nv, nt = np.frombuffer(raw[4:12], dtype=np.int32)
        assert nv > 100 and len(raw) == 12 + 24 * nv + 12 * nt
