"""The macro.dat / centroid.dat reductions on the HIP path against the
oracle: FP64 results equal bit for bit (same per-cell terms, same blocked
summation order), and the average velocities that compute_v_centroid keeps
for the wall repulsion leave the two runs identical afterwards."""
import numpy as np
import pytest

from oracle.oracle import OracleSim
from rmt3d_amd import _abi, configs
from rmt3d_amd.lib import Simulation

pytestmark = pytest.mark.gpu


def run_pair(cfg, objs, steps):
    g, o = Simulation(cfg, objs), OracleSim(cfg, objs)
    eg, eo = g.initialize(), o.initialize()
    g.prepare_run(eg)
    o.prepare_run(eo)
    for _ in range(steps):
        assert g.step() == 0 and o.step() == 0
    return g, o


@pytest.mark.parametrize("case", ["cavity", "stretched", "gravity"])
def test_macro_data_matches_oracle(case):
    if case == "stretched":
        cfg, objs = configs.stretched_sphere(24)
        steps = 0
    else:
        cfg, objs = configs.cavity_sphere(24)
        steps = 4
        if case == "gravity":
            cfg.gravity = 1.5
            objs[0].srho = 1.2
    print("macro", case, flush=True)
    g, o = run_pair(cfg, objs, steps)
    a, b = g.macro_data(), o.macro_data()
    for k in _abi.MACRO_FIELDS:
        assert a[k] == b[k], (k, a[k], b[k])
    if case == "stretched":
        assert a["elastic"] > 0 and a["max_dev_detF"] < 1e-12
    if case == "gravity":
        assert a["potential"] > 0


def test_centroids_match_and_runs_stay_identical():
    # a sphere close to the bottom wall, inside the repulsion distance
    cfg, objs = configs.cavity_sphere(24)
    for i, v in enumerate([0.0, 1.0, 0.0, 1.0, 0.0, 1.0]):
        cfg.wall_pos[i] = v
    objs[0].basic[2] = 0.22
    _abi.finalize_config(cfg)
    g, o = run_pair(cfg, objs, 3)
    cg, co = g.centroids(), o.centroids()
    assert np.array_equal(cg, co)
    assert abs(cg[0, 2] - 0.22) < 0.02
    for _ in range(3):
        assert g.step() == 0 and o.step() == 0
    fg, fo = g.download(), o.download()
    for k in ("u", "v", "w", "p", "rm_x", "band_x"):
        assert np.array_equal(fg[k], fo[k]), k
    assert np.array_equal(g.centroids(), o.centroids())


def test_macro_mixed_precision_close_to_fp64():
    cfg, objs = configs.cavity_sphere(24)
    g64 = Simulation(cfg, objs)
    gmx = Simulation(cfg, objs, _abi.PREC_MIXED)
    for s in (g64, gmx):
        s.prepare_run(s.initialize())
        for _ in range(4):
            assert s.step() == 0
    a, b = g64.macro_data(), gmx.macro_data()
    for k in ("kinetic", "momentum_x", "elastic"):
        assert abs(a[k] - b[k]) <= 1e-6 * max(abs(a[k]), 1e-12), k
