"""The serving loop: concurrent requests on one GPU give the same results as
each request run alone (FP64 is bit-reproducible), admission keeps the device
bytes held within the budget, and the footprint estimate matches what a
context allocates."""
import numpy as np
import pytest

from rmt3d_amd import _abi, serve
from rmt3d_amd.lib import Simulation

pytestmark = pytest.mark.gpu


def _alone(cfg, objs, steps, precision=_abi.PREC_FP64):
    s = Simulation(cfg, objs, precision)
    e = s.initialize()
    s.prepare_run(e)
    for _ in range(steps):
        assert s.step() == 0
    f = s.download(band=False)
    s.close()
    return f


@pytest.mark.parametrize("precision", [_abi.PREC_FP64, _abi.PREC_MIXED])
def test_results_same_alone_and_in_the_mix(precision):
    """Bitwise on both paths: a request's stage order and arithmetic do not
    depend on the other requests.  The autotuner's choices (launch shapes,
    and on the mixed path the V-cycle schedule and fast-diagonalisation
    form) are made once per grid shape in the process and shared, so a
    request alone and in the mix runs the same schedule."""
    rng = np.random.default_rng(7)
    reqs = [serve.mixed_request(rng, steps=6, grids=(24, 32), max_spheres=4) for _ in range(5)]
    srv = serve.Server(16e9, max_active=3)
    ids = [srv.submit(cfg, objs, st, precision) for cfg, objs, st in reqs]
    for r in ids:
        done, info = srv.wait(r, 300.0)
        assert done and info["status"] == _abi.REQ_DONE, (info, srv.error(r))
    stats = srv.stats()
    print("serving stats", stats)
    assert stats["peak_active"] >= 2
    for (cfg, objs, st), r in zip(reqs, ids):
        got = srv.result(r)
        ref = _alone(cfg, objs, st, precision)
        for k in ("u", "v", "w", "p"):
            assert np.array_equal(got[k], ref[k]), (r, k, np.max(np.abs(got[k] - ref[k])))
        assert np.array_equal(got["rm_c"], ref["rm_c"])
        srv.release(r)
    srv.close()


def test_admission_respects_budget_and_estimate():
    rng = np.random.default_rng(11)
    reqs = [serve.mixed_request(rng, steps=2, grids=(32,), max_spheres=3) for _ in range(4)]
    ests = [serve.estimate_bytes(cfg, objs) for cfg, objs, _ in reqs]
    budget = 1.5 * max(ests)            # room for one request at a time
    srv = serve.Server(budget, max_active=4)
    ids = [srv.submit(cfg, objs, st) for cfg, objs, st in reqs]
    infos = [srv.wait(r, 300.0)[1] for r in ids]
    stats = srv.stats()
    print("budget", budget, "stats", stats, "est", ests, "held", [i["device_bytes"] for i in infos])
    for est, info in zip(ests, infos):
        assert info["status"] == _abi.REQ_DONE
        assert abs(info["device_bytes"] - est) <= 0.1 * est
    assert stats["peak_bytes"] <= budget * 1.1
    assert stats["peak_active"] == 1
    # requests ran one after another, in arrival order
    order = sorted(range(len(ids)), key=lambda a: infos[a]["t_admit"])
    assert order == list(range(len(ids)))
    srv.close()


def test_poisson_run_reports_latency():
    r = serve.poisson_run(rate=4.0, n_requests=6, seed=3, steps=4, max_active=3, budget_bytes=16e9,
                          grids=(24, 32), max_spheres=3)
    print("poisson", r)
    assert r["completed"] == 6 and r["failed"] == 0
    assert 0 < r["p50_s"] <= r["p99_s"]
    assert r["aggregate_cell_steps_per_s"] > 0
