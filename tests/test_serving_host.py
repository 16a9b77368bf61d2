"""Host-side parts of the serving loop that need no GPU: the footprint
estimate, request generation, and admission refusing a request larger than
the whole budget."""
import numpy as np
import pytest

from rmt3d_amd import configs, serve
from rmt3d_amd.lib import Rmt3dError


def test_estimate_grows_with_grid_and_objects():
    e = [serve.estimate_bytes(*configs.stretched_sphere(n)) for n in (16, 32, 64)]
    assert 0 < e[0] < e[1] < e[2]
    cfg, objs = configs.cavity_sphere(32)
    one = serve.estimate_bytes(cfg, objs)
    cfg.n_obj = 2
    two = serve.estimate_bytes(cfg, objs + [configs.sphere((0.2, 0.2, 0.2), 0.1)])
    assert two > one


def test_mixed_requests_are_valid():
    rng = np.random.default_rng(0)
    for _ in range(50):
        cfg, objs, steps = serve.mixed_request(rng)
        assert cfg.sys_size[0] in serve.GRIDS and 1 <= len(objs) <= 8 and cfg.n_obj == len(objs)
        assert 0.5 <= cfg.bcv[0][5] <= 2.0
        for a, o in enumerate(objs):
            c, r = np.array(o.basic[:3]), o.basic[3]
            assert 0.05 <= r <= 0.15 and 0.5 <= o.shear_mod <= 5.0
            assert np.all(c - r >= 0.05 - 1e-12) and np.all(c + r <= 0.95 + 1e-12)
            for o2 in objs[a + 1:]:
                assert np.linalg.norm(c - np.array(o2.basic[:3])) > r + o2.basic[3]


def test_server_rejects_request_over_budget():
    cfg, objs = configs.stretched_sphere(64)
    srv = serve.Server(1e6, max_active=1)
    with pytest.raises(Rmt3dError):
        srv.submit(cfg, objs, 1)
    srv.close()
