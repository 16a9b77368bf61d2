"""Serving many independent FSI solves on one GPU (the C ABI's rmt3d_server).

Server wraps the C++ serving loop: FIFO admission against a device-memory
budget, up to max_active concurrent solves (one context, stream and worker
thread each), per-request submit/admit/done timestamps.

mixed_request() draws one request of the serving workload: a unit box with
no-slip walls and a lid moving along x at a random speed in [0.5, 2], grid
drawn from {32, 48, 64, 96, 128}^3, 1-8 neo-Hookean spheres with radius in
[0.05, 0.15] and G in [0.5, 5] placed by random sequential addition without
overlap, 200 steps.  poisson_run() submits such requests at Poisson arrival
times and reports p50/p99 time-to-completion and aggregate cell-steps/s.
"""
import ctypes as C
import time

import numpy as np

from . import _abi
from .lib import Rmt3dError, lib

GRIDS = (32, 48, 64, 96, 128)


def estimate_bytes(cfg, objs, precision=_abi.PREC_FP64):
    n = max(1, len(objs))
    arr = (_abi.Object * n)(*objs) if objs else (_abi.Object * 1)()
    b = C.c_double()
    rc = lib().rmt3d_estimate_bytes(C.byref(cfg), arr, precision, C.byref(b))
    if rc:
        raise Rmt3dError("rmt3d_estimate_bytes failed with %d" % rc)
    return b.value


class Server:
    def __init__(self, hbm_budget_bytes, max_active=4):
        h = C.c_void_p()
        rc = lib().rmt3d_server_create(float(hbm_budget_bytes), int(max_active), C.byref(h))
        if rc:
            raise Rmt3dError("rmt3d_server_create failed with %d" % rc)
        self.h = h
        self._cfg = {}

    def submit(self, cfg, objs, steps, precision=_abi.PREC_FP64):
        n = max(1, len(objs))
        arr = (_abi.Object * n)(*objs) if objs else (_abi.Object * 1)()
        req = C.c_int()
        rc = lib().rmt3d_server_submit(self.h, C.byref(cfg), arr, precision, int(steps), C.byref(req))
        if rc:
            raise Rmt3dError("rmt3d_server_submit failed with %d" % rc)
        self._cfg[req.value] = tuple(cfg.sys_size)
        return req.value

    @staticmethod
    def _info_dict(info):
        return {k: getattr(info, k) for k, _ in _abi.RequestInfo._fields_}

    def wait(self, req, timeout=-1.0):
        info = _abi.RequestInfo()
        rc = lib().rmt3d_server_wait(self.h, req, float(timeout), C.byref(info))
        if rc < 0:
            raise Rmt3dError("rmt3d_server_wait failed with %d" % rc)
        return rc == 0, self._info_dict(info)

    def info(self, req):
        info = _abi.RequestInfo()
        rc = lib().rmt3d_server_info(self.h, req, C.byref(info))
        if rc:
            raise Rmt3dError("rmt3d_server_info failed with %d" % rc)
        return self._info_dict(info)

    def error(self, req):
        return lib().rmt3d_server_error(self.h, req).decode()

    def result(self, req):
        m, n, o = self._cfg[req]
        Ng = (m + 4) * (n + 4) * (o + 4)
        f = {k: np.zeros(Ng) for k in ("u", "v", "w", "p")}
        f["rm_c"] = np.zeros(Ng, dtype=np.uint32)
        st = _abi.State()
        for k in ("u", "v", "w", "p"):
            setattr(st, k, f[k].ctypes.data_as(C.POINTER(C.c_double)))
        st.rm_c = f["rm_c"].ctypes.data_as(C.POINTER(C.c_uint32))
        rc = lib().rmt3d_server_result(self.h, req, C.byref(st))
        if rc:
            raise Rmt3dError("rmt3d_server_result failed with %d" % rc)
        return f

    def release(self, req):
        lib().rmt3d_server_release(self.h, req)
        self._cfg.pop(req, None)

    def stats(self):
        pb, pa, now = C.c_double(), C.c_int(), C.c_double()
        lib().rmt3d_server_stats(self.h, C.byref(pb), C.byref(pa), C.byref(now))
        return {"peak_bytes": pb.value, "peak_active": pa.value, "now": now.value}

    def now(self):
        return self.stats()["now"]

    def close(self):
        if self.h:
            lib().rmt3d_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def place_spheres(rng, count, rmin, rmax, margin=0.05, gap=0.05, tries=10000):
    """Random sequential addition of non-overlapping spheres in the unit box,
    each at least `margin` from the walls and `gap` from the others."""
    out = []
    for _ in range(tries):
        if len(out) == count:
            break
        r = rng.uniform(rmin, rmax)
        c = rng.uniform(r + margin, 1 - r - margin, 3)
        if all(np.linalg.norm(c - c2) > r + r2 + gap for c2, r2 in out):
            out.append((c, r))
    return out


def mixed_request(rng, steps=200, grids=GRIDS, max_spheres=8):
    """One request of the serving workload: (cfg, objs, steps)."""
    from . import configs
    n = int(rng.choice(grids))
    cfg, _ = configs.cavity_sphere(n, steps=steps)
    lid = rng.uniform(0.5, 2.0)
    for comp in range(3):
        cfg.bcv[comp][5] = (lid, 0.0, 0.0)[comp]
    spheres = place_spheres(rng, int(rng.integers(1, max_spheres + 1)), 0.05, 0.15)
    objs = [configs.sphere(tuple(c), r, G=float(rng.uniform(0.5, 5.0))) for c, r in spheres]
    cfg.n_obj = len(objs)
    return cfg, objs, steps


def poisson_run(rate, n_requests, seed=20261202, steps=200, max_active=4, budget_bytes=96e9,
                precision=_abi.PREC_FP64, grids=GRIDS, max_spheres=8):
    """Submit n_requests at Poisson arrival times of the given rate (per
    second); return latency percentiles and aggregate throughput."""
    rng = np.random.default_rng(seed)
    reqs = [mixed_request(rng, steps, grids, max_spheres) for _ in range(n_requests)]
    gaps = rng.exponential(1.0 / rate, n_requests)
    srv = Server(budget_bytes, max_active)
    ids = []
    t0 = time.perf_counter()
    arrival = 0.0
    for (cfg, objs, st), g in zip(reqs, gaps):
        arrival += g
        delay = arrival - (time.perf_counter() - t0)
        if delay > 0:
            time.sleep(delay)
        ids.append(srv.submit(cfg, objs, st, precision))
    infos = []
    for r in ids:
        _, info = srv.wait(r)
        infos.append(info)
    stats = srv.stats()
    srv.close()
    ok = [i for i in infos if i["status"] == _abi.REQ_DONE]
    lat = np.array([i["t_done"] - i["t_submit"] for i in ok]) if ok else np.zeros(1)
    t_first = min(i["t_submit"] for i in infos)
    t_last = max(i["t_done"] for i in infos)
    cell_steps = sum(i["cell_steps"] for i in infos)
    return {
        "arrival_rate_per_s": rate, "requests": n_requests, "completed": len(ok),
        "failed": n_requests - len(ok), "steps_per_request": steps, "max_active": max_active,
        "hbm_budget_gb": budget_bytes / 1e9, "peak_device_gb": round(stats["peak_bytes"] / 1e9, 3),
        "peak_active": stats["peak_active"],
        "p50_s": float(np.percentile(lat, 50)), "p99_s": float(np.percentile(lat, 99)),
        "mean_queue_s": float(np.mean([i["t_admit"] - i["t_submit"] for i in ok])) if ok else 0.0,
        "aggregate_cell_steps_per_s": cell_steps / (t_last - t_first) if t_last > t_first else 0.0,
        "span_s": t_last - t_first,
    }
