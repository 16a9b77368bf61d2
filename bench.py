#!/usr/bin/env python3
"""Headline benchmark: FSI cell-steps/s of the per-timestep fluid-structure
step on one MI355X.

Workload (one "step" = one fluid_3d::step_forward over the whole grid): the
128^3 wall-bounded box with one sphere (radius 0.25, centred) pre-stretched to
F = diag(1.3, 1.3^-1/2, 1.3^-1/2) relaxing in quiescent fluid, mu = 0.01,
G = 1 (rmt3d_amd/configs.py: stretched_sphere).  Inputs are resident in HBM
when the timed region starts; the timed region is exactly --steps steps,
bracketed by device synchronisation.

The JSON line carries:
  value        whole-job cell-steps/s of the measured precision path
  roofline     the dominant kernel's algorithmic HBM bytes per launch over its
               average launch time, timed live with HIP events on the stream the
               kernel runs on
  cpu_baseline the FP64 oracle (C restatement of the reference) on a bounded
               sample of the same workload on this host
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "FSI cell-steps/s at 128^3 (mixed vs FP64, same error); V-cycle HBM GB/s; p99"


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=128)
    ap.add_argument("--precision", choices=["fp64", "mixed"], default="fp64")
    ap.add_argument("--cpu-steps", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true")
    return ap.parse_args()


def fem_colour_bytes(nodes, word):
    """Algorithmic HBM bytes of one colour pass of the nodal 27-point
    Gauss-Seidel smoother: the other seven colours' unknowns are read once
    (7/8 N words), the pass's own rhs read and unknowns written (1/8 N each)."""
    return word * nodes * (7.0 / 8 + 1.0 / 8 + 1.0 / 8)


def run_gpu(args):
    from rmt3d_amd import _abi, configs
    from rmt3d_amd.lib import Simulation
    cfg, objs = configs.stretched_sphere(args.n, steps=args.steps + args.warmup + 10)
    prec = _abi.PREC_MIXED if args.precision == "mixed" else _abi.PREC_FP64
    sim = Simulation(cfg, objs, prec)
    e = sim.initialize()
    sim.prepare_run(e)
    for _ in range(args.warmup):
        rc = sim.step()
        if rc:
            raise RuntimeError("step returned %d" % rc)
    sim.synchronize()
    sim.profile(True)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        rc = sim.step()
        if rc:
            raise RuntimeError("step returned %d" % rc)
    sim.synchronize()
    t1 = time.perf_counter()
    ms_fem, n_fem = sim.profile_read(0)
    ms_mac, n_mac = sim.profile_read(1)
    sim.profile(False)
    elapsed = t1 - t0
    m, n, o = cfg.sys_size
    cells = m * n * o
    word = 8
    nodes = (m + 1) * (n + 1) * (o + 1)
    avg_s = (ms_fem / max(n_fem, 1)) * 1e-3
    bytes_launch = fem_colour_bytes(nodes, word)
    achieved = bytes_launch / avg_s / 1e9 if avg_s > 0 else 0.0
    roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
            "kernel": "fem smoother colour pass (finest level, %d^3 nodes)" % (m + 1),
            "bytes_per_launch": bytes_launch, "avg_launch_us": round(avg_s * 1e6, 3),
            "launches": n_fem}
    res = {"value": cells * args.steps / elapsed, "elapsed": elapsed, "roofline": roof,
           "ms_fem_smooth": ms_fem, "ms_mac_smooth": ms_mac, "n_mac": n_mac,
           "mac_cycles": sim.diag(1), "fem_cycles": sim.diag(2), "dt": sim.dt}
    sim.close()
    return res


def run_cpu(args):
    import multiprocessing
    from oracle.oracle import OracleSim
    from rmt3d_amd import configs
    cfg, objs = configs.stretched_sphere(args.n)
    S = OracleSim(cfg, objs)
    e = S.initialize()
    S.prepare_run(0)
    t0 = time.perf_counter()
    for _ in range(args.cpu_steps):
        S.step()
    t1 = time.perf_counter()
    m, n, o = cfg.sys_size
    S.close()
    return {"value": m * n * o * args.cpu_steps / (t1 - t0), "unit": "cell-steps/s", "cores": 1,
            "kind": "port",
            "sample": "%d step(s) of the %d^3 stretched-sphere workload after initialize(), FP64 oracle "
                      "(plain-C restatement of the reference step), one thread; host has %d logical CPUs"
                      % (args.cpu_steps, args.n, multiprocessing.cpu_count())}


def main():
    args = parse()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py: this build runs the FSI step on one GPU only; --gpus %d is not supported "
              "(multi-GPU sharding is design only)" % max(args.gpus, world), file=sys.stderr)
        sys.exit(2)
    g = run_gpu(args)
    line = {
        "metric": METRIC,
        "value": round(g["value"], 1),
        "unit": "cell-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * g["elapsed"] / args.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f32+f64",
        "data": "synthetic",
        "config": {"workload": "stretched_sphere_%d" % args.n, "grid": [args.n] * 3, "objects": 1,
                   "lambda": 1.3, "G": 1.0, "mu": 0.01, "nlayers": 4, "precision": args.precision,
                   "dt": g["dt"], "mac_vcycles_last_step": g["mac_cycles"],
                   "fem_vcycles_last_step": g["fem_cycles"]},
        "roofline": g["roofline"],
    }
    if not args.no_cpu:
        line["cpu_baseline"] = run_cpu(args)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
