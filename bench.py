#!/usr/bin/env python3
"""Headline benchmark: FSI cell-steps/s of the per-timestep fluid-structure
step on one MI355X.

Workload (one "step" = one fluid_3d::step_forward over the whole grid): the
128^3 wall-bounded box with one sphere (radius 0.25, centred) pre-stretched to
F = diag(1.3, 1.3^-1/2, 1.3^-1/2) relaxing in quiescent fluid, mu = 0.01,
G = 1 (rmt3d_amd/configs.py: stretched_sphere).  Inputs are resident in HBM
when the timed region starts; the timed region is exactly --steps steps,
bracketed by device synchronisation.

By default the FP64 path and the mixed path (FP32 multigrid preconditioner
inside FP64 defect correction) are timed one after the other from the same
initial state; value is the mixed path, the FP64 rate, the speedup and the
mixed path's relative L2 distance from the FP64 fields after the same steps
sit in config.

The JSON line carries:
  value        whole-job cell-steps/s of the measured precision path
  roofline     the dominant kernel's algorithmic HBM bytes per launch over its
               average launch time, timed with HIP events on the stream the
               kernel runs on, over --kernel-steps steps run right after the
               timed region with per-launch events (the timed region replays
               the multigrid V-cycles as captured graphs, where single
               launches cannot be bracketed; each launch stamps its own start
               and end events, so the figure is the kernel's execution time);
               traffic is the HBM bytes per launch that rocprofv3 --pmc
               measured for the same kernel (profiles/hbm_traffic_*.json),
               null if that summary covers another kernel
  cpu_baseline the FP64 oracle (C restatement of the reference) on a bounded
               sample of the same workload on this host
  config.serving  the serving workload: Poisson arrivals (--serve-rates per
               second, swept; --serve-requests requests per rate) of
               independent mixed-precision solves, grid from
               {32,48,64,96,128}^3, 1-8 spheres, lid speed 0.5-2, 200 steps
               each, up to --serve-max-active at a time; p50/p99
               time-to-completion and aggregate cell-steps/s per rate
               (rmt3d_amd/serve.py)
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
    ap.add_argument("--precision", choices=["both", "fp64", "mixed"], default="both",
                    help="both: time the FP64 path, then the mixed path; value is the mixed path")
    ap.add_argument("--kernel-steps", type=int, default=3,
                    help="steps of the per-kernel event window after the timed region")
    ap.add_argument("--cpu-steps", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run a plain 1 GiB device-to-device copy (the reference kernel for "
                         "scaling FETCH_SIZE / WRITE_SIZE counters to bytes)")
    ap.add_argument("--io", action="store_true",
                    help="instead of the step: time the checkpoint write (device gather + files) and the "
                         "resume (files + device scatter) of the FP64 state after --warmup steps")
    ap.add_argument("--io-reps", type=int, default=5)
    ap.add_argument("--shear", action="store_true",
                    help="time the 256^3 shear workload (64 spheres, periodic x and y) instead")
    ap.add_argument("--serve-rates", type=float, nargs="+", default=[1.0, 4.0, 16.0],
                    help="Poisson arrival rates (per second) to sweep")
    ap.add_argument("--serve-requests", type=int, default=16)
    ap.add_argument("--serve-max-active", type=int, default=6)
    return ap.parse_args()


# Nodal 27-point smoother, four colours by (j,k) parity: a colour pass reads
# the unknowns of the three other colours (3/4 N words), its own rhs (1/4 N)
# and writes its own unknowns (1/4 N).
FEM_COLOUR_WORDS = 3.0 / 4 + 1.0 / 4 + 1.0 / 4
# np consecutive colour passes in one launch (k_sweep): every unknown of the
# level read once, the np colours' rhs read and their unknowns written.


def sweep_words(np_):
    return 1.0 + np_ / 4.0 + np_ / 4.0


# One V-cycle, per finest unknown: the finest level's smoothing sweeps
# (red-black MAC: 2 colour passes of 1.5 words; FEM: 4 colour passes of
# FEM_COLOUR_WORDS), residual 3, restriction 1.25, prolongation 2.125, and
# the fast-diagonalisation solve of the first coarse level, which ends the
# hierarchy of the 128^3 workload (six contractions reading and writing 1/8
# as many unknowns: 1.5 words).  At the reference's 2 + 2 sweeps: MAC 19.9,
# FEM 27.9 words.
def vcycle_words(fem, down, up):
    """Algorithmic words per finest unknown of one V-cycle: the sweeps, the
    finest residual (not computed when no sweep precedes the coarse
    correction: the right-hand side is restricted as it is), restriction,
    prolongation and the coarse solve's share."""
    per_sweep = 4 * FEM_COLOUR_WORDS if fem else 2 * 1.5
    return (down + up) * per_sweep + (3.0 if down > 0 else 0.0) + 1.25 + 2.125 + 1.5
# One FP64 defect-correction step of the mixed path adds the FP64 residual
# (read x, b, write r) per finest unknown.
DEFECT_WORDS64 = 3.0


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "hbm_traffic_fem_smooth_129.json")


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 --pmc
    summary (FETCH_SIZE and WRITE_SIZE, each scaled by what the counter reads
    for a plain 1 GiB copy), or None when the summary covers another kernel."""
    try:
        with open(TRAFFIC_FILE) as f:
            t = json.load(f)
    except (OSError, ValueError):
        return None
    return t.get("traffic_bytes_per_launch") if t.get("kernel") == kernel else None


def run_gpu(args, precision):
    from rmt3d_amd import _abi, configs
    from rmt3d_amd.lib import Simulation
    cfg, objs = configs.stretched_sphere(args.n, steps=args.steps + args.warmup + 10)
    prec = _abi.PREC_MIXED if precision == "mixed" else _abi.PREC_FP64
    sim = Simulation(cfg, objs, prec)
    e = sim.initialize()
    sim.prepare_run(e)
    for _ in range(args.warmup):
        rc = sim.step()
        if rc:
            raise RuntimeError("step returned %d" % rc)
    sim.synchronize()
    # timed region: stage buckets and whole V-cycles are event-timed; the
    # V-cycles replay their captured graphs
    sim.profile(1)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        rc = sim.step()
        if rc:
            raise RuntimeError("step returned %d" % rc)
    sim.synchronize()
    t1 = time.perf_counter()
    prof = {k: sim.profile_read(i) for i, k in enumerate(
        ("fem_smooth", "mac_smooth", "stress", "advect", "fem_vcycle", "mac_vcycle"))}
    stages = {k: round(v / args.steps, 3) for k, v in sim.stage_times().items()}
    # per-kernel window right after: the same steps with every launch of the
    # reported kernel bracketed by events (V-cycles launched one kernel at a
    # time, which leaves each kernel's own duration unchanged)
    sim.profile(2)
    for _ in range(args.kernel_steps):
        rc = sim.step()
        if rc:
            raise RuntimeError("step returned %d" % rc)
    sim.synchronize()
    for i, k in enumerate(("fem_smooth", "mac_smooth", "stress", "advect")):
        prof[k] = sim.profile_read(i)
    sim.profile(0)
    npass = int(sim.diag(17))
    elapsed = t1 - t0
    m, n, o = cfg.sys_size
    cells = m * n * o
    nodes = (m + 1) * (n + 1) * (o + 1)
    word = 4 if precision == "mixed" else 8
    ms, launches = prof["fem_smooth"]
    avg_s = ms / max(launches, 1) * 1e-3
    tname = "float" if word == 4 else "double"
    if npass <= 1:
        bytes_launch = word * nodes * FEM_COLOUR_WORDS
        kname = "k_fem_smooth<%s> colour pass, finest level (%d^3 nodes)" % (tname, m + 1)
    else:
        bytes_launch = word * nodes * sweep_words(npass)
        kname = "k_sweep<%s, FEM> %d colour passes, finest level (%d^3 nodes)" % (tname, npass, m + 1)
    achieved = bytes_launch / avg_s / 1e9 if avg_s > 0 else 0.0
    roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": measured_traffic(kname),
            "kernel": kname,
            "bytes_per_launch": bytes_launch, "avg_launch_us": round(avg_s * 1e6, 3),
            "launches": launches}
    vc = {}
    for key, N in (("fem", nodes), ("mac", cells)):
        ms_v, nv = prof[key + "_vcycle"]
        t_v = ms_v / max(nv, 1) * 1e-3
        sched = int(sim.diag(34 if key == "fem" else 33))
        words = vcycle_words(key == "fem", sched // 10, sched % 10)
        b = words * N * word + (DEFECT_WORDS64 * N * 8 if precision == "mixed" else 0)
        vc[key] = {"us": round(t_v * 1e6, 2), "gbs": round(b / t_v / 1e9, 1) if t_v > 0 else 0.0,
                   "frac": round(b / t_v / 1e9 / HBM_PEAK_GBS, 4) if t_v > 0 else 0.0, "cycles": nv}
    fields = sim.download(band=False)
    res = {"value": cells * args.steps / elapsed, "elapsed": elapsed, "roofline": roof, "vcycle": vc,
           "mac_cycles": sim.diag(1), "fem_cycles": sim.diag(2), "dt": sim.dt,
           "autotuned": {"mac_passes_per_launch": sim.diag(16), "fem_passes_per_launch": sim.diag(17),
                         "mac_colour_pass_workgroup": sim.diag(24), "fem_colour_pass_workgroup": sim.diag(25),
                         "mac_fdm_outputs_per_item_xyz": int(sim.diag(26)),
                         "fem_fdm_outputs_per_item_xyz": int(sim.diag(27)),
                         "mac_fdm_form": int(sim.diag(31)), "fem_fdm_form": int(sim.diag(32)),
                         "mac_vcycle_sweeps_down_up": int(sim.diag(33)),
                         "fem_vcycle_sweeps_down_up": int(sim.diag(34))},
           "stress_ms_per_step": prof["stress"][0] / max(args.kernel_steps, 1), "stage_ms_per_step": stages,
           "fields": {k: fields[k] for k in ("u", "v", "w", "p")}}
    sim.close()
    return res


def run_cpu(args):
    import multiprocessing
    from oracle.oracle import OracleSim, lib
    from rmt3d_amd import configs
    cfg, objs = configs.stretched_sphere(args.n)
    S = OracleSim(cfg, objs)
    e = S.initialize()
    S.prepare_run(e)
    t0 = time.perf_counter()
    for _ in range(args.cpu_steps):
        S.step()
    t1 = time.perf_counter()
    m, n, o = cfg.sys_size
    S.close()
    threads = lib().oracle_threads()
    return {"value": m * n * o * args.cpu_steps / (t1 - t0), "unit": "cell-steps/s", "cores": threads,
            "kind": "port",
            "sample": "%d step(s) of the %d^3 stretched-sphere workload after initialize() and the init_iter "
                      "warm-up, FP64 oracle (plain-C restatement of the reference step) with OpenMP over %d "
                      "threads on its per-cell and per-colour loops (extrapolation, upwinding and FEM rhs "
                      "serial); host reports %d logical CPUs"
                      % (args.cpu_steps, args.n, threads, multiprocessing.cpu_count())}


def run_io(args):
    """Checkpoint files of the --n^3 stretched-sphere state (fluid_3d::
    write_chk_pt / initialize_from_chk_point).  One JSON line: value is the
    checkpoint bytes written per second of rmt3d_write_checkpoint (device
    gather, copy to the host, eight files into a scratch directory);
    roofline is the gather kernel k_chk_pack against HBM; cpu_baseline is the
    oracle's gather of the same state (the reference's slice loop restated)."""
    import shutil
    import tempfile
    import numpy as np
    from oracle.oracle import OracleSim
    from rmt3d_amd import _abi, configs
    from rmt3d_amd import lib as rlib
    from rmt3d_amd.lib import Simulation
    cfg, objs = configs.stretched_sphere(args.n, steps=args.warmup + 10)
    work = tempfile.mkdtemp(prefix="rmt3d_io_")
    try:
        out = os.path.join(work, "run.out")
        chk = os.path.join(out, "chk.00001")
        os.makedirs(chk)
        cfg.dirname = out.encode()
        sim = Simulation(cfg, objs, _abi.PREC_FP64)
        sim.prepare_run(sim.initialize())
        for _ in range(args.warmup):
            if sim.step():
                raise RuntimeError("step failed")
        sim.synchronize()
        sim.write_checkpoint(args.warmup, 1, chk)       # first launch loads the kernels
        sim.profile(1)
        t0 = time.perf_counter()
        for _ in range(args.io_reps):
            sim.write_checkpoint(args.warmup, 1, chk)
        t1 = time.perf_counter()
        pack_ms, pack_n = sim.profile_read(6)
        # macro.dat / centroid.dat reductions of the same state
        sim.macro_data()
        sim.centroids()
        tm0 = time.perf_counter()
        for _ in range(args.io_reps):
            sim.macro_data()
            sim.centroids()
        tm1 = time.perf_counter()
        mac_ms, mac_n = sim.profile_read(8)
        cen_ms, cen_n = sim.profile_read(9)
        st = sim.download()
        sim.profile(0)
        m, n, o = cfg.sys_size
        L = (m + 1) * (n + 1)
        fm, fn = m + 1, n + 1
        N = m * n * o
        nband = len(st["band_cell"])
        n_solid = int(np.count_nonzero(st["rm_c"]))
        file_bytes = 8 * 8 * o * L
        # gather: 8 words written per slot; per cell vel 3 words, c and cnt
        # 4 B each; per node p 1 word; x 3 words of each primary solid cell;
        # start 4 B, c 4 B and x 3 words of each band entry
        pack_bytes = file_bytes + N * (24 + 4 + 4) + o * fm * fn * 8 + n_solid * 24 + nband * 36
        pack_s = pack_ms / max(pack_n, 1) * 1e-3
        achieved = pack_bytes / pack_s / 1e9
        # macro partials: per cell vel 3 words, six face velocities, c 4 B;
        # per solid cell en 3 + det F 3 words of its own map and of its three
        # upper neighbours' maps, and the band count of each of those cells
        macro_bytes = N * (24 + 48 + 4) + n_solid * (48 + 3 * (48 + 4 + 4))
        # centroid partials: c of every cell; x 3 and vel 3 words per solid cell
        cen_bytes = N * 4 + n_solid * 48
        # resume from those files
        c2, o2 = rlib.parse_cfg(os.path.join(chk, "recover.cfg"))
        h = Simulation(c2, o2, _abi.PREC_FP64)
        h.profile(1)
        t2 = time.perf_counter()
        e = h.initialize()
        h.synchronize()
        t3 = time.perf_counter()
        unpack_ms, unpack_n = h.profile_read(7)
        h.close()
        if e != 0:
            raise RuntimeError("resume returned %d" % e)
        # scatter: per cell 7 words read, vel 3, x 3, xpred 3 words and c
        # written; per node 1 word read and written
        unpack_bytes = N * (56 + 72 + 4) + o * fm * fn * 16
        # the oracle gathers the same state: files must match bit for bit
        orc = OracleSim(cfg, objs)
        orc.upload(**{k: st[k] for k in ("u", "v", "w", "p", "rm_c", "rm_x", "rm_xpred")})
        orc.set_band(st["band_cell"], st["band_c"], st["band_x"].reshape(3, -1), st["band_xpred"].reshape(3, -1))
        t4 = time.perf_counter()
        img = orc.chk_images()
        t5 = time.perf_counter()
        orc.macro_data()
        orc.centroids()
        t6 = time.perf_counter()
        orc.close()
        names = ("vx", "vy", "vz", "p", "X", "Y", "Z", "OID")
        match = all(np.array_equal(np.fromfile(os.path.join(chk, nm)), img[f].ravel()) for f, nm in enumerate(names))
        sim.close()
    finally:
        shutil.rmtree(work, ignore_errors=True)
    wall = (t1 - t0) / args.io_reps
    line = {
        "metric": "checkpoint write bytes/s at %d^3 (write_chk_pt files)" % args.n,
        "value": round(file_bytes / wall, 1),
        "unit": "B/s",
        "n_gpus": 1,
        "steps": args.io_reps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * wall, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "stretched_sphere_%d checkpoint" % args.n, "file_bytes": file_bytes,
                   "band_entries": nband, "solid_cells": n_solid, "files_match_oracle": bool(match),
                   "resume_ms": round(1e3 * (t3 - t2), 3),
                   "scatter": {"kernel": "k_chk_unpack", "avg_launch_us": round(unpack_ms / max(unpack_n, 1) * 1e3, 3),
                               "bytes_per_launch": unpack_bytes,
                               "achieved_gbs": round(unpack_bytes / (unpack_ms / max(unpack_n, 1) * 1e-3) / 1e9, 1)
                               if unpack_ms > 0 else 0.0},
                   "macro": {"kernel": "k_macro_partials", "avg_launch_us": round(mac_ms / max(mac_n, 1) * 1e3, 3),
                             "bytes_per_launch": macro_bytes,
                             "achieved_gbs": round(macro_bytes / (mac_ms / max(mac_n, 1) * 1e-3) / 1e9, 1)
                             if mac_ms > 0 else 0.0,
                             "ms_per_record": round(1e3 * (tm1 - tm0) / args.io_reps, 3),
                             "oracle_ms_per_record": round(1e3 * (t6 - t5), 3)},
                   "centroids": {"kernel": "k_centroid_partials",
                                 "avg_launch_us": round(cen_ms / max(cen_n, 1) * 1e3, 3),
                                 "bytes_per_launch": cen_bytes,
                                 "achieved_gbs": round(cen_bytes / (cen_ms / max(cen_n, 1) * 1e-3) / 1e9, 1)
                                 if cen_ms > 0 else 0.0}},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None, "kernel": "k_chk_pack",
                     "bytes_per_launch": pack_bytes, "avg_launch_us": round(pack_s * 1e6, 3), "launches": pack_n},
        "cpu_baseline": {"value": round(file_bytes / (t5 - t4), 1), "unit": "B/s", "cores": 1, "kind": "port",
                         "sample": "one gather of the same %d^3 state by the oracle's restatement of "
                                   "write_chk_pt's slice loop (files not written)" % args.n},
    }
    print(json.dumps(line), flush=True)


def _rel(a, b):
    import numpy as np
    d = np.linalg.norm(a - b)
    n = np.linalg.norm(b)
    return float(d / n) if n > 0 else float(d)


def calibrate_copy():
    """A plain device copy of 1 GiB, five times: read 1 GiB and write 1 GiB
    per launch.  Under rocprofv3 --pmc FETCH_SIZE (or WRITE_SIZE) its counter
    reading per dispatch gives the factor from counter units to bytes."""
    import torch
    n = 1 << 30
    a = torch.ones(n // 4, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(5):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_per_launch": n, "launches": 5}), flush=True)


def run_shear(args):
    """The largest workload, 256^3 with 64 neo-Hookean spheres between
    moving walls (periodic x and y), on the mixed and the FP64 path: after
    initialize() and --warmup steps, the wall time of --steps steps (each
    step returns after its status read-back).  One JSON line."""
    from rmt3d_amd import _abi, configs
    from rmt3d_amd import lib as rlib
    res = {}
    for prec, name in ((_abi.PREC_MIXED, "mixed"), (_abi.PREC_FP64, "fp64")):
        cfg, objs = configs.shear_spheres(256)
        S = rlib.Simulation(cfg, objs, precision=prec)
        S.prepare_run(S.initialize())
        for _ in range(args.warmup):
            rc = S.step()
            if rc:
                raise RuntimeError("shear %s warm-up step returned %d" % (name, rc))
        S.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            rc = S.step()
            if rc:
                raise RuntimeError("shear %s step returned %d" % (name, rc))
        S.synchronize()
        dt = (time.perf_counter() - t0) / args.steps
        res[name] = {"ms_per_step": round(dt * 1e3, 3), "cell_steps_per_s": round(256 ** 3 / dt, 1),
                     "mac_vcycles_last_step": S.diag(1), "fem_vcycles_last_step": S.diag(2),
                     "levels_mac_fem": [S.diag(10), S.diag(11)], "fdm_level": [S.diag(18), S.diag(19)],
                     "extrapolation_object_batches": S.diag(30)}
        print("bench: shear %s %.1f ms/step" % (name, dt * 1e3), file=sys.stderr, flush=True)
        del S
    print(json.dumps({"metric": "FSI cell-steps/s, 256^3 shear workload (64 spheres)",
                      "value": res["mixed"]["cell_steps_per_s"], "unit": "cell-steps/s", "n_gpus": 1,
                      "steps": args.steps, "warmup": args.warmup, "higher_is_better": True,
                      "data": "synthetic", "config": {"workload": "shear_spheres_256", "grid": [256, 256, 256],
                                                      "objects": 64, "mixed": res["mixed"], "fp64": res["fp64"]}}))


def main():
    args = parse()
    if args.calibrate_copy:
        calibrate_copy()
        return
    if args.io:
        run_io(args)
        return
    if args.shear:
        run_shear(args)
        return
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py: this build runs the FSI step on one GPU only; --gpus %d is not supported "
              "(multi-GPU sharding is design only)" % max(args.gpus, world), file=sys.stderr)
        sys.exit(2)
    runs = {}
    for prec in (("fp64", "mixed") if args.precision == "both" else (args.precision,)):
        runs[prec] = run_gpu(args, prec)
        print("bench: %s path %.1f ms/step" % (prec, 1e3 * runs[prec]["elapsed"] / args.steps),
              file=sys.stderr, flush=True)
    head = "mixed" if "mixed" in runs else "fp64"
    g = runs[head]
    config = {"workload": "stretched_sphere_%d" % args.n, "grid": [args.n] * 3, "objects": 1,
              "lambda": 1.3, "G": 1.0, "mu": 0.01, "nlayers": 4, "precision": head,
              "dt": g["dt"], "mac_vcycles_last_step": g["mac_cycles"],
              "fem_vcycles_last_step": g["fem_cycles"], "vcycle": g["vcycle"], "autotuned": g["autotuned"],
              "stress_ms_per_step": round(g["stress_ms_per_step"], 3),
              "stage_ms_per_step": g["stage_ms_per_step"]}
    if len(runs) == 2:
        f64, fmx = runs["fp64"]["fields"], runs["mixed"]["fields"]
        import numpy as np
        vel64 = np.concatenate([f64[k] for k in ("u", "v", "w")])
        velmx = np.concatenate([fmx[k] for k in ("u", "v", "w")])
        config["fp64_cell_steps_per_s"] = round(runs["fp64"]["value"], 1)
        config["fp64_ms_per_step"] = round(1e3 * runs["fp64"]["elapsed"] / args.steps, 4)
        config["mixed_speedup_vs_fp64"] = round(runs["mixed"]["value"] / runs["fp64"]["value"], 3)
        config["mixed_rel_l2_vs_fp64"] = {"vel": _rel(velmx, vel64), "p": _rel(fmx["p"], f64["p"]),
                                          "steps": args.warmup + args.steps + args.kernel_steps}
        config["fp64_vcycle"] = runs["fp64"]["vcycle"]
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
        "dtype": "f32+f64" if head == "mixed" else "f64",
        "data": "synthetic",
        "config": config,
        "roofline": g["roofline"],
    }
    if not args.no_serve:
        from rmt3d_amd import _abi, serve
        # Poisson arrivals of mixed-precision requests at a swept rate, from
        # a light load to past saturation (requests queue for a slot)
        sweep = []
        for rate in args.serve_rates:
            r = serve.poisson_run(rate=rate, n_requests=args.serve_requests, precision=_abi.PREC_MIXED,
                                  max_active=args.serve_max_active)
            sweep.append(r)
            print("bench: serving rate %g/s p50 %.2f s p99 %.2f s, %.3g cell-steps/s, peak %d active" %
                  (rate, r["p50_s"], r["p99_s"], r["aggregate_cell_steps_per_s"], r["peak_active"]),
                  file=sys.stderr, flush=True)
        config["serving"] = {"precision": "mixed", "sweep": sweep}
    if not args.no_cpu:
        line["cpu_baseline"] = run_cpu(args)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
