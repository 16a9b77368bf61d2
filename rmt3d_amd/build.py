"""Build librmt3d.so (HIP kernels + C++ host + C ABI) for gfx950 with hipcc.

Objects go to rmt3d_amd/_build/, the library to rmt3d_amd/librmt3d.so and
the drop-in driver to rmt3d_amd/run_sim.  Every call recompiles what is out
of date; force=True recompiles everything.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "librmt3d.so")
DRIVER = os.path.join(HERE, "run_sim")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

SOURCES = ["k_fluid.hip", "k_mg.hip", "k_band.hip", "k_io.hip", "k_diag.hip", "k_tracers.hip", "k_contour.hip", "fluid3d.cpp", "sim_params.cpp", "chk_pt.cpp",
           "capi.cpp", "serve.cpp"]
HEADERS = ["dev.h", "launch.h", "fluid3d.h", "detmath.h", "mc_table.h"]
# FP64 parity path: no contraction so the arithmetic order is the reference's;
# packed FP32 instructions stay disabled on this hardware pool.
FLAGS = ["--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++17", "-ffp-contract=off",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-Wno-unused-result", "-Wno-unused-value", "-I" + os.path.join(HERE, "..", "include")]


def _newer(src, dst, deps):
    if not os.path.exists(dst):
        return True
    t = os.path.getmtime(dst)
    return any(os.path.getmtime(p) > t for p in [src] + deps)


def _compile(src, force):
    s = os.path.join(CSRC, src)
    o = os.path.join(BUILD, src + ".o")
    deps = [os.path.join(CSRC, h) for h in HEADERS] + [os.path.join(HERE, "..", "include", "rmt3d.h")]
    if not force and not _newer(s, o, deps):
        return o
    lang = ["-x", "hip"] if src.endswith(".hip") else []
    cmd = [HIPCC] + FLAGS + lang + ["-c", s, "-o", o]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed for %s:\n%s\n%s" % (src, r.stdout, r.stderr[-6000:]))
    return o


def build(force=False, verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    with ThreadPoolExecutor(max_workers=min(6, os.cpu_count() or 1)) as ex:
        objs = list(ex.map(lambda s: _compile(s, force), SOURCES))
    if force or not os.path.exists(LIB) or any(os.path.getmtime(o) > os.path.getmtime(LIB) for o in objs):
        cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", LIB] + objs
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError("link failed:\n%s" % r.stderr[-4000:])
    drv_src = os.path.join(CSRC, "run_sim.cpp")
    if os.path.exists(drv_src) and (force or _newer(drv_src, DRIVER, [LIB])):
        cmd = [HIPCC, "-O2", "-std=c++17", "-I" + os.path.join(HERE, "..", "include"), drv_src,
               "-o", DRIVER, "-L" + HERE, "-lrmt3d", "-Wl,-rpath,$ORIGIN"]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError("driver link failed:\n%s" % r.stderr[-4000:])
    if verbose:
        print("built", LIB)
    return LIB


if __name__ == "__main__":
    build(force="--force" in sys.argv, verbose=True)
