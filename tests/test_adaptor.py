"""The p_class adaptor of the integration guide compiles and links.

rmt3d_amd/integration/fluid_3d_gpu.hh is what a maintainer adds to run the
reference's stepper and run_sim on the library.  tests/adaptor/
stepper_harness.cc declares stand-ins of the reference types it touches and
makes every p_class call of stepper::run / step / write_files / dump_all /
collect_macro_data and of run_sim's main, in their order; it must compile
without warnings and link against librmt3d.so (it is not run: that needs a
GPU).
"""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_DIR = os.path.join(ROOT, "rmt3d_amd")


def test_adaptor_compiles_and_links(tmp_path):
    src = os.path.join(ROOT, "tests", "adaptor", "stepper_harness.cc")
    cmd = ["g++", "-std=c++17", "-Wall", "-Wextra", "-Werror", "-I" + os.path.join(ROOT, "include"),
           "-I" + os.path.join(LIB_DIR, "integration"), src]
    if os.path.exists(os.path.join(LIB_DIR, "librmt3d.so")):
        cmd += ["-o", str(tmp_path / "harness"), "-L" + LIB_DIR, "-lrmt3d"]
    else:
        cmd += ["-fsyntax-only"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_integration_guide_shows_the_adaptor():
    """INTEGRATION.md carries the adaptor header verbatim."""
    with open(os.path.join(LIB_DIR, "integration", "fluid_3d_gpu.hh")) as f:
        hdr = f.read()
    with open(os.path.join(ROOT, "INTEGRATION.md")) as f:
        doc = f.read()
    body = hdr[hdr.index("class fluid_3d_gpu {"):hdr.index("#endif")].strip()
    assert body in doc
