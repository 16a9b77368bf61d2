"""The C++ command-line driver (rmt3d_amd/run_sim) runs a .cfg end to end
through the C ABI, and its final state is the library's."""
import os
import subprocess

import pytest

from rmt3d_amd import configs

pytestmark = pytest.mark.gpu

DRIVER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "rmt3d_amd", "run_sim")


@pytest.mark.parametrize("mixed", [False, True])
def test_run_sim_driver(tmp_path, mixed):
    cfg, objs = configs.cavity_sphere(24)
    path = str(tmp_path / "cavity24.cfg")
    configs.write_cfg(cfg, objs, path, comment="driver test")
    cmd = [DRIVER, path, "--steps", "6"] + (["--mixed"] if mixed else [])
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=120)
    print(r.stdout[-2000:], r.stderr[-2000:])
    assert r.returncode == 0
    assert "6 steps" in r.stdout and "cell-steps/s" in r.stdout
    assert "stopped" not in r.stdout
