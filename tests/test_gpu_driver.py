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


def test_run_sim_outputs_and_resume(tmp_path):
    # dump_code 4|8: slices of out_flag's fields at output_ind and
    # checkpoints every frame (stepper.hh:580-610); the written checkpoint is
    # byte-identical to the oracle's after the same steps, and a run resumed
    # from its recover.cfg continues like the oracle resumed from the same
    # files
    import numpy as np
    from oracle.oracle import OracleSim
    from rmt3d_amd import lib as rlib
    cfg, objs = configs.cavity_sphere(24)
    cfg.dump_code = 12
    path = str(tmp_path / "cav.cfg")
    configs.write_cfg(cfg, objs, path, comment="driver output test")
    r = subprocess.run([DRIVER, path, "--steps", "20", "--quiet"], capture_output=True, text=True, timeout=300)
    print(r.stdout[-2000:], r.stderr[-2000:])
    assert r.returncode == 0 and "20 steps" in r.stdout
    out = tmp_path / "cav.out"
    names = ("vx", "vy", "vz", "p", "X", "Y", "Z", "OID")
    for f in range(3):
        assert (out / ("chk.%05d" % f) / "recover.cfg").exists()
        for name in ("vx", "vy", "vz", "p"):          # out_flag 135
            assert (out / ("%s.%05d" % (name, f))).exists()
    c1, o1 = rlib.parse_cfg(path)
    q = OracleSim(c1, o1)
    rc, dt, tot = q.prepare_run(q.initialize())
    for s in range(20):
        assert q.step() == 0
    ref = tmp_path / "oracle_chk"
    ref.mkdir()
    q.write_chk_pt(str(ref))
    for name in names:
        assert (out / "chk.00002" / name).read_bytes() == (ref / name).read_bytes(), name
    sl = q.slice(2, c1.output_ind, 0)
    want = "".join("".join("%17.14g " % v for v in row) + "\n" for row in sl)
    assert (out / "vx.00002").read_text() == want
    # resume from frame 1 for one frame: recover.cfg keeps dt to %g, so the
    # resumed stepper may take a different step count per frame
    rec = str(out / "chk.00001" / "recover.cfg")
    c2, o2 = rlib.parse_cfg(rec)
    assert c2.chk_num == 1 and c2.chk_step == 10
    p = OracleSim(c2, o2)
    assert p.initialize() == 0
    rc, dt2, tot2 = p.prepare_run(0)
    spf = int(round((c2.T / (c2.frames - 1)) / dt2))
    r = subprocess.run([DRIVER, rec, "--steps", str(spf), "--quiet"], capture_output=True, text=True, timeout=300)
    print(r.stdout[-2000:], r.stderr[-2000:])
    assert r.returncode == 0 and "%d steps" % spf in r.stdout
    for s in range(spf):
        assert p.step() == 0
    ref2 = tmp_path / "oracle_chk2"
    ref2.mkdir()
    p.write_chk_pt(str(ref2))
    res = out / "chk.00001" / "recover.out" / "chk.00002"
    for name in names:
        assert (res / name).read_bytes() == (ref2 / name).read_bytes(), name
    text = (res / "recover.cfg").read_text()
    assert "chk_step                  %d#" % (10 + spf) in text


def test_run_sim_macro_files(tmp_path):
    # dump_code 32: one macro.dat record per frame plus the initial one (the
    # reference writes them without line breaks) and one centroid.dat line
    # per record
    cfg, objs = configs.cavity_sphere(24)
    cfg.dump_code = 32
    path = str(tmp_path / "m.cfg")
    configs.write_cfg(cfg, objs, path)
    r = subprocess.run([DRIVER, path, "--steps", "20", "--quiet"], capture_output=True, text=True, timeout=300)
    print(r.stdout[-2000:], r.stderr[-2000:])
    assert r.returncode == 0
    lines = (tmp_path / "m.out" / "macro.dat").read_text().split("\n")
    assert lines[0].split()[0] == "#Time" and len(lines) == 2
    assert len(lines[1].split()) == 3 * 18
    cent = (tmp_path / "m.out" / "centroid.dat").read_text().strip().split("\n")
    assert len(cent) == 4 and all(len(l.split()) == 9 for l in cent[1:])
