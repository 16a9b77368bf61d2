"""The C-ABI library loads (no GPU needed) and exports every entry point
include/rmt3d.h declares."""
import ctypes
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    src = open(os.path.join(ROOT, "include", "rmt3d.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(rmt3d_[a-z0-9_]+)\s*\(", src)))


def test_header_declares_entry_points():
    syms = declared_symbols()
    assert "rmt3d_step" in syms and "rmt3d_mg_solve" in syms and len(syms) >= 20


def test_library_exports_every_declared_symbol():
    from rmt3d_amd.lib import LIB_PATH
    assert os.path.exists(LIB_PATH), "librmt3d.so not built"
    L = ctypes.CDLL(LIB_PATH)
    missing = [s for s in declared_symbols() if not hasattr(L, s)]
    assert not missing, missing
