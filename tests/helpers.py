"""Shared helpers of the parity tests: run the HIP path and the oracle on the
same configuration and compare their states."""
import numpy as np

from oracle.oracle import OracleSim
from rmt3d_amd import _abi
from rmt3d_amd.lib import Simulation


def rel_l2(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    den = np.sqrt(np.sum(b * b))
    num = np.sqrt(np.sum((a - b) ** 2))
    if den == 0:
        return num
    return num / den


def interior_mask(cfg, pad=0):
    m, n, o = cfg.sys_size
    k, j, i = np.meshgrid(np.arange(o + 4), np.arange(n + 4), np.arange(m + 4), indexing="ij")
    lo, hi = 2 - pad, (m + 2 + pad, n + 2 + pad, o + 2 + pad)
    msk = (i >= lo) & (i < hi[0]) & (j >= lo) & (j < hi[1]) & (k >= lo) & (k < hi[2])
    return msk.ravel()


def pair(cfg, objs, precision=_abi.PREC_FP64):
    return Simulation(cfg, objs, precision), OracleSim(cfg, objs)


def compare_states(fg, fo, cfg, tol=1e-10, fields=("u", "v", "w", "p")):
    """Return dict of per-field relative L2 errors over interior cells, and
    assert integer state (primary ids, band cells and codes) bit-exact."""
    out = {}
    msk = interior_mask(cfg)
    for k in fields:
        out[k] = rel_l2(fg[k][msk], fo[k][msk])
    # velocity as a vector field (a component that vanishes by symmetry has no
    # meaningful relative error of its own)
    vg = np.concatenate([fg[k][msk] for k in ("u", "v", "w")])
    vo = np.concatenate([fo[k][msk] for k in ("u", "v", "w")])
    out["vel"] = rel_l2(vg, vo)
    out["rm_c_equal"] = bool(np.array_equal(fg["rm_c"], fo["rm_c"]))
    if "band_cell" in fg and "band_cell" in fo:
        out["band_equal"] = (len(fg["band_cell"]) == len(fo["band_cell"]) and
                             bool(np.array_equal(fg["band_cell"], fo["band_cell"])) and
                             bool(np.array_equal(fg["band_c"], fo["band_c"])))
        if out["band_equal"] and len(fg["band_x"]):
            out["band_x"] = rel_l2(fg["band_x"], fo["band_x"])
    return out
