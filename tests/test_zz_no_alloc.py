"""rmt3d_slice_noalloc, the allocation-free slice, against rmt3d_slice.

Both entry points run the same gather kernel; the allocation-free one stages
the slice in the library's fixed per-GPU scratch instead of a buffer
allocated per call.  Two calls in a row, the second with fewer values (and a
second context), must each equal rmt3d_slice exactly, so nothing of the first
call's larger slice leaks into the second.
"""
import numpy as np
import pytest

from rmt3d_amd import _abi, configs
from rmt3d_amd.lib import Rmt3dError, Simulation

pytestmark = pytest.mark.gpu


def _stepped(shape, steps=2):
    cfg, objs = configs.stretched_sphere(shape[0], steps=10)
    # cubic cells of side 1/32 (or 1/16): the box is shape / shape[1] long
    for a in range(3):
        cfg.sys_size[a] = shape[a]
        cfg.xyz_bounds[2 * a] = 0.0
        cfg.xyz_bounds[2 * a + 1] = shape[a] / shape[1]
    _abi.finalize_config(cfg)
    g = Simulation(cfg, objs)
    e = g.initialize()
    g.prepare_run(e)
    for _ in range(steps):
        assert g.step() == 0
    return g


def test_slice_noalloc_matches_slice_twice():
    g = _stepped((40, 32, 36))
    print("no-alloc slice: 40x32x36 context, z then x slices", flush=True)
    for t in range(22):
        big_ref = g.slice(2, 11, t)
        big = g.slice(2, 11, t, noalloc=True)
        assert big.shape == big_ref.shape
        assert np.array_equal(big, big_ref), ("z slice", t)
        small_ref = g.slice(0, 17, t)
        small = g.slice(0, 17, t, noalloc=True)
        assert small.size < big.size
        assert np.array_equal(small, small_ref), ("x slice", t)
    h = _stepped((16, 16, 16), steps=1)
    print("no-alloc slice: second context 16^3", flush=True)
    for t in (0, 3, 7, 9, 12):
        ref = h.slice(1, 5, t)
        got = h.slice(1, 5, t, noalloc=True)
        assert np.array_equal(got, ref), ("second context", t)
    # a refused point gives the zero slice on both
    assert not g.slice(2, 99, 0, noalloc=True).any()
    h.close()
    g.close()


def test_slice_noalloc_rejects_short_output():
    g = _stepped((16, 16, 16), steps=0)
    with pytest.raises(Rmt3dError):
        import ctypes as C
        from rmt3d_amd.lib import lib
        out = np.zeros(4)
        ln = C.c_int(4)
        rc = lib().rmt3d_slice_noalloc(g.h, 2, 3, 0, -1, out.ctypes.data_as(C.POINTER(C.c_double)), C.byref(ln))
        g._check(rc, "slice_noalloc")
    g.close()
