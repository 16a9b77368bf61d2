"""ctypes binding of librmt3d.so (the C ABI of include/rmt3d.h).

There is no fallback: if the HIP library is missing or fails to load, every
entry point raises.
"""
import ctypes as C
import os

import numpy as np

from . import _abi

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "librmt3d.so")

_lib = None

EXPORTS = [
    "rmt3d_parse_cfg", "rmt3d_default_config", "rmt3d_create", "rmt3d_destroy", "rmt3d_last_error",
    "rmt3d_initialize", "rmt3d_prepare_run", "rmt3d_init_iter", "rmt3d_step", "rmt3d_set_dt",
    "rmt3d_get_dt", "rmt3d_get_time", "rmt3d_set_time", "rmt3d_upload_state", "rmt3d_download_state",
    "rmt3d_band_count", "rmt3d_mg_solve", "rmt3d_mg_dims", "rmt3d_diag", "rmt3d_stage_times",
    "rmt3d_stream", "rmt3d_profile", "rmt3d_profile_read", "rmt3d_device_synchronize",
    "rmt3d_estimate_bytes", "rmt3d_server_create", "rmt3d_server_submit", "rmt3d_server_wait",
    "rmt3d_server_info", "rmt3d_server_error", "rmt3d_server_result", "rmt3d_server_release",
    "rmt3d_server_stats", "rmt3d_server_destroy", "rmt3d_write_checkpoint", "rmt3d_write_params",
    "rmt3d_slice", "rmt3d_slice_noalloc", "rmt3d_write_slice", "rmt3d_macro_data", "rmt3d_centroids", "rmt3d_write_tracers",
    "rmt3d_tracer_positions", "rmt3d_write_contours", "rmt3d_contour",
]


class Rmt3dError(RuntimeError):
    pass


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise Rmt3dError("librmt3d.so is not built (run __graft_entry__.build())")
        L = C.CDLL(LIB_PATH)
        P, D = C.c_void_p, C.c_double
        dp, ip = C.POINTER(C.c_double), C.POINTER(C.c_int)
        L.rmt3d_parse_cfg.argtypes = [C.c_char_p, C.POINTER(_abi.Config), C.POINTER(_abi.Object), C.c_int]
        L.rmt3d_default_config.argtypes = [C.POINTER(_abi.Config)]
        L.rmt3d_default_config.restype = None
        L.rmt3d_create.argtypes = [C.POINTER(_abi.Config), C.POINTER(_abi.Object), C.c_int, C.POINTER(P)]
        L.rmt3d_destroy.argtypes = [P]
        L.rmt3d_destroy.restype = None
        L.rmt3d_last_error.argtypes = [P]
        L.rmt3d_last_error.restype = C.c_char_p
        for f in ("rmt3d_initialize", "rmt3d_band_count"):
            getattr(L, f).argtypes = [P]
        L.rmt3d_prepare_run.argtypes = [P, C.c_int, dp, ip]
        L.rmt3d_init_iter.argtypes = [P, C.c_int]
        L.rmt3d_step.argtypes = [P, C.c_int]
        L.rmt3d_set_dt.argtypes = [P, D]
        L.rmt3d_set_time.argtypes = [P, D]
        L.rmt3d_get_dt.argtypes = [P]
        L.rmt3d_get_dt.restype = D
        L.rmt3d_get_time.argtypes = [P]
        L.rmt3d_get_time.restype = D
        L.rmt3d_upload_state.argtypes = [P, C.POINTER(_abi.State)]
        L.rmt3d_download_state.argtypes = [P, C.POINTER(_abi.State)]
        L.rmt3d_mg_solve.argtypes = [P, C.c_int, dp, dp, D, ip, dp]
        L.rmt3d_mg_dims.argtypes = [P, C.c_int, ip]
        L.rmt3d_diag.argtypes = [P, C.c_int, dp]
        L.rmt3d_stage_times.argtypes = [P, dp, C.c_int]
        L.rmt3d_stream.argtypes = [P]
        L.rmt3d_stream.restype = P
        L.rmt3d_profile.argtypes = [P, C.c_int]
        L.rmt3d_profile_read.argtypes = [P, C.c_int, dp, ip]
        L.rmt3d_device_synchronize.argtypes = [P]
        L.rmt3d_estimate_bytes.argtypes = [C.POINTER(_abi.Config), C.POINTER(_abi.Object), C.c_int, dp]
        L.rmt3d_server_create.argtypes = [D, C.c_int, C.POINTER(P)]
        L.rmt3d_server_submit.argtypes = [P, C.POINTER(_abi.Config), C.POINTER(_abi.Object), C.c_int, C.c_int, ip]
        L.rmt3d_server_wait.argtypes = [P, C.c_int, D, C.POINTER(_abi.RequestInfo)]
        L.rmt3d_server_info.argtypes = [P, C.c_int, C.POINTER(_abi.RequestInfo)]
        L.rmt3d_server_error.argtypes = [P, C.c_int]
        L.rmt3d_server_error.restype = C.c_char_p
        L.rmt3d_server_result.argtypes = [P, C.c_int, C.POINTER(_abi.State)]
        L.rmt3d_server_release.argtypes = [P, C.c_int]
        L.rmt3d_server_stats.argtypes = [P, dp, ip, dp]
        L.rmt3d_server_destroy.argtypes = [P]
        L.rmt3d_server_destroy.restype = None
        L.rmt3d_macro_data.argtypes = [P, C.POINTER(_abi.Macro)]
        L.rmt3d_write_tracers.argtypes = [P, C.c_char_p, C.c_int]
        L.rmt3d_write_contours.argtypes = [P, C.c_char_p]
        fp = C.POINTER(C.c_float)
        L.rmt3d_contour.argtypes = [P, C.c_int, fp, fp, ip, ip, ip]
        L.rmt3d_tracer_positions.argtypes = [P, C.POINTER(C.c_float), ip]
        L.rmt3d_centroids.argtypes = [P, dp]
        L.rmt3d_write_checkpoint.argtypes = [P, C.c_int, C.c_int, C.c_char_p]
        L.rmt3d_write_params.argtypes = [C.POINTER(_abi.Config), C.POINTER(_abi.Object), C.c_char_p]
        L.rmt3d_slice.argtypes = [P, C.c_int, C.c_int, C.c_int, C.c_int, dp, ip]
        L.rmt3d_slice_noalloc.argtypes = [P, C.c_int, C.c_int, C.c_int, C.c_int, dp, ip]
        L.rmt3d_write_slice.argtypes = [P, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_char_p]
        _lib = L
    return _lib


def parse_cfg(path, max_obj=256):
    cfg = _abi.Config()
    objs = (_abi.Object * max_obj)()
    rc = lib().rmt3d_parse_cfg(path.encode(), C.byref(cfg), objs, max_obj)
    if rc:
        raise Rmt3dError("rmt3d_parse_cfg(%s) failed with %d" % (path, rc))
    return cfg, [objs[i] for i in range(cfg.n_obj)]


def write_params(cfg, objs, dirname):
    """sim_params::write_params: <dirname>/recover.cfg."""
    n = max(1, len(objs))
    arr = (_abi.Object * n)(*objs) if objs else (_abi.Object * 1)()
    rc = lib().rmt3d_write_params(C.byref(cfg), arr, os.fsencode(dirname))
    if rc:
        raise Rmt3dError("rmt3d_write_params(%s) failed with %d" % (dirname, rc))


CHK_FILES = ("vx", "vy", "vz", "p", "X", "Y", "Z", "OID")


def compress_checkpoint(dirname):
    """Replace the eight checkpoint files by .xz files, as the reference's
    `xz` step after write_chk_pt does (fluid_3d_chk_pt.cc:54-60)."""
    import lzma
    for name in CHK_FILES:
        path = os.path.join(dirname, name)
        with open(path, "rb") as f:
            data = f.read()
        with lzma.open(path + ".xz", "wb", format=lzma.FORMAT_XZ, check=lzma.CHECK_CRC64) as f:
            f.write(data)
        os.remove(path)


def decompress_checkpoint(dirname):
    """The reference's `xz -d` before reading (fluid_3d_chk_pt.cc:109-114):
    every <name>.xz present is expanded to <name> and removed."""
    import lzma
    for name in CHK_FILES:
        path = os.path.join(dirname, name)
        if os.path.exists(path + ".xz"):
            with lzma.open(path + ".xz", "rb") as f:
                data = f.read()
            with open(path, "wb") as f:
                f.write(data)
            os.remove(path + ".xz")


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


class Simulation:
    """One device-resident simulation (fluid_3d on the GPU)."""

    def __init__(self, cfg, objs, precision=_abi.PREC_FP64):
        n = max(1, len(objs))
        self._objs = (_abi.Object * n)(*objs) if objs else (_abi.Object * 1)()
        self.cfg = cfg
        h = C.c_void_p()
        rc = lib().rmt3d_create(C.byref(cfg), self._objs, precision, C.byref(h))
        if rc:
            raise Rmt3dError("rmt3d_create failed (%d): %s" % (rc, lib().rmt3d_last_error(None).decode()))
        self.h = h
        m, nn, o = cfg.sys_size
        self.Ng = (m + 4) * (nn + 4) * (o + 4)

    def _check(self, rc, what):
        if rc < 0:
            raise Rmt3dError("%s failed (%d): %s" % (what, rc, lib().rmt3d_last_error(self.h).decode()))
        return rc

    def close(self):
        if getattr(self, "h", None):
            lib().rmt3d_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def initialize(self):
        return self._check(lib().rmt3d_initialize(self.h), "initialize")

    def init_iter(self, e):
        return self._check(lib().rmt3d_init_iter(self.h, e), "init_iter")

    def prepare_run(self, init_err):
        dt = C.c_double()
        tot = C.c_int()
        self._check(lib().rmt3d_prepare_run(self.h, init_err, C.byref(dt), C.byref(tot)), "prepare_run")
        return 0, dt.value, tot.value

    def step(self, debug=0):
        return self._check(lib().rmt3d_step(self.h, debug), "step")

    @property
    def dt(self):
        return lib().rmt3d_get_dt(self.h)

    @property
    def time(self):
        return lib().rmt3d_get_time(self.h)

    def set_dt(self, dt):
        lib().rmt3d_set_dt(self.h, dt)

    def set_time(self, t):
        lib().rmt3d_set_time(self.h, t)

    def diag(self, which):
        out = C.c_double()
        self._check(lib().rmt3d_diag(self.h, which, C.byref(out)), "diag")
        return out.value

    def step_counters(self):
        """Branch counters of the last step: (collision contributions, primary
        maps redrawn into the band, band maps redrawn onto the primary grid)."""
        return tuple(int(self.diag(w)) for w in (20, 21, 22))

    def div_u(self):
        return self.diag(0)

    def band_count(self):
        return self._check(lib().rmt3d_band_count(self.h), "band_count")

    def download(self, band=True):
        Ng = self.Ng
        f = {k: np.zeros(Ng) for k in ("u", "v", "w", "p", "dp", "q")}
        f["dvel"] = np.zeros(3 * Ng)
        f["fvel"] = np.zeros(18 * Ng)
        f["sigma"] = np.zeros(9 * Ng)
        f["rm_c"] = np.zeros(Ng, dtype=np.uint32)
        f["rm_x"] = np.zeros(3 * Ng)
        f["rm_xpred"] = np.zeros(3 * Ng)
        st = _abi.State()
        for k in ("u", "v", "w", "p", "dp", "q", "dvel", "fvel", "sigma", "rm_x", "rm_xpred"):
            setattr(st, k, _dp(f[k]))
        st.rm_c = f["rm_c"].ctypes.data_as(C.POINTER(C.c_uint32))
        if band:
            nb = self.band_count()
            f["band_cell"] = np.zeros(nb, dtype=np.int32)
            f["band_c"] = np.zeros(nb, dtype=np.uint32)
            f["band_x"] = np.zeros(3 * nb)
            f["band_xpred"] = np.zeros(3 * nb)
            st.nband = nb
            st.band_cell = f["band_cell"].ctypes.data_as(C.POINTER(C.c_int32))
            st.band_c = f["band_c"].ctypes.data_as(C.POINTER(C.c_uint32))
            st.band_x = _dp(f["band_x"])
            st.band_xpred = _dp(f["band_xpred"])
        self._check(lib().rmt3d_download_state(self.h, C.byref(st)), "download")
        return f

    def upload(self, **fields):
        st = _abi.State()
        keep = []
        for k, a in fields.items():
            if k == "rm_c":
                a = np.ascontiguousarray(a, dtype=np.uint32)
                st.rm_c = a.ctypes.data_as(C.POINTER(C.c_uint32))
            else:
                a = np.ascontiguousarray(a, dtype=np.float64)
                setattr(st, k, _dp(a))
            keep.append(a)
        self._check(lib().rmt3d_upload_state(self.h, C.byref(st)), "upload")

    def set_band(self, cell, c, x, xpred):
        """Replace the band (cells ascending; x, xpred as (3, n))."""
        st = _abi.State()
        cell = np.ascontiguousarray(cell, dtype=np.int32)
        c = np.ascontiguousarray(c, dtype=np.uint32)
        x = np.ascontiguousarray(x, dtype=np.float64)
        xpred = np.ascontiguousarray(xpred, dtype=np.float64)
        st.nband = len(cell)
        st.band_cell = cell.ctypes.data_as(C.POINTER(C.c_int32))
        st.band_c = c.ctypes.data_as(C.POINTER(C.c_uint32))
        st.band_x = _dp(x)
        st.band_xpred = _dp(xpred)
        self._check(lib().rmt3d_upload_state(self.h, C.byref(st)), "upload band")

    # ---- diagnostics (stepper::collect_macro_data) ----
    def macro_data(self):
        m = _abi.Macro()
        self._check(lib().rmt3d_macro_data(self.h, C.byref(m)), "macro_data")
        return {k: getattr(m, k) for k in _abi.MACRO_FIELDS}

    def centroids(self):
        """compute_v_centroid: (n_obj, 6) centroid x y z, velocity x y z;
        the average velocities are kept for the wall repulsion."""
        k = max(1, self.cfg.n_obj)
        out = np.zeros(6 * k)
        self._check(lib().rmt3d_centroids(self.h, _dp(out)), "centroids")
        return out.reshape(k, 6)[:self.cfg.n_obj]

    # ---- contours ----
    def contour(self, obj):
        """Object obj's (1-based) contour: (pts, normals, triangles)."""
        nv, nt = C.c_int(0), C.c_int(0)
        self._check(lib().rmt3d_contour(self.h, obj, None, None, C.byref(nv), None, C.byref(nt)), "contour")
        pts = np.zeros((max(nv.value, 1), 3), dtype=np.float32)
        nrm = np.zeros_like(pts)
        tri = np.zeros((max(nt.value, 1), 3), dtype=np.int32)
        fp = C.POINTER(C.c_float)
        self._check(lib().rmt3d_contour(self.h, obj, pts.ctypes.data_as(fp), nrm.ctypes.data_as(fp), C.byref(nv),
                                        tri.ctypes.data_as(C.POINTER(C.c_int)), C.byref(nt)), "contour")
        return pts[:nv.value], nrm[:nv.value], tri[:nt.value]

    def write_contours(self, filename):
        self._check(lib().rmt3d_write_contours(self.h, os.fsencode(filename)), "write_contours")

    # ---- tracers ----
    def tracer_positions(self):
        """tracers::print's positions: (n, 3) float32 by tracer id."""
        n = C.c_int(0)
        self._check(lib().rmt3d_tracer_positions(self.h, None, C.byref(n)), "tracer_positions")
        out = np.zeros((max(n.value, 1), 3), dtype=np.float32)
        self._check(lib().rmt3d_tracer_positions(self.h, out.ctypes.data_as(C.POINTER(C.c_float)), C.byref(n)),
                    "tracer_positions")
        return out[:n.value]

    def write_tracers(self, filename, binary=True):
        self._check(lib().rmt3d_write_tracers(self.h, os.fsencode(filename), int(binary)), "write_tracers")

    # ---- checkpoint and slice formats ----
    def write_checkpoint(self, snum, fnum, dirname):
        """fluid_3d::write_chk_pt into the existing directory dirname."""
        self._check(lib().rmt3d_write_checkpoint(self.h, snum, fnum, os.fsencode(dirname)), "write_checkpoint")

    def slice(self, dim, point, o_type, obj_id=-1, noalloc=False):
        """fluid_3d::slice: the global slice as a (d2, d1) array (noalloc:
        through rmt3d_slice_noalloc, staged in the library's fixed scratch)."""
        m, n, o = self.cfg.sys_size
        cap = (max(m, n, o) + 1) ** 2
        out = np.zeros(cap)
        ln = C.c_int(cap)
        fn = lib().rmt3d_slice_noalloc if noalloc else lib().rmt3d_slice
        self._check(fn(self.h, dim, point, o_type, obj_id, _dp(out), C.byref(ln)), "slice")
        corner = 3 <= o_type < 6 or o_type in (7, 18)
        g = [self.cfg.sys_size[d] + (1 if corner and not self.cfg.prd[d] else 0) for d in range(3)]
        rest = [d for d in range(3) if d != dim]
        return out[:ln.value].reshape(g[rest[1]], g[rest[0]])

    def write_slice(self, dim, point, o_type, filename, obj_id=-1, fmt=0):
        """fluid_3d::write_slice (fmt 0 matrix, 1 text, 2 gnuplot binary)."""
        self._check(lib().rmt3d_write_slice(self.h, dim, point, o_type, obj_id, fmt, os.fsencode(filename)),
                    "write_slice")

    def mg_dims(self, which):
        d = (C.c_int * 3)()
        self._check(lib().rmt3d_mg_dims(self.h, which, d), "mg_dims")
        return tuple(d)

    def mg_solve(self, which, rhs, x0, tol=0.0):
        x = np.ascontiguousarray(x0, dtype=np.float64).copy()
        r = np.ascontiguousarray(rhs, dtype=np.float64)
        nc = C.c_int()
        res = C.c_double()
        self._check(lib().rmt3d_mg_solve(self.h, which, _dp(r), _dp(x), tol, C.byref(nc), C.byref(res)),
                    "mg_solve")
        return x, nc.value, res.value

    def profile(self, level=1):
        """0 off, 1 stage buckets and V-cycles, 2 also per-kernel classes
        (V-cycles then launch eagerly instead of replaying their graph)."""
        self._check(lib().rmt3d_profile(self.h, int(level)), "profile")

    def profile_read(self, which):
        ms = C.c_double()
        n = C.c_int()
        self._check(lib().rmt3d_profile_read(self.h, which, C.byref(ms), C.byref(n)), "profile_read")
        return ms.value, n.value

    STAGES = ("extrapolation", "pressure_poisson", "communication", "stresses", "mac", "advection")

    def stage_times(self):
        """Device ms per stage bucket over the steps run while profiling."""
        ms = (C.c_double * 6)()
        self._check(lib().rmt3d_stage_times(self.h, ms, 6), "rmt3d_stage_times")
        return dict(zip(self.STAGES, list(ms)))

    def synchronize(self):
        self._check(lib().rmt3d_device_synchronize(self.h), "synchronize")
