"""ctypes mirror of the C ABI structs declared in include/rmt3d.h.

The field order and types follow the header exactly; the structs are the
POD mirror of the reference's sim_params fields on the step path
(Sim/sim_params.hh) and of its per-object .cfg keys (Sim/sim_params.cc:275-314).
"""
import ctypes as C

N_EXTRA_SPECS = 10
MAX_VEL_PROF = 8

SIM_FTEST, SIM_STEST, SIM_OBJECTS, SIM_MILLS = 0, 1, 2, 3
BASIC_SPHERE, SPHERE = 0, 1
PREC_FP64, PREC_MIXED = 0, 1
# step return codes (Utilities/defs.hh:47-50)
NORMAL_EXIT, REFMAP_UPDATE_ERR, PROJECTION_ERR = 0, 1, 2


class Object(C.Structure):
    _fields_ = [
        ("type", C.c_int),
        ("basic", C.c_double * 4),
        ("extra", C.c_double * N_EXTRA_SPECS),
        ("srho", C.c_double),
        ("shear_mod", C.c_double),
    ]


class Config(C.Structure):
    _fields_ = [
        ("sys_size", C.c_int * 3),
        ("prd", C.c_int * 3),
        ("xyz_bounds", C.c_double * 6),
        ("bct", (C.c_int * 6) * 3),
        ("bcv", (C.c_double * 6) * 3),
        ("walls", C.c_int * 6),
        ("wall_pos", C.c_double * 6),
        ("wall_dist", C.c_double),
        ("fmu", C.c_double),
        ("frho", C.c_double),
        ("fdt_pad", C.c_double),
        ("dt", C.c_double),
        ("T", C.c_double),
        ("frames", C.c_int),
        ("stop_short", C.c_int),
        ("num_iters", C.c_int),
        ("nlayers", C.c_int),
        ("max_extrap_rs", C.c_int),
        ("wt_n", C.c_double),
        ("ex_visc_mult", C.c_double),
        ("ev_trans_mult", C.c_double),
        ("sdt_pad", C.c_double),
        ("dt_ex_pad", C.c_double),
        ("weight_fac", C.c_double),
        ("gravity", C.c_double),
        ("godunov", C.c_int),
        ("impl", C.c_int),
        ("has_exact", C.c_int),
        ("sim_type", C.c_int),
        ("vel_prof_num", C.c_int),
        ("vel_profile", C.c_double * (2 * MAX_VEL_PROF)),
        ("dump_code", C.c_int),
        ("ntracers", C.c_int),
        ("debug_flag", C.c_int),
        ("chk_num", C.c_int),
        ("chk_step", C.c_int),
        ("n_obj", C.c_int),
        ("suppress", C.c_int),
        ("display", C.c_int),
        ("chkpt_freq", C.c_int),
        ("omp_num_thr", C.c_int),
        ("debug_obj_id", C.c_int),
        ("obj_body", C.c_int),
        ("data_fmt", C.c_int),
        ("out_flag", C.c_int),
        ("output_dim", C.c_int),
        ("output_ind", C.c_int),
        ("current_time", C.c_double),
        ("dirname", C.c_char * 256),
        ("recover_dirname", C.c_char * 256),
        ("chk_dirname", C.c_char * 256),
    ]


MACRO_FIELDS = ("divu_max", "divu_min", "divu_l2", "momentum_x", "momentum_y", "momentum_z", "energy", "kinetic",
                "elastic", "potential", "power", "max_dev_detF", "dev_detF_l2")


class Macro(C.Structure):
    """rmt3d_macro (include/rmt3d.h): the macro.dat quantities."""
    _fields_ = [(k, C.c_double) for k in MACRO_FIELDS]


class State(C.Structure):
    _fields_ = [
        ("u", C.POINTER(C.c_double)),
        ("v", C.POINTER(C.c_double)),
        ("w", C.POINTER(C.c_double)),
        ("p", C.POINTER(C.c_double)),
        ("dp", C.POINTER(C.c_double)),
        ("q", C.POINTER(C.c_double)),
        ("dvel", C.POINTER(C.c_double)),
        ("fvel", C.POINTER(C.c_double)),
        ("sigma", C.POINTER(C.c_double)),
        ("rm_c", C.POINTER(C.c_uint32)),
        ("rm_x", C.POINTER(C.c_double)),
        ("rm_xpred", C.POINTER(C.c_double)),
        ("nband", C.c_int),
        ("band_cell", C.POINTER(C.c_int32)),
        ("band_c", C.POINTER(C.c_uint32)),
        ("band_x", C.POINTER(C.c_double)),
        ("band_xpred", C.POINTER(C.c_double)),
    ]


def default_config():
    """Defaults of sim_params (Sim/sim_params.cc:10-41, 326-372) for a
    periodic 96^3 box; callers overwrite what they need and call
    finalize_config()."""
    c = Config()
    for i in range(3):
        c.sys_size[i] = 96
        c.prd[i] = 1
    for i, v in enumerate([0, 1, 0, 1, 0, 1]):
        c.xyz_bounds[i] = v
        c.wall_pos[i] = v
    c.wall_dist = 0
    c.fmu, c.frho, c.fdt_pad = 0.002, 1.0, 0.5
    c.dt, c.T = 1.0, 1.0
    c.frames, c.stop_short, c.num_iters = 100, -1, 5
    c.nlayers, c.max_extrap_rs = 8, 5
    c.wt_n, c.ex_visc_mult, c.ev_trans_mult = 2.5, 1.0, 1.0
    c.sdt_pad, c.dt_ex_pad, c.weight_fac, c.gravity = 0.25, 0.8, 0.25, 0.0
    c.godunov, c.impl, c.has_exact = 1, 0, 0
    c.sim_type = SIM_FTEST
    c.dump_code, c.ntracers = 3, 1000
    c.n_obj = 0
    c.chkpt_freq, c.omp_num_thr, c.debug_obj_id, c.obj_body = 1, 1, -1, -1
    c.out_flag, c.output_dim, c.output_ind = 1 | 2 | 4 | 128, 2, -1
    return c


def finalize_config(c):
    """The post-parse consolidation of sim_params (Sim/sim_params.cc:326-372):
    boundary types decide periodicity, walls follow, dx from the bounds and
    wall_dist >= min dh."""
    for d in range(3):
        s = c.bct[0][2 * d] + c.bct[1][2 * d] + c.bct[2][2 * d]
        if 0 <= s < 3:
            for comp in range(3):
                c.bct[comp][2 * d] = 0
                c.bct[comp][2 * d + 1] = 0
            c.prd[d] = 1
        else:
            c.prd[d] = 0
    for d in range(3):
        if not c.prd[d]:
            c.walls[2 * d] = c.walls[2 * d + 1] = 1
    if c.output_ind == -1 and 0 <= c.output_dim <= 2:
        c.output_ind = c.sys_size[c.output_dim] // 2
    dh = min((c.xyz_bounds[2 * d + 1] - c.xyz_bounds[2 * d]) / c.sys_size[d] for d in range(3))
    c.wall_dist = max(dh, c.wall_dist)
    return c


REQ_QUEUED, REQ_RUNNING, REQ_DONE, REQ_FAILED = 0, 1, 2, 3


class RequestInfo(C.Structure):
    """rmt3d_request_info (include/rmt3d.h)."""
    _fields_ = [
        ("status", C.c_int),
        ("rc", C.c_int),
        ("steps_done", C.c_int),
        ("cell_steps", C.c_longlong),
        ("t_submit", C.c_double),
        ("t_admit", C.c_double),
        ("t_done", C.c_double),
        ("device_bytes", C.c_double),
        ("sim_time", C.c_double),
    ]
