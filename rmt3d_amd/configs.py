"""Workload definitions of the hot-path benchmarks, as rmt3d_config structs.

Each builder restates a reference sample .cfg (sim_configs/*.cfg) with the
overrides of the workload it names; the same workloads ship as .cfg files in
rmt3d_amd/sim_configs/ for the C++ driver.

  taylor_green(n)      n^3 triply periodic unit cube, fmu 0.01, fdt_pad 0.4,
                       no objects; the Taylor-Green field is uploaded as the
                       initial state (the reference injects it through a
                       checkpoint, F3D_Source/fluid_3d_chk_pt.cc:74-163).
  cavity_sphere(n)     lid_driven_with_sphere.cfg at n^3: no-slip walls, lid
                       (1,0,0) on z-high, fmu 0.01, one basic sphere r=0.2 at
                       (0.5,0.5,0.4), G=1, nlayers 4.
  stretched_sphere(n)  pre-stretched_sphere.cfg at n^3 with no-slip walls,
                       fmu 0.01, sphere r=0.25 at the centre pre-stretched to
                       F=diag(1.3, 1.3^-1/2, 1.3^-1/2), G=1, nlayers 4.
  shear_spheres(n, k)  n^3 box periodic in x and y, no-slip walls at z = 0
                       and 1 moving at -0.5 and +0.5 along x (simple shear),
                       fmu 0.01; k spheres with radius uniform in
                       [0.04, 0.08] and G log-uniform in [0.5, 5], density
                       ratio 1, placed by random sequential addition without
                       overlap from a fixed seed, each kept inside the box
                       (the reference map of a sphere is not shifted across
                       a periodic boundary, fluid_3d_comm.cc).
"""
import math

from ._abi import (BASIC_SPHERE, SIM_OBJECTS, SPHERE, Object, default_config,
                   finalize_config)


def _walls(c, lid=None):
    for f in range(6):
        for comp in range(3):
            c.bct[comp][f] = 1
            c.bcv[comp][f] = 0.0
    if lid is not None:
        for comp in range(3):
            c.bcv[comp][5] = lid[comp]


def sphere(center, radius, G=1.0, srho=1.0, stretch_dir=None, sqrt_lam=None):
    o = Object()
    o.type = BASIC_SPHERE if stretch_dir is None else SPHERE
    for i in range(3):
        o.basic[i] = center[i]
    o.basic[3] = radius
    if stretch_dir is None:
        o.extra[0] = 1.0              # sim_params' default extra_specs row
    else:
        o.extra[0] = float(stretch_dir)
        o.extra[1] = 1.0 if sqrt_lam is None else sqrt_lam
    o.srho = srho
    o.shear_mod = G
    return o


def taylor_green(n=32, steps=20):
    c = default_config()
    for i in range(3):
        c.sys_size[i] = n
    c.fmu = 0.01
    c.fdt_pad = 0.4
    c.dt = 10.0
    c.sim_type = SIM_OBJECTS
    c.n_obj = 0
    c.stop_short = steps
    c.dump_code = 0
    finalize_config(c)
    return c, []


def cavity_sphere(n=64, steps=500):
    c = default_config()
    for i in range(3):
        c.sys_size[i] = n
    _walls(c, lid=(1.0, 0.0, 0.0))
    c.fmu = 0.01
    c.dt = 10.0
    c.T = 50.0
    c.frames = 500
    c.sim_type = SIM_OBJECTS
    c.nlayers = 4
    c.wt_n = 1.5
    c.ex_visc_mult = 1.0
    c.ev_trans_mult = 0.0
    c.stop_short = steps
    c.dump_code = 0
    for i, v in enumerate([-0.2, 1.2, -0.2, 1.2, -0.2, 1.2]):
        c.wall_pos[i] = v
    c.n_obj = 1
    finalize_config(c)
    return c, [sphere((0.5, 0.5, 0.4), 0.2, G=1.0)]


def stretched_sphere(n=128, steps=1000, lam=1.3):
    c = default_config()
    for i in range(3):
        c.sys_size[i] = n
    _walls(c)
    c.fmu = 0.01
    c.dt = 10.0
    c.T = 8.0
    c.frames = 500
    c.sim_type = SIM_OBJECTS
    c.nlayers = 4
    c.wt_n = 1.0
    c.ex_visc_mult = 1.0
    c.ev_trans_mult = 0.0
    c.stop_short = steps
    c.dump_code = 0
    c.n_obj = 1
    finalize_config(c)
    return c, [sphere((0.5, 0.5, 0.5), 0.25, G=1.0, stretch_dir=0, sqrt_lam=math.sqrt(lam))]


def taylor_green_field(cfg):
    """u = (sin 2pi x cos 2pi y cos 2pi z, -cos 2pi x sin 2pi y cos 2pi z, 0) on
    the padded grid (cell centres, ghosts by periodic wrap)."""
    import numpy as np
    m, n, o = cfg.sys_size
    b = cfg.xyz_bounds
    dx = (b[1] - b[0]) / m
    x = b[0] + (np.arange(m + 4) - 2 + 0.5) * dx
    y = b[2] + (np.arange(n + 4) - 2 + 0.5) * ((b[3] - b[2]) / n)
    z = b[4] + (np.arange(o + 4) - 2 + 0.5) * ((b[5] - b[4]) / o)
    Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
    tp = 2 * np.pi
    u = np.sin(tp * X) * np.cos(tp * Y) * np.cos(tp * Z)
    v = -np.cos(tp * X) * np.sin(tp * Y) * np.cos(tp * Z)
    w = np.zeros_like(u)
    return u.ravel(), v.ravel(), w.ravel()


def shear_spheres(n=256, count=64, steps=300, seed=20261202, rmin=0.04, rmax=0.08):
    import numpy as np
    rng = np.random.default_rng(seed)
    c = default_config()
    for i in range(3):
        c.sys_size[i] = n
    c.prd[0] = c.prd[1] = 1
    for f in (4, 5):
        for comp in range(3):
            c.bct[comp][f] = 1
            c.bcv[comp][f] = 0.0
    c.bcv[0][4] = -0.5
    c.bcv[0][5] = 0.5
    c.fmu = 0.01
    c.dt = 10.0
    c.T = 30.0
    c.frames = 300
    c.sim_type = SIM_OBJECTS
    c.nlayers = 4
    c.wt_n = 1.5
    c.ex_visc_mult = 1.0
    c.ev_trans_mult = 0.0
    c.stop_short = steps
    c.dump_code = 0
    objs = []
    placed = []
    for _ in range(200000):
        if len(placed) == count:
            break
        r = rng.uniform(rmin, rmax)
        ctr = rng.uniform(r + 0.02, 1 - r - 0.02, 3)
        if all(np.linalg.norm(ctr - c2) > r + r2 + 0.02 for c2, r2 in placed):
            placed.append((ctr, r))
    for ctr, r in placed:
        G = float(np.exp(rng.uniform(np.log(0.5), np.log(5.0))))
        objs.append(sphere(tuple(ctr), r, G=G))
    c.n_obj = len(objs)
    finalize_config(c)
    return c, objs


_SIM_TYPES = {0: "ftest", 1: "stest", 2: "objects", 3: "object_mills"}


def write_cfg(cfg, objs, path, comment=""):
    """Write cfg/objs as a .cfg file in the reference's keyword format
    (Sim/sim_params.cc), readable by rmt3d_parse_cfg and by the reference."""
    def f(v):
        return repr(float(v))
    lines = []
    for ln in comment.splitlines():
        lines.append("# " + ln)
    lines.append("#Keyword                  Input")
    put = lambda k, *v: lines.append("%-25s %s" % (k, " ".join(str(x) for x in v)))
    put("frames", cfg.frames)
    put("dump_code", cfg.dump_code)
    put("x_prd", cfg.prd[0]); put("y_prd", cfg.prd[1]); put("z_prd", cfg.prd[2])
    for fi, name in enumerate(("bc_xl", "bc_xh", "bc_yl", "bc_yh", "bc_zl", "bc_zh")):
        put(name, *([cfg.bct[c][fi] for c in range(3)] + [f(cfg.bcv[c][fi]) for c in range(3)]))
    put("walls", *[cfg.walls[i] for i in range(6)])
    put("wall_pos", *[f(cfg.wall_pos[i]) for i in range(6)])
    put("sys_size", *list(cfg.sys_size))
    put("xyz_bounds", *[f(cfg.xyz_bounds[i]) for i in range(6)])
    put("dt", f(cfg.dt))
    put("fdt_pad", f(cfg.fdt_pad))
    put("T", f(cfg.T))
    put("stop_short", cfg.stop_short)
    put("sim_type", _SIM_TYPES[cfg.sim_type])
    put("fmu", f(cfg.fmu))
    put("nlayers", cfg.nlayers)
    put("max_extrap_rs", cfg.max_extrap_rs)
    put("wt_n", f(cfg.wt_n))
    put("ex_visc_mult", f(cfg.ex_visc_mult))
    put("ev_trans_mult", f(cfg.ev_trans_mult))
    put("weight_fac", f(cfg.weight_fac))
    put("n_obj", len(objs))
    for i, o in enumerate(objs):
        put("object_list", i, o.type)
        put("basic_specs", i, *[f(o.basic[a]) for a in range(4)])
        put("extra_specs", i, *[f(o.extra[a]) for a in range(10)])
        put("srho", i, f(o.srho))
        put("shear_mod", i, f(o.shear_mod))
    with open(path, "w") as fh:
        fh.write("\n".join(lines) + "\n")


def manufactured(n=32, fmu=0.01):
    """sim_ftest (Sim/sim_manager.cc:139-178): the manufactured solution
    u = cos(2 pi t) (-sin 2pi x cos 2pi y cos 2pi z, 1/2 cos sin cos, 1/2 cos cos sin)
    with its forcing, on the triply periodic unit cube (the field is
    1-periodic; the reference's sample ftest.cfg uses walls, where this
    velocity does not vanish)."""
    c = default_config()
    for i in range(3):
        c.sys_size[i] = n
    c.fmu = fmu
    c.dt = 10.0
    c.fdt_pad = 0.5
    c.T = 1.0
    c.frames = 100
    c.sim_type = 0
    c.n_obj = 0
    c.dump_code = 0
    finalize_config(c)
    return c, []


def manufactured_velocity(cfg, t):
    """The exact velocity of manufactured() at time t on the padded grid."""
    import numpy as np
    m, n, o = cfg.sys_size
    b = cfg.xyz_bounds
    x = b[0] + (np.arange(m + 4) - 1.5) * ((b[1] - b[0]) / m)
    y = b[2] + (np.arange(n + 4) - 1.5) * ((b[3] - b[2]) / n)
    z = b[4] + (np.arange(o + 4) - 1.5) * ((b[5] - b[4]) / o)
    Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
    tp = 2 * np.pi
    ct = np.cos(tp * t)
    u = -np.sin(tp * X) * np.cos(tp * Y) * np.cos(tp * Z) * ct
    v = 0.5 * np.cos(tp * X) * np.sin(tp * Y) * np.cos(tp * Z) * ct
    w = 0.5 * np.cos(tp * X) * np.cos(tp * Y) * np.sin(tp * Z) * ct
    return u.ravel(), v.ravel(), w.ravel()
