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
