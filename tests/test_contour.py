"""Contour output of the oracle (CPU only): the marching-cubes table (two
independent derivations of the same face rule agree, and its triangle
count per box is the reference's), and the surface of a sphere is a closed,
consistently oriented triangulation of the zero level set."""
import importlib.util
import os
import re

import numpy as np
import pytest

from oracle.oracle import OracleSim, mc_table
from rmt3d_amd import configs

HERE = os.path.dirname(os.path.abspath(__file__))
GEN = os.path.join(os.path.dirname(HERE), "rmt3d_amd", "csrc", "gen_mc_table.py")
REF_IO = "/root/reference/F3D_Source/fluid_3d_io.cc"


def _gen():
    spec = importlib.util.spec_from_file_location("gen_mc_table", GEN)
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_table_generators_agree_with_header():
    g = _gen()
    py = g.table()
    assert mc_table() == py
    hdr = open(os.path.join(os.path.dirname(GEN), "mc_table.h")).read()
    ntri = [int(x) for x in re.search(r"mc_ntri\[256\] = \{([^}]*)\}", hdr).group(1).split(",")]
    assert ntri == [len(t) for t in py]


def test_table_polygons_close_and_cover_crossed_edges():
    g = _gen()
    for key, tris in enumerate(g.table()):
        pos = [(key >> c) & 1 for c in range(8)]
        crossed = {e for e, (a, b) in enumerate(g.EDGES) if pos[a] != pos[b]}
        used = {e for t in tris for e in t}
        assert used == crossed, key
        # each triangle's normal points to the positive ends of the edges
        # its vertices lie on
        for t in tris:
            n = g.cross(g.sub(g.mid(t[1]), g.mid(t[0])), g.sub(g.mid(t[2]), g.mid(t[0])))
            s_ = 0.0
            for e in t:
                a, b = g.EDGES[e]
                pc = a if pos[a] else b
                s_ += g.dot(n, g.sub(g.CORNERS[pc], g.mid(e)))
            assert s_ >= 0, (key, t)


@pytest.mark.skipif(not os.path.exists(REF_IO), reason="reference source absent")
def test_triangle_counts_equal_reference():
    src = open(REF_IO).read()
    blk = src[src.index("n_poly[256]={") + 13:]
    ref = [int(x) for x in re.findall(r"-?\d+", blk[:blk.index("}")])]
    assert ref == [len(t) for t in _gen().table()]


def test_sphere_surface_is_closed_and_on_the_level_set():
    cfg, objs = configs.cavity_sphere(24)
    S = OracleSim(cfg, objs)
    S.initialize()
    pts, nrm, tri = S.contour(1)
    assert len(pts) > 100 and len(tri) > 100
    c, R, h = np.array([0.5, 0.5, 0.4]), 0.2, 1.0 / 24
    r = np.linalg.norm(pts.astype(np.float64) - c, axis=1)
    assert np.all(np.abs(r - R) < 0.5 * h * h / R + 1e-6)
    assert np.allclose(np.linalg.norm(nrm, axis=1), 1, atol=1e-6)
    assert np.all(np.einsum("ij,ij->i", nrm.astype(np.float64), pts - c) > 0)
    # closed and consistently oriented: each directed side once, reversed once
    sides = {}
    for a, b, d in tri:
        for e in ((a, b), (b, d), (d, a)):
            assert e not in sides
            sides[e] = 1
    assert all((b, a) in sides for (a, b) in sides)
    V, E, F = len(pts), len(sides) // 2, len(tri)
    assert V - E + F == 2
    # triangle normals point outwards
    p = pts.astype(np.float64)
    tn = np.cross(p[tri[:, 1]] - p[tri[:, 0]], p[tri[:, 2]] - p[tri[:, 0]])
    assert np.all(np.einsum("ij,ij->i", tn, p[tri].mean(axis=1) - c) > 0)


