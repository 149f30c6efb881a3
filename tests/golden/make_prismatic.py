"""Writes the golden prismatic slabs from the reference's compiled code.

Runs where the reference tree is present: it needs oracle/_ref/libdisco_ref.so
(``make -C oracle ref``: Code/prismatic.cpp and Code/loop.cpp compiled from
their sources with ref_shim.cpp).  prismatic.npz holds, per case, the slab the
reference's prismaticSlab (Code/prismatic.cpp:30-74) selects -- atom indices,
sides and (type, planes, height, spacing, b) -- for a loop built by the
reference's own constructors on a lattice of generateLattice.  The inputs are
the case table below (loops, lattices, box) and are rebuilt by the tests.

    python tests/golden/make_prismatic.py
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from disco_amd import geometry as G  # noqa: E402

BCC, FCC = ("BCC", 2.8553, 12), ("FCC", 3.615, 10)

# (lattice, shape, params, centre offset in cells, m, n)
#   params as the reference shim reads them: circle [R]; rect [hm, hp, corner,
#   side xyz]; parallelogram [h1, h2, dir1, dir2]; polygon [sides, R, corner,
#   axis xyz]; polygon_vertices [count, corner, axis xyz, vertices]
CASES = [
    (BCC, "circle", [9.0], [0.3, -0.2, 0.1], [1, 1, 1], [1, 1, 1]),              # vacancy, 3 planes
    (BCC, "circle", [9.0], [0.3, -0.2, 0.1], [-1, -1, -1], [1, 1, 1]),           # interstitial
    (BCC, "rect", [6.0, 4.0, 1.5], [0.0, 0.1, -0.2], [2, 0, 0], [1, 0, 0]),      # <100>, 2 planes
    (BCC, "circle", [7.5], [-0.4, 0.2, 0.3], [1, 1, 1], [1, 1, 0]),              # mixed, 1 plane
    (BCC, "polygon", [6, 8.0, 2.0], [0.2, 0.2, -0.1], [-1, -1, -1], [1, 1, 1]),  # rounded hexagon
    (BCC, "parallelogram", [7.0, 5.0, 1, -1, 0, 1, 0, -1], [0.1, -0.3, 0.2], [1, 1, 1], [1, 1, 1]),
    (BCC, "circle", [6.0], [5.9, 5.8, -5.95], [2, 2, 2], [1, 1, 1]),             # 6 planes, at the box edge
    (BCC, "circle", [8.0], [0.0, 0.0, 0.0], [1, -1, 1], [1, 1, 0]),              # glide: no slab
    (FCC, "circle", [8.0], [0.2, -0.1, 0.3], [1, 1, 0], [1, 1, 0]),              # 2 planes
    (FCC, "rect", [7.0, 5.0, 2.0, 1, -1, 0], [-0.2, 0.3, 0.1], [0, 0, 2], [0, 0, 1]),  # Side axis
    (FCC, "polygon_vertices", None, [0.3, 0.1, -0.2], [-1, -1, 0], [1, 1, 0]),    # irregular
    (FCC, "polygon", [5, 9.0, 0.0, 0, 0, 1], [4.9, -4.8, 4.95], [1, 1, 0], [1, 1, 0]),  # axis, edge
]


def _irregular(m, n):
    """A convex pentagon (offsets from the centre) in the plane of n."""
    u, _ = G.default_axes(G._unit(m), G._unit(n))
    w = G._cross(G._unit(n), u)
    ang = [0.1, 1.3, 2.6, 3.9, 5.0]
    rad = [9.0, 7.0, 8.5, 6.5, 8.0]
    return [[r * np.cos(a) * u[i] + 0.7 * r * np.sin(a) * w[i] for i in range(3)]
            for a, r in zip(ang, rad)]


def case(k):
    """(G.Loop, shape code, params, lattice kind, a0, cells, spacing, box)."""
    (kind, a0, cells), shape, p, off, m, n = CASES[k]
    c0 = [o * a0 for o in off]
    c, spacing = G.native_snap_to_mid_plane(c0, n, a0, cells, cells, cells, kind)
    if shape == "circle":
        lp = G.Loop("circle", c, m, n, a0, radius=p[0])
    elif shape == "rect":
        side = list(p[3:6]) if len(p) > 3 else []
        lp = G.Loop("rect", c, m, n, a0, half_m=p[0], half_p=p[1], corner_r=p[2], axis=side)
    elif shape == "parallelogram":
        lp = G.Loop("parallelogram", c, m, n, a0, half_m=p[0], half_p=p[1], dir1=p[2:5],
                    dir2=p[5:8])
    elif shape == "polygon":
        axis = list(p[3:6]) if len(p) > 3 else []
        lp = G.Loop("polygon", c, m, n, a0, vertices=G.native_regular_polygon(int(p[0]), p[1], m, n),
                    corner_r=p[2], axis=axis)
    else:
        V = _irregular(m, n)
        p = [len(V), 0.0, 0.0, 0.0, 0.0, *np.ravel(V)]
        lp = G.Loop("polygon", c, m, n, a0, vertices=V)
    return lp, shape, list(p), kind, a0, cells, spacing, [cells * a0] * 3


def main():
    import oracle
    out = {}
    for k in range(len(CASES)):
        lp, shape, p, kind, a0, cells, spacing, box = case(k)
        x, y, z = G.lattice(kind, cells, cells, cells, a0)
        atoms, side, info = oracle.ref_prismatic_slab(shape, p, lp.center, lp.m, lp.n, a0,
                                                      spacing, x, y, z, box)
        out[f"atoms{k}"] = atoms
        out[f"side{k}"] = side.astype(np.int8)
        out[f"info{k}"] = info
        print(f"case {k}: {shape} type {int(info[0])} planes {int(info[1])}: {atoms.size} atoms")
    np.savez_compressed(os.path.join(HERE, "prismatic.npz"), **out)


if __name__ == "__main__":
    main()
