"""Pins of the prismatic slab oracle (oracle/disco_prismatic.c): the golden
slabs of the reference's prismaticSlab, the compiled reference itself on random
loops of every shape (where its build is present), and an exact count for a
loop aligned with the lattice."""
import os
import sys

import numpy as np
import pytest

import oracle
from conftest import GOLDEN, golden
from disco_amd import geometry as G

sys.path.insert(0, GOLDEN)
import make_prismatic as MP  # noqa: E402


@pytest.mark.parametrize("k", range(len(MP.CASES)))
def test_oracle_matches_golden_slabs(k):
    g = golden("prismatic.npz")
    lp, shape, p, kind, a0, cells, spacing, box = MP.case(k)
    x, y, z = G.lattice(kind, cells, cells, cells, a0)
    st = G.native_loop_state(lp)
    atoms, side, info = oracle.prismatic_slab(st, spacing, x, y, z, box)
    assert np.array_equal(atoms, g[f"atoms{k}"])
    assert np.array_equal(side, g[f"side{k}"].astype(np.int32))
    assert np.array_equal(info, g[f"info{k}"])
    assert oracle.matter_type(st) == int(info[0])


def random_prismatic_loops(rng, count):
    """Prismatic and mixed loops of every shape whose |b.n| is a whole number of
    plane spacings, in BCC and FCC: (G.Loop, shape code, params, lattice kind,
    a0, cells, spacing, box)."""
    combos = {"BCC": [([1, 1, 1], [1, 1, 1]), ([-1, -1, -1], [1, 1, 1]), ([2, 0, 0], [1, 0, 0]),
                      ([1, 1, 1], [1, 1, 0]), ([-1, 1, 1], [0, 1, 1]), ([2, 2, 2], [1, 1, 1])],
              "FCC": [([1, 1, 0], [1, 1, 0]), ([-1, -1, 0], [1, 1, 0]), ([0, 0, 2], [0, 0, 1]),
                      ([0, 1, 1], [0, 1, 1]), ([-2, 0, 0], [1, 0, 0])]}
    out = []
    for t in range(count):
        kind = "BCC" if t % 2 == 0 else "FCC"
        a0 = 2.8553 if kind == "BCC" else 3.615
        cells = int(rng.integers(8, 15))
        m, n = combos[kind][int(rng.integers(len(combos[kind])))]
        c0 = list(rng.uniform(-0.5, 0.5, 3) * cells * a0)
        c, spacing = G.native_snap_to_mid_plane(c0, n, a0, cells, cells, cells, kind)
        R = float(rng.uniform(3.0, 0.45 * cells * a0))
        u, _ = G.default_axes(G._unit(m), G._unit(n))
        w = G._cross(G._unit(n), u)
        sh = t % 5
        if sh == 0:
            lp, code, p = G.Loop("circle", c, m, n, a0, radius=R), "circle", [R]
        elif sh == 1:
            hp, rc = R * float(rng.uniform(0.4, 1)), float(rng.uniform(0, 0.5)) * R
            ang = float(rng.uniform(0, 3))
            side = [np.cos(ang) * u[i] + np.sin(ang) * w[i] for i in range(3)] if t % 3 else []
            lp = G.Loop("rect", c, m, n, a0, half_m=R, half_p=hp, corner_r=rc, axis=side)
            code, p = "rect", [R, hp, rc, *(side or [0.0, 0.0, 0.0])]
        elif sh == 2:
            ang = float(rng.uniform(0.4, 1.3))
            d2 = [np.cos(ang) * u[i] + np.sin(ang) * w[i] for i in range(3)]
            lp = G.Loop("parallelogram", c, m, n, a0, half_m=R, half_p=0.6 * R, dir1=u, dir2=d2)
            code, p = "parallelogram", [R, 0.6 * R, *u, *d2]
        elif sh == 3:
            sides, rc = int(rng.integers(3, 10)), float(rng.uniform(0, 0.4)) * R
            lp = G.Loop("polygon", c, m, n, a0,
                        vertices=G.native_regular_polygon(sides, R, m, n), corner_r=rc)
            code, p = "polygon", [sides, R, rc]
        else:
            nv = int(rng.integers(3, 8))
            ang = np.linspace(0, 2 * np.pi, nv, endpoint=False) + rng.uniform(0, 0.5, nv)
            V = [[R * np.cos(a) * u[i] + 0.7 * R * np.sin(a) * w[i] for i in range(3)] for a in ang]
            rc = float(rng.uniform(0, 0.2)) * R
            lp = G.Loop("polygon", c, m, n, a0, vertices=V, corner_r=rc)
            code, p = "polygon_vertices", [nv, rc, 0.0, 0.0, 0.0, *np.ravel(V)]
        out.append((lp, code, p, kind, a0, cells, spacing, [cells * a0] * 3))
    return out


@pytest.mark.skipif(not oracle.ref_available(), reason="reference build not present")
def test_oracle_matches_compiled_reference_on_random_loops():
    rng = np.random.default_rng(31)
    total = 0
    for lp, code, p, kind, a0, cells, spacing, box in random_prismatic_loops(rng, 60):
        x, y, z = G.lattice(kind, cells, cells, cells, a0)
        ra, rs, ri = oracle.ref_prismatic_slab(code, p, lp.center, lp.m, lp.n, a0, spacing, x, y,
                                               z, box)
        oa, os_, oi = oracle.prismatic_slab(G.native_loop_state(lp), spacing, x, y, z, box)
        assert np.array_equal(oa, ra) and np.array_equal(os_, rs) and np.array_equal(oi, ri), code
        total += ra.size
    assert total > 2000


def test_aligned_rectangle_slab_is_exact():
    """A rectangular vacancy loop on (100) of BCC, b = a0 [100]: the slab is the
    2 atomic planes behind and in front of the cut whose in-plane positions lie
    inside the rectangle -- counted here from the lattice directly."""
    a0, cells = 2.8553, 14
    x, y, z = G.lattice("BCC", cells, cells, cells, a0)
    c, spacing = G.native_snap_to_mid_plane([0.3, -0.1, 0.2], [1, 0, 0], a0, cells, cells, cells,
                                            "BCC")
    assert spacing == a0 / 2
    hm, hp = 3.3 * a0 / 2, 2.3 * a0 / 2            # half widths between lattice rows
    lp = G.Loop("rect", c, [2, 0, 0], [1, 0, 0], a0, half_m=hm, half_p=hp)
    atoms, side, info = oracle.prismatic_slab(G.native_loop_state(lp), spacing, x, y, z,
                                              [cells * a0] * 3)
    assert info[0] == 1 and info[1] == 2                 # vacancy, 2 planes
    # default axes for m // n = x: u = z, v = y
    d = x - c[0]
    want = np.nonzero((np.abs(d) < spacing) & (np.abs(z - c[2]) <= hm) &
                      (np.abs(y - c[1]) <= hp))[0]
    assert np.array_equal(atoms, want)
    assert np.array_equal(side, np.where(d[want] > 0, 1, -1))
    assert atoms.size >= 12


def test_whole_number_of_planes_is_required():
    a0 = 3.615
    c, spacing = G.native_snap_to_mid_plane([0, 0, 0], [1, 0, 0], a0, 8, 8, 8, "FCC")
    lp = G.Loop("circle", c, [0.5, 0, 0], [1, 0, 0], a0, radius=5.0)  # |b.n| = d / 2
    x, y, z = G.lattice("FCC", 8, 8, 8, a0)
    with pytest.raises(ValueError):
        oracle.prismatic_slab(G.native_loop_state(lp), spacing, x, y, z, [8 * a0] * 3)


def test_host_matter_type_matches_oracle():
    from disco_amd import prismatic as P
    rng = np.random.default_rng(3)
    for lp, *_ in random_prismatic_loops(rng, 30):
        st = G.native_loop_state(lp)
        assert P.matter_type(st) == oracle.matter_type(st)
    c = [0.0, 0.0, 0.0]
    glide = G.native_loop_state(G.Loop("circle", c, [1, -1, 0], [1, 1, 1], 2.8553, radius=5.0))
    assert P.matter_type(glide) == oracle.matter_type(glide) == 0


def hexagon(R, m, n, vdir):
    """Regular hexagon offsets (first vertex along vdir projected on the plane,
    or the default axis), counter-clockwise about n."""
    nh = np.asarray(n, float) / np.linalg.norm(n)
    if vdir:
        u = np.asarray(vdir, float)
        u = u - u.dot(nh) * nh
    else:
        u = np.asarray(G.default_axes(G._unit(m), G._unit(n))[0], float)
    u /= np.linalg.norm(u)
    v = np.cross(u, nh)
    return [list(R * (np.cos(t) * u + np.sin(t) * v)) for t in np.arange(6) * np.pi / 3]


# The slab cases of the reference's own unit test (Code/tests/test_prismatic.cpp:
# 224-245): A0 = 3.3024, 24^3 cells, centre snapped from (1.1, -0.7, 0.4),
# circles and hexagons of R = 18 A (corner radius 0 or 3, vertex axis <1-10> or
# default); (lattice, shape, m, n, corner, vertex axis, expected type)
T3 = 2.0 / 3.0
REF_CASES = [
    ("BCC", "circle", [1, 1, 1], [1, 1, 1], 0, None, 1),
    ("BCC", "circle", [-1, -1, -1], [1, 1, 1], 0, None, 2),
    ("BCC", "circle", [1, -1, 1], [1, -1, 0], 0, None, 1),
    ("BCC", "circle", [-1, 1, -1], [1, -1, 0], 0, None, 2),
    ("BCC", "circle", [0, 0, 2], [0, 0, 1], 0, None, 1),
    ("BCC", "hexagon", [-1, -1, -1], [1, 1, 1], 0.0, [1, -1, 0], 2),
    ("BCC", "hexagon", [1, 1, 1], [1, 1, 1], 3.0, [1, -1, 0], 1),
    ("BCC", "hexagon", [-1, -1, -1], [1, 1, 1], 3.0, None, 2),
    ("FCC", "circle", [1, 1, 0], [1, 1, 0], 0, None, 1),
    ("FCC", "circle", [-1, -1, 0], [1, 1, 0], 0, None, 2),
    ("FCC", "circle", [1, 0, 1], [1, 1, 1], 0, None, 1),
    ("FCC", "circle", [-1, 0, -1], [1, 1, 1], 0, None, 2),
    ("FCC", "circle", [0, 0, 2], [0, 0, 1], 0, None, 1),
    ("FCC", "hexagon", [-1, -1, 0], [1, 1, 0], 3.0, None, 2),
    ("FCC", "circle", [T3, T3, T3], [1, 1, 1], 0, None, 1),          # Frank, vacancy
    ("FCC", "circle", [-T3, -T3, -T3], [1, 1, 1], 0, None, 2),       # Frank, interstitial
]


@pytest.mark.parametrize("case", range(len(REF_CASES)))
def test_oracle_passes_the_reference_unit_test_checks(case):
    """The oracle's slab for each case of Code/tests/test_prismatic.cpp satisfies
    that test's checks: type, |b.n| / d planes, size within 15 % of rho A |b.n|,
    atoms inside the slab bounds on their reported side."""
    kind, shape, m, n, rc, vdir, expect = REF_CASES[case]
    a0, N = 3.3024, 24
    x, y, z = G.lattice(kind, N, N, N, a0)
    c, d = G.native_snap_to_mid_plane([1.1, -0.7, 0.4], n, a0, N, N, N, kind)
    R = 18.0
    if shape == "circle":
        lp = G.Loop("circle", c, m, n, a0, radius=R)
        area = np.pi * R * R
    else:
        lp = G.Loop("polygon", c, m, n, a0, vertices=hexagon(R, m, n, vdir), corner_r=rc,
                    axis=vdir or [])
        area = 1.5 * np.sqrt(3) * R * R - rc * rc * (6 / np.sqrt(3) - np.pi)
    st = G.native_loop_state(lp)
    atoms, side, info = oracle.prismatic_slab(st, d, x, y, z, [N * a0] * 3)
    bn = abs(np.dot(st.bmag * np.array(st.m[:]), np.array(st.n[:])))
    assert int(info[0]) == expect == oracle.matter_type(st)
    assert int(info[1]) == round(bn / d)
    nominal = (2 if kind == "BCC" else 4) / a0 ** 3 * area * bn
    assert abs(atoms.size - nominal) < 0.15 * nominal
    s = (x[atoms] - c[0]) * st.n[0] + (y[atoms] - c[1]) * st.n[1] + (z[atoms] - c[2]) * st.n[2]
    assert np.all(s / d >= -np.floor(info[1] / 2) - 1e-9) and np.all(s / d < np.ceil(info[1] / 2))
    assert np.array_equal(s > 0, side > 0)
