"""The C++ loop-geometry producer (include/disco_geometry.h): bit-identical to
the reference's golden geometry and -- where the reference build is present --
to the compiled reference itself on random loops of every shape (rounded
corners, explicit axes, arbitrary convex polygons); within 1e-12 of the numpy
restatement, which rounds every product on its own."""
import numpy as np
import pytest

import oracle
from conftest import golden
from disco_amd import geometry as G
from test_geometry import CASES, _make


def _native_case(k):
    """Case k with the polygon vertices from the producer's regularPolygon (the
    golden file used the reference's)."""
    lp, eps = _make(k)
    shape, p, m, n, a0, _ = CASES[k]
    if shape == "polygon":
        lp = G.Loop("polygon", lp.center, m, n, a0,
                    vertices=G.native_regular_polygon(int(p[0]), p[1], m, n))
    return lp, eps


@pytest.mark.parametrize("k", range(len(CASES)))
def test_native_bit_identical_to_golden(k):
    g = golden("geometry.npz")
    lp, eps = _native_case(k)
    P = G.native_polyline(lp, eps)
    assert np.array_equal(P, g[f"poly{k}"])
    assert np.abs(P - _make(k)[0].polyline(eps)).max() < 1e-12
    L = G.native_loopons(lp, 1.585)
    assert np.array_equal(L, g[f"loopons{k}"])
    assert np.abs(L - _make(k)[0].loopons(1.585)).max() < 1e-12
    assert np.array_equal(np.array(G.native_phys(lp, 1.585)), g[f"phys{k}"])


def test_native_polygon_and_snap():
    g = golden("geometry.npz")
    for sides, R, m, n, key in ((64, 30.0, [1, 1, 1], [1, -1, 0], "rpoly64"),
                                (6, 44.0, [0, 1, -1], [1, 1, 1], "rpoly6")):
        P = G.native_regular_polygon(sides, R, m, n)
        assert np.array_equal(P, g[key])
        assert np.abs(P - np.array(G.regular_polygon(sides, R, m, n))).max() < 1e-13
    for k in range(16):
        lat = "FCC" if k % 2 else "BCC"
        n = [[1, 1, 1], [1, -1, 0], [1, 1, -1], [0, 1, 1]][k % 4]
        a0 = 3.615 if lat == "FCC" else 2.8553
        c, sp = G.native_snap_to_mid_plane(list(g["snap_in"][k]), n, a0, 40, 50, 60, lat)
        assert np.array_equal(np.array([*c, sp]), g["snap_out"][k])
        c2, sp2 = G.snap_to_mid_plane(list(g["snap_in"][k]), n, a0, 40, 50, 60, lat)
        assert np.abs(np.array(c) - np.array(c2)).max() < 1e-12 and sp == sp2


def _random_loops(rng, count):
    """Loops of every shape as the batched workload draws them, plus regular polygons."""
    out = []
    a0 = 2.8553
    for _ in range(count):
        n = G.BCC_110[int(rng.integers(6))]
        m = G._in_plane(G._BCC_B, n)[int(rng.integers(2))]
        c, _ = G.snap_to_mid_plane(list(rng.uniform(-50, 50, 3)), n, a0, 40, 40, 40, "BCC")
        size = float(rng.uniform(8, 40))
        kind = int(rng.integers(4))
        if kind == 0:
            out.append((G.Loop("circle", c, list(m), list(n), a0, radius=size),
                        int(rng.integers(16, 129)), None))
        elif kind == 1:
            out.append((G.Loop("rect", c, list(m), list(n), a0, half_m=size,
                               half_p=size * float(rng.uniform(0.5, 1))), None,
                        int(rng.integers(16, 129))))
        elif kind == 2:
            u, _v = G.default_axes(G._unit(m), G._unit(n))
            w = G._cross(G._unit(n), u)
            ang = float(rng.uniform(0.35, 1.2))
            d2 = [np.cos(ang) * u[i] + np.sin(ang) * w[i] for i in range(3)]
            out.append((G.Loop("parallelogram", c, list(m), list(n), a0, half_m=size,
                               half_p=size * 0.7, dir1=u, dir2=d2), None,
                        int(rng.integers(16, 129))))
        else:
            sides = int(rng.integers(3, 13))
            out.append((G.Loop("polygon", c, list(m), list(n), a0,
                               vertices=G.regular_polygon(sides, size, m, n)), None, None))
    return out


def test_native_batch_equals_single_calls():
    rng = np.random.default_rng(17)
    items = _random_loops(rng, 300)
    loops = [lp for lp, _, _ in items]
    V, off = G.native_polylines(loops, 0.01, [ns for _, ns, _ in items],
                                [sd for _, _, sd in items])
    assert off[-1] == len(V)
    for k, (lp, ns, sd) in enumerate(items):
        assert np.array_equal(V[off[k]:off[k + 1]], G.native_polyline(lp, 0.01, ns, sd)), k
        want = lp.subdivided_polyline(sd) if sd else lp.polyline(0.01, ns)
        assert np.abs(V[off[k]:off[k + 1]] - want).max() < 1e-12, k
        assert np.abs(G.native_loopons(lp, 1.3) - lp.loopons(1.3)).max() < 1e-12, k


def _ref_random_shapes(rng, count):
    """(Loop, reference shape code, reference params) of every shape the
    reference builds, with rounded corners and explicit axes."""
    out = []
    a0 = 2.8553
    for t in range(count):
        n = [int(v) for v in G.BCC_110[int(rng.integers(6))]]
        m = [int(v) for v in G._in_plane(G._BCC_B, n)[int(rng.integers(2))]]
        if t % 3 == 2:                       # prismatic / mixed: m off the plane
            m = [1, 1, 1] if np.dot([1, 1, 1], n) else [1, -1, 1]
        c = list(rng.uniform(-40, 40, 3))
        R = float(rng.uniform(5, 45))
        kind = t % 6
        u, _ = G.default_axes(G._unit(m), G._unit(n))
        w = G._cross(G._unit(n), u)
        if kind == 0:
            out.append((G.Loop("circle", c, m, n, a0, radius=R), "circle", [R]))
        elif kind == 1:
            hm, hp, rc = R, R * float(rng.uniform(0.4, 1.0)), float(rng.uniform(0, 0.6)) * R
            side = [0.0, 0.0, 0.0]
            if t % 4 == 1:
                ang = float(rng.uniform(0, 3))
                side = [np.cos(ang) * u[i] + np.sin(ang) * w[i] for i in range(3)]
            out.append((G.Loop("rect", c, m, n, a0, half_m=hm, half_p=hp, corner_r=rc,
                               axis=side if any(side) else []),
                        "rect", [hm, hp, rc, *side]))
        elif kind == 2:
            ang = float(rng.uniform(0.35, 1.2))
            d2 = [np.cos(ang) * u[i] + np.sin(ang) * w[i] for i in range(3)]
            out.append((G.Loop("parallelogram", c, m, n, a0, half_m=R, half_p=0.7 * R, dir1=u,
                               dir2=d2), "parallelogram", [R, 0.7 * R, *u, *d2]))
        elif kind == 3:
            sides, rc = int(rng.integers(3, 13)), float(rng.uniform(0, 0.5)) * R
            out.append((G.Loop("polygon", c, m, n, a0,
                               vertices=G.native_regular_polygon(sides, R, m, n), corner_r=rc),
                        "polygon", [sides, R, rc]))
        else:
            # an irregular convex polygon (random angles on an ellipse), any axis
            nv = int(rng.integers(3, 9))
            ang = np.sort(rng.uniform(0, 2 * np.pi, nv))
            if np.diff(np.concatenate([ang, [ang[0] + 2 * np.pi]])).min() < 0.2:
                ang = np.linspace(0, 2 * np.pi, nv, endpoint=False) + ang[0]
            V = [[R * np.cos(a) * u[i] + 0.6 * R * np.sin(a) * w[i] for i in range(3)] for a in ang]
            axis = [0.0, 0.0, 0.0] if kind == 4 else list(rng.normal(0, 1, 3))
            rc = float(rng.uniform(0, 0.3)) * R if kind == 5 else 0.0
            out.append((G.Loop("polygon", c, m, n, a0, vertices=V, corner_r=rc,
                               axis=axis if any(axis) else []),
                        "polygon_vertices", [nv, rc, *axis, *np.ravel(V)]))
    return out


@pytest.mark.skipif(not oracle.ref_available(), reason="reference build not present")
def test_native_bit_identical_to_compiled_reference():
    rng = np.random.default_rng(23)
    a0 = 2.8553
    for lp, code, p in _ref_random_shapes(rng, 120):
        c, m, n = lp.center, lp.m, lp.n
        for eps in (0.01, 0.2):
            assert np.array_equal(G.native_polyline(lp, eps),
                                  oracle.ref_polyline(code, p, c, m, n, a0, eps)), (code, p)
        assert np.array_equal(G.native_loopons(lp, 1.6),
                              oracle.ref_loopons(code, p, 1.6, c, m, n, a0)), (code, p)
        phys, _ = oracle.ref_loop_phys(code, p, 1.6, c, m, n, a0)
        assert np.array_equal(np.array(G.native_phys(lp, 1.6)), phys), (code, p)
    for k in range(40):
        m = G._in_plane(G._BCC_B, G.BCC_110[k % 6])[k % 2]
        sides = 3 + k % 14
        assert np.array_equal(G.native_regular_polygon(sides, 10.0 + k, m, G.BCC_110[k % 6]),
                              oracle.ref_regular_polygon(sides, 10.0 + k, m, G.BCC_110[k % 6]))
        lat = "FCC" if k % 2 else "BCC"
        ctr = list(rng.uniform(-60, 60, 3))
        c1, s1 = G.native_snap_to_mid_plane(ctr, G.BCC_110[k % 6], a0, 40, 50, 60, lat)
        c2, s2 = oracle.ref_snap(ctr, G.BCC_110[k % 6], a0, 40, 50, 60, lat)
        assert np.array_equal(np.array(c1), c2) and s1 == s2


def test_native_rejects_bad_input():
    with pytest.raises(ValueError):
        G.native_polyline(G.Loop("circle", [0, 0, 0], [1, 1, 1], [1, -1, 0], 2.9, radius=0.0))
    with pytest.raises(ValueError):
        G.native_regular_polygon(2, 10.0, [1, 1, 1], [1, -1, 0])
    with pytest.raises(ValueError):
        G.native_snap_to_mid_plane([0, 0, 0], [0.5, 1, 0], 2.9, 10, 10, 10, "BCC")
    with pytest.raises(ValueError):
        G.native_snap_to_mid_plane([0, 0, 0], [0, 0, 0], 2.9, 10, 10, 10, "BCC")


@pytest.mark.parametrize("dims", [(12, 12, 12), (11, 13, 9)])
@pytest.mark.parametrize("kind", ["BCC", "FCC"])
def test_snap_reference_unit_test(dims, kind):
    """Code/tests/test_snap.cpp on the producer: the snapped centre moves along n
    only, the reported spacing is the distance between the lattice's planes
    around it (brute force over the sites), and it sits on their mid-plane."""
    a0, c = 3.3024, [1.3, -2.7, 0.9]
    x, y, z = G.lattice(kind, *dims, a0)
    for n in ([1, 1, 0], [1, -1, 2], [1, 2, -3], [1, 0, 0], [2, 2, 0], [1, 1, 1], [1, -1, 3]):
        s, d = G.native_snap_to_mid_plane(c, n, a0, *dims, kind)
        nh = np.asarray(n, float) / np.linalg.norm(n)
        ps, pc = float(np.dot(s, nh)), float(np.dot(c, nh))
        assert np.abs((np.array(s) - c) - (ps - pc) * nh).max() < 1e-9
        p = x * nh[0] + y * nh[1] + z * nh[2]
        below, above = p[p <= ps].max(), p[p > ps].min()
        assert abs((above - below) - d) < 1e-9
        assert abs((ps - below) - d / 2) < 1e-9


def test_snap_centre_on_a_plane_goes_to_the_plus_side():
    s, d = G.native_snap_to_mid_plane([0.0, 0.0, 0.0], [1, 1, 0], 3.3024, 12, 12, 12, "BCC")
    assert abs((s[0] + s[1]) / np.sqrt(2.0) - d / 2) < 1e-9
