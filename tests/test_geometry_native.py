"""The C++ loop-geometry producer (include/disco_geometry.h) against the numpy
restatement (bit for bit), the reference's golden geometry, and -- where the
reference build is present -- the compiled reference itself on random loops."""
import numpy as np
import pytest

import oracle
from conftest import golden
from disco_amd import geometry as G
from test_geometry import CASES, _make


@pytest.mark.parametrize("k", range(len(CASES)))
def test_native_equals_restatement_and_golden(k):
    g = golden("geometry.npz")
    lp, eps = _make(k)
    P = G.native_polyline(lp, eps)
    assert np.array_equal(P, lp.polyline(eps))
    assert np.abs(P - g[f"poly{k}"]).max() < 1e-12
    L = G.native_loopons(lp, 1.585)
    assert np.array_equal(L, lp.loopons(1.585))
    assert np.abs(L - g[f"loopons{k}"]).max() < 1e-12
    assert np.array_equal(np.array(G.native_phys(lp, 1.585)), np.array(lp.phys(1.585)))
    assert np.abs(np.array(G.native_phys(lp, 1.585)) - g[f"phys{k}"]).max() < 1e-14


def test_native_polygon_and_snap():
    g = golden("geometry.npz")
    for sides, R, m, n, key in ((64, 30.0, [1, 1, 1], [1, -1, 0], "rpoly64"),
                                (6, 44.0, [0, 1, -1], [1, 1, 1], "rpoly6")):
        P = G.native_regular_polygon(sides, R, m, n)
        assert np.array_equal(P, np.array(G.regular_polygon(sides, R, m, n)))
        assert np.abs(P - g[key]).max() < 1e-13
    for k in range(16):
        lat = "FCC" if k % 2 else "BCC"
        n = [[1, 1, 1], [1, -1, 0], [1, 1, -1], [0, 1, 1]][k % 4]
        a0 = 3.615 if lat == "FCC" else 2.8553
        c, sp = G.native_snap_to_mid_plane(list(g["snap_in"][k]), n, a0, 40, 50, 60, lat)
        c2, sp2 = G.snap_to_mid_plane(list(g["snap_in"][k]), n, a0, 40, 50, 60, lat)
        assert c == c2 and sp == sp2
        assert np.abs(np.array([*c, sp]) - g["snap_out"][k]).max() < 1e-12


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


def test_native_batch_equals_restatement():
    rng = np.random.default_rng(17)
    items = _random_loops(rng, 300)
    loops = [lp for lp, _, _ in items]
    V, off = G.native_polylines(loops, 0.01, [ns for _, ns, _ in items],
                                [sd for _, _, sd in items])
    assert off[-1] == len(V)
    for k, (lp, ns, sd) in enumerate(items):
        want = lp.subdivided_polyline(sd) if sd else lp.polyline(0.01, ns)
        assert np.array_equal(V[off[k]:off[k + 1]], want), k
        assert np.array_equal(G.native_loopons(lp, 1.3), lp.loopons(1.3)), k


@pytest.mark.skipif(not oracle.ref_available(), reason="reference build not present")
def test_native_equals_compiled_reference_on_random_loops():
    rng = np.random.default_rng(23)
    for _ in range(60):
        a0 = 2.8553
        n = G.BCC_110[int(rng.integers(6))]
        m = G._in_plane(G._BCC_B, n)[int(rng.integers(2))]
        c = list(rng.uniform(-40, 40, 3))
        R = float(rng.uniform(5, 45))
        eps = float(rng.uniform(0.005, 0.2))
        lp = G.Loop("circle", c, list(m), list(n), a0, radius=R)
        assert np.abs(G.native_polyline(lp, eps) -
                      oracle.ref_polyline("circle", [R], c, m, n, a0, eps)).max() < 1e-12
        assert np.abs(G.native_loopons(lp, 1.6) -
                      oracle.ref_loopons("circle", [R], 1.6, c, m, n, a0)).max() < 1e-12
        sides = int(rng.integers(3, 10))
        pg = G.Loop("polygon", c, list(m), list(n), a0,
                    vertices=G.regular_polygon(sides, R, m, n))
        assert np.abs(G.native_polyline(pg) -
                      oracle.ref_polyline("polygon", [sides, R, 0.0], c, m, n, a0, 0.01)).max() < 1e-12


def test_native_rejects_bad_input():
    with pytest.raises(ValueError):
        G.native_polyline(G.Loop("circle", [0, 0, 0], [1, 1, 1], [1, -1, 0], 2.9, radius=0.0))
    with pytest.raises(ValueError):
        G.native_regular_polygon(2, 10.0, [1, 1, 1], [1, -1, 0])
    with pytest.raises(ValueError):
        G.native_snap_to_mid_plane([0, 0, 0], [0.5, 1, 0], 2.9, 10, 10, 10, "BCC")
    with pytest.raises(ValueError):
        G.native_snap_to_mid_plane([0, 0, 0], [0, 0, 0], 2.9, 10, 10, 10, "BCC")
