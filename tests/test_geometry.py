"""The workload producer (disco_amd.geometry) against the reference's loop
geometry (golden outputs of Code/loop.cpp and Code/lattice.hpp)."""
import numpy as np
import pytest

import oracle
from conftest import golden
from disco_amd import geometry as G

CASES = [("circle", [30.0], [1, 1, 1], [1, -1, 0], 2.8553, 0.036),
         ("circle", [25.0], [1, -1, 0], [1, 1, 1], 3.615, 0.05),
         ("polygon", [64, 30.0, 0.0], [1, 1, 1], [1, -1, 0], 2.8553, 0.01),
         ("polygon", [6, 44.0, 0.0], [0, 1, -1], [1, 1, 1], 3.615, 0.01),
         ("rect", [12.0, 7.0, 0.0], [1, 1, 1], [1, -1, 0], 2.8553, 0.01),
         ("parallelogram", [9.0, 6.0, 1, 1, 1, 1, 1, -1], [1, 1, 1], [1, -1, 0], 2.8553, 0.01)]


def _make(k):
    shape, p, m, n, a0, eps = CASES[k]
    c = list(golden("geometry.npz")["case_centers"][k])
    if shape == "circle":
        return G.Loop("circle", c, m, n, a0, radius=p[0]), eps
    if shape == "polygon":
        return G.Loop("polygon", c, m, n, a0, vertices=G.regular_polygon(int(p[0]), p[1], m, n)), eps
    if shape == "rect":
        return G.Loop("rect", c, m, n, a0, half_m=p[0], half_p=p[1]), eps
    return G.Loop("parallelogram", c, m, n, a0, half_m=p[0], half_p=p[1], dir1=p[2:5],
                  dir2=p[5:8]), eps


@pytest.mark.parametrize("k", range(len(CASES)))
def test_polyline_matches_reference(k):
    g = golden("geometry.npz")
    lp, eps = _make(k)
    P = lp.polyline(eps)
    ref = g[f"poly{k}"]
    assert P.shape == ref.shape
    assert np.abs(P - ref).max() < 1e-12


@pytest.mark.parametrize("k", range(len(CASES)))
def test_loop_physics_and_loopons_match_reference(k):
    g = golden("geometry.npz")
    lp, _ = _make(k)
    assert np.abs(np.array(lp.phys(1.585)) - g[f"phys{k}"]).max() < 1e-14
    L = lp.loopons(1.585)
    ref = g[f"loopons{k}"]
    assert L.shape == ref.shape
    assert np.abs(L - ref).max() < 1e-12


def test_regular_polygon_matches_reference():
    g = golden("geometry.npz")
    assert np.abs(np.array(G.regular_polygon(64, 30.0, [1, 1, 1], [1, -1, 0])) - g["rpoly64"]).max() < 1e-13
    assert np.abs(np.array(G.regular_polygon(6, 44.0, [0, 1, -1], [1, 1, 1])) - g["rpoly6"]).max() < 1e-13


def test_snap_to_mid_plane_matches_reference():
    g = golden("geometry.npz")
    for k in range(16):
        lat = "FCC" if k % 2 else "BCC"
        n = [[1, 1, 1], [1, -1, 0], [1, 1, -1], [0, 1, 1]][k % 4]
        a0 = 3.615 if lat == "FCC" else 2.8553
        c, sp = G.snap_to_mid_plane(list(g["snap_in"][k]), n, a0, 40, 50, 60, lat)
        assert np.abs(np.array([*c, sp]) - g["snap_out"][k]).max() < 1e-12


def test_lattice_is_bit_identical_to_reference():
    g = golden("geometry.npz")
    for lat, a0 in (("BCC", 3.0), ("FCC", 3.615)):
        x, y, z = G.lattice(lat, 3, 4, 5, a0)
        assert np.array_equal(np.stack([x, y, z]), g[f"lattice_{lat}"])


@pytest.mark.skipif(not oracle.ref_available(), reason="reference build not present")
@pytest.mark.parametrize("lat,a0,cells", [("BCC", 2.8553, 40), ("BCC", 2.8553, 60),
                                          ("FCC", 3.615, 100)])
def test_lattice_bit_identical_to_compiled_reference_at_workload_sizes(lat, a0, cells):
    """The lattices of the 128,000-, 432,000- and 4,000,000-atom workloads against the
    reference's generateLattice (Code/lattice.hpp:35-51) compiled from its sources:
    the device lattice kernel is tested against G.lattice at these sizes."""
    x, y, z = G.lattice(lat, cells, cells, cells, a0)
    rx, ry, rz = oracle.ref_lattice(lat, cells, cells, cells, a0)
    assert np.array_equal(x, rx) and np.array_equal(y, ry) and np.array_equal(z, rz)


def test_c1_loop_centres_match_the_measured_snap():
    """Centres of the C1 loops after the reference's mid-plane snap, as the
    compiled reference places them (snap distance d = 2.019002 A)."""
    s = G.config_c1()
    c = np.array([[l.cx, l.cy, l.cz] for l in s.loops])
    assert np.abs(c - np.array([[-27.839175, -29.266825, -28.553],
                                [29.266825, 27.839175, 28.553]])).max() < 1e-5
    assert s.n == 128000 and all(len(l.vx) == 64 for l in s.loops)


def test_workload_shapes():
    s = G.config_c2()
    assert s.n == 4_000_000 and len(s.loops) == 50 and s.img == 2
    nseg = sorted(len(l.vx) for l in s.loops)
    assert nseg.count(48) == 30 and nseg.count(6) == 20
    for l in s.loops:       # glide loops: b in the plane
        assert abs(np.dot(l.burgers(), l.phys[6:9])) < 1e-12
    c3 = G.config_c3()
    assert c3.n == 432_000 and len(c3.loops) == 20
    assert all(abs(len(l.x) - 2001) <= 8 for l in c3.loops)
    rng = np.random.default_rng(0)
    s4 = G.config_c4_sample(rng)
    assert 1 <= len(s4.loops) <= 16 and all(16 <= len(l.vx) <= 128 for l in s4.loops)
