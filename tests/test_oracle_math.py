"""Pins of the FP64 oracle by what the mathematics fixes, independent of the
reference's code: closed forms, a brute-force quadrature, invariances and the
image-sum decomposition.  Each is chosen so that a plausible slip in the
oracle (a dropped term, a sign, a swapped operand) fails at least one."""
import math

import numpy as np
import pytest

import oracle
from disco_amd import geometry as G

NU = 0.35
K1 = 1.0 / (8.0 * math.pi * (1.0 - NU))


def _circle_polygon(R, N, c=(0.0, 0.0, 0.0)):
    """N-gon in z = 0 with the disc's area (area-preserving radius)."""
    Rp = R * math.sqrt(2 * math.pi / (N * math.sin(2 * math.pi / N)))
    t = 2 * math.pi * np.arange(N) / N
    return np.stack([c[0] + Rp * np.cos(t), c[1] + Rp * np.sin(t), c[2] + 0 * t], 1)


# ------------------------------------------------------------------ closed forms
def _prismatic_axis_uz(b, R, h):
    """u_z on the axis of a prismatic circular loop (b along n = z), from the
    Kroupa surface integral done in closed form."""
    s, rr = math.copysign(1.0, h), math.sqrt(R * R + h * h)
    return -(b / (4 * (1 - NU))) * ((1 - 2 * NU) * (s - h / rr) + (s - h ** 3 / rr ** 3))


def _glide_axis_ux(b, R, h):
    """u_x on the axis of a glide circular loop (b along x, n = z)."""
    s, rr = math.copysign(1.0, h), math.sqrt(R * R + h * h)
    return -K1 * b * math.pi * (2 * (1 - 2 * NU) * (s - h / rr) + 2 * s - 3 * h / rr
                                + h ** 3 / rr ** 3)


@pytest.mark.parametrize("h", [-7.0, -1.0, 0.5, 2.0, 15.0])
def test_prismatic_loop_on_axis_closed_form(h):
    b, R = 2.5, 12.0
    errs = []
    for N in (256, 1024, 4096):
        u = oracle.barnett_loop([0.0, 0.0, h], _circle_polygon(R, N), [0, 0, 0], [0, 0, b])
        errs.append(abs(u[2] - _prismatic_axis_uz(b, R, h)))
        assert abs(u[0]) < 1e-12 and abs(u[1]) < 1e-12
    assert errs[-1] < 1e-6 * b
    assert errs[2] < errs[1] < errs[0] or errs[0] < 1e-12       # converges as N grows


@pytest.mark.parametrize("h", [-4.0, -1.0, 1.0, 3.0, 20.0])
def test_glide_loop_on_axis_closed_form(h):
    b, R = 2.5, 10.0
    errs = []
    for N in (256, 1024, 4096):
        u = oracle.barnett_loop([0.0, 0.0, h], _circle_polygon(R, N), [0, 0, 0], [b, 0, 0])
        errs.append(abs(u[0] - _glide_axis_ux(b, R, h)))
        assert abs(u[1]) < 1e-12 and abs(u[2]) < 1e-12
    assert errs[-1] < 1e-6 * b
    assert errs[2] < errs[1] < errs[0] or errs[0] < 1e-12


# ------------------------------------------------------------------ quadrature
# Dunavant's 7-point degree-5 rule on a triangle (barycentric points, weights).
_DW = np.array([0.225, 0.1323941527, 0.1323941527, 0.1323941527,
                0.1259391805, 0.1259391805, 0.1259391805])
_a1, _b1, _a2, _b2 = 0.0597158717, 0.4701420641, 0.7974269853, 0.1012865073
_DB = np.array([[1 / 3, 1 / 3, 1 / 3], [_a1, _b1, _b1], [_b1, _a1, _b1], [_b1, _b1, _a1],
                [_a2, _b2, _b2], [_b2, _a2, _b2], [_b2, _b2, _a2]])


def _kroupa_field(p, X, W, phys):
    """Sum over surface points X (weights W = dS) of the Kroupa term (numpy);
    checked against the oracle's kroupa_pair below."""
    b, m, n = np.asarray(phys[0:3]), np.asarray(phys[3:6]), np.asarray(phys[6:9])
    r = p[None, :] - X
    rho = np.linalg.norm(r, axis=1)
    mr, nr = r @ m, r @ n
    k2 = (1 - 2 * NU) / rho
    fac = -K1 * W / rho ** 2
    tail = 3 * mr * nr / rho ** 3
    inner = k2[:, None] * (n[None, :] * mr[:, None] + m[None, :] * nr[:, None]
                           - r * float(m @ n)) + tail[:, None] * r
    return b * (fac[:, None] * inner).sum(0)


def _quad_points(V, c, sub):
    """Quadrature points and weights over the fan triangles (c, V_k, V_k+1)."""
    pts, wts = [], []
    N = len(V)
    for k in range(N):
        e1, e2 = (V[k] - c) / sub, (V[(k + 1) % N] - c) / sub
        area = 0.5 * np.linalg.norm(np.cross(e1, e2))
        tri = []
        for i in range(sub):
            for j in range(sub - i):
                tri.append((c + i * e1 + j * e2, c + (i + 1) * e1 + j * e2, c + i * e1 + (j + 1) * e2))
                if i + j < sub - 1:
                    tri.append((c + (i + 1) * e1 + j * e2, c + (i + 1) * e1 + (j + 1) * e2,
                                c + i * e1 + (j + 1) * e2))
        T = np.array(tri)                                   # (ntri, 3, 3)
        pts.append(np.einsum("gk,tkd->tgd", _DB, T).reshape(-1, 3))
        wts.append(np.tile(_DW * area, len(tri)))
    return np.concatenate(pts), np.concatenate(wts)


def test_numpy_kroupa_integrand_matches_oracle_pair():
    rng = np.random.default_rng(3)
    for _ in range(20):
        phys = [2.0, 2.0, 2.0, *(v := rng.normal(size=3)) / np.linalg.norm(v),
                *(w := rng.normal(size=3)) / np.linalg.norm(w), 0.7]
        p, x = rng.normal(size=3) * 5, rng.normal(size=(1, 3)) * 5
        got = _kroupa_field(p, x, np.array([phys[9]]), phys)
        assert np.abs(got - oracle.kroupa_pair(p - x[0], phys)).max() < 1e-13


@pytest.mark.parametrize("shape", ["circle-glide", "hexagon-prismatic", "rect-mixed"])
def test_line_integral_vs_quadrature(shape):
    """Line integral == Dunavant quadrature of the Kroupa kernel over the cut
    (the reference's own check, Code/tests/test_barnett.cpp:51-146)."""
    a0 = 3.3024
    c = [0.3, -0.2, 0.1]
    if shape == "circle-glide":
        lp = G.Loop("circle", c, [1, -1, 1], [1, 1, 0], a0, radius=8.0)
    elif shape == "hexagon-prismatic":
        lp = G.Loop("polygon", c, [1, 1, 1], [1, 1, 1], a0,
                    vertices=G.regular_polygon(6, 8.0, [1, 1, 1], [1, 1, 1]))
    else:
        lp = G.Loop("rect", c, [1, 1, 1], [1, -1, 0], a0, half_m=9.0, half_p=6.0)
    V = lp.polyline(0.01)
    phys = lp.phys()
    X, W = _quad_points(V, np.array(c), 40)
    rng = np.random.default_rng(7)
    nh = np.array(lp.nh)
    e1 = np.array(lp.u)
    e2 = np.cross(nh, e1)
    for _ in range(5):
        a, b = (2 * rng.random(2) - 1) * 10.0
        h = (1 + 2 * rng.random()) * (1 if rng.random() < 0.5 else -1)
        p = np.array(c) + a * e1 + b * e2 + h * nh
        ul = oracle.barnett_loop(p, V, c, lp.burgers())
        uq = _kroupa_field(p, X, W, phys)
        assert np.abs(ul - uq).max() < 1e-6, (ul, uq)


# ------------------------------------------------------------------ invariants
def _loop(shape="circle"):
    a0, c = 2.8553, [1.0, -2.0, 0.5]
    if shape == "circle":
        return G.Loop("circle", c, [1, 1, 1], [1, -1, 0], a0, radius=12.0)
    return G.Loop("rect", c, [1, 1, 1], [1, -1, 0], a0, half_m=9.0, half_p=5.0)


def test_jump_across_cut_is_minus_b_inside_zero_outside():
    lp = _loop()
    V, c, b, n = lp.polyline(0.01), np.array(lp.center), np.array(lp.burgers()), np.array(lp.nh)
    h = 1e-7
    up = oracle.barnett_loop(c + h * n, V, c, b)
    um = oracle.barnett_loop(c - h * n, V, c, b)
    assert np.abs(up - um + b).max() < 1e-5
    out = c + 3 * 12.0 * np.array(lp.u)
    up = oracle.barnett_loop(out + h * n, V, c, b)
    um = oracle.barnett_loop(out - h * n, V, c, b)
    assert np.abs(up - um).max() < 1e-5


def test_apex_and_start_vertex_invariance():
    lp = _loop()
    V, c, b = lp.polyline(0.01), np.array(lp.center), np.array(lp.burgers())
    Vr = np.roll(V, -len(V) // 3, axis=0)
    c2 = c + 3.0 * np.array(lp.u) - 2.0 * np.array(lp.v)
    rng = np.random.default_rng(11)
    for _ in range(8):
        p = c + rng.normal(size=3) * 8.0
        assert np.abs(oracle.barnett_loop(p, V, c, b) - oracle.barnett_loop(p, Vr, c2, b)).max() < 1e-11


def test_collinear_subdivision_is_exact():
    lp = _loop("rect")
    V, c, b = lp.polyline(), np.array(lp.center), np.array(lp.burgers())
    Vs = lp.subdivided_polyline(37)
    assert len(Vs) == 37
    rng = np.random.default_rng(12)
    for _ in range(8):
        p = c + rng.normal(size=3) * 10.0
        assert np.abs(oracle.barnett_loop(p, V, c, b) - oracle.barnett_loop(p, Vs, c, b)).max() < 1e-12


def test_linear_in_burgers_vector():
    lp = _loop()
    V, c = lp.polyline(0.01), np.array(lp.center)
    p = c + np.array([3.0, -4.0, 2.5])
    b1, b2 = np.array([1.0, 0.2, -0.5]), np.array([-0.3, 2.0, 0.7])
    u = oracle.barnett_loop(p, V, c, 2 * b1 - 3 * b2)
    assert np.abs(u - (2 * oracle.barnett_loop(p, V, c, b1) - 3 * oracle.barnett_loop(p, V, c, b2))).max() < 1e-13


def test_far_field_matches_one_loopon():
    """A 2 A loop seen from 105 A is one loopon of the same area
    (Code/tests/test_barnett.cpp:228-245)."""
    lp = G.Loop("circle", [0, 0, 0], [1, -1, 1], [1, 1, 0], 3.3024, radius=2.0)
    V = lp.polyline(0.001)
    phys = list(lp.phys())
    phys[9] = math.pi * 4.0
    p = np.array([60.0, -35.0, 80.0])
    ul = oracle.barnett_loop(p, V, [0, 0, 0], lp.burgers())
    uk = oracle.kroupa_pair(p, phys)
    assert np.abs(ul - uk).max() / np.abs(uk).max() < 1e-3


def test_loopon_sum_converges_to_line_integral():
    """Code/tests/test_barnett.cpp:193-226: prismatic hexagon R = 10."""
    c = [0.2, -0.1, 0.3]
    lp = G.Loop("polygon", c, [-1, -1, -1], [1, 1, 1], 3.3024,
                vertices=G.regular_polygon(6, 10.0, [-1, -1, -1], [1, 1, 1]))
    V = lp.polyline()
    nh = np.array(lp.nh)
    rng = np.random.default_rng(2024)
    pts = []
    for _ in range(8):
        p = np.array(c) + rng.uniform(-13, 13, 3)
        s = (p - c) @ nh
        if abs(s) < 1.5:
            p = p + ((1.5 - s) if s > 0 else (-1.5 - s)) * nh
        pts.append(p)
    pts = np.array(pts)
    ul = np.array([oracle.barnett_loop(p, V, c, lp.burgers()) for p in pts])
    errs = []
    for da in (0.8, 0.4, 0.2, 0.1):
        ld = lp.loopon_data(da)
        uk = np.stack(oracle.loopon_sum(pts[:, 0], pts[:, 1], pts[:, 2], [ld], (1e4,) * 3, 0), 1)
        errs.append(np.abs(ul - uk).max())
    assert all(b < a for a, b in zip(errs, errs[1:])), errs     # monotone
    assert errs[-1] < errs[0] / 4                                # refinement pays
    assert errs[-1] < 5e-3, errs


def test_kroupa_pair_is_odd_and_skips_coincident():
    rng = np.random.default_rng(4)
    phys = [2.0, 2.0, 2.0, 0.6, 0.8, 0.0, 0.0, 0.6, 0.8, 1.3]
    for _ in range(10):
        r = rng.normal(size=3) * 4
        assert np.abs(oracle.kroupa_pair(r, phys) + oracle.kroupa_pair(-r, phys)).max() < 1e-15
    assert np.all(oracle.kroupa_pair([1e-11, 0, 0], phys) == 0.0)
    assert np.all(oracle.kroupa_pair([0, 0, 0], phys) == 0.0)


# ------------------------------------------------------------------ image sums
@pytest.mark.parametrize("img", [1, 2])
def test_image_sum_equals_sum_of_single_images(img):
    """+-img sum == sum over delta of the single-image field of the atom moved
    by -delta (the reference's image convention, Code/backend_cpu.cpp:41-56)."""
    s = G.config_c1()
    idx = np.arange(0, s.n, 5003)
    x, y, z = s.x[idx], s.y[idx], s.z[idx]
    Lx, Ly, Lz = s.box
    tot = np.stack(oracle.line_sum(x, y, z, s.loops, s.box, img))
    acc = np.zeros_like(tot)
    for i in range(-img, img + 1):
        for j in range(-img, img + 1):
            for k in range(-img, img + 1):
                acc += np.stack(oracle.line_sum(x - i * Lx, y - j * Ly, z - k * Lz, s.loops,
                                                s.box, 0))
    assert np.abs(tot - acc).max() < 1e-12
    # loopons the same way
    c3 = G.config_c3(nloops=1, da=4.0)
    xi, yi, zi = c3.x[:64], c3.y[:64], c3.z[:64]
    tot = np.stack(oracle.loopon_sum(xi, yi, zi, c3.loops, c3.box, img))
    acc = np.zeros_like(tot)
    L3 = c3.box
    for i in range(-img, img + 1):
        for j in range(-img, img + 1):
            for k in range(-img, img + 1):
                acc += np.stack(oracle.loopon_sum(xi - i * L3[0], yi - j * L3[1], zi - k * L3[2],
                                                  c3.loops, c3.box, 0))
    assert np.abs(tot - acc).max() < 1e-12


def test_line_sum_is_additive_over_loops():
    s = G.config_c1()
    idx = np.arange(0, s.n, 4001)
    x, y, z = s.x[idx], s.y[idx], s.z[idx]
    both = np.stack(oracle.line_sum(x, y, z, s.loops, s.box, 1))
    one = np.stack(oracle.line_sum(x, y, z, s.loops[:1], s.box, 1))
    two = np.stack(oracle.line_sum(x, y, z, s.loops[1:], s.box, 1))
    assert np.abs(both - one - two).max() < 1e-13
