"""The oracle on closed polylines that are not simple flat polygons (the
shapes of polylines.py), pinned three ways: against the reference's own
barnettLoop/kroupaCPU compiled from its sources, against a quadrature of the
Kroupa kernel over the fan (each fan triangle with its own normal), and
against what the cut fixes: moving the apex changes u by b times the change
of winding number, and crossing the fan at a point of winding w jumps u by
-w b (Code/barnett.hpp:44-107 sums every fan triangle's atan2, so its solid
angle reaches +-4 pi)."""
import numpy as np
import pytest

import oracle
import polylines as PL
from test_oracle_math import _DB, _DW, _kroupa_field


@pytest.mark.skipif(not oracle.ref_available(), reason="reference build not present")
@pytest.mark.parametrize("name", PL.SHAPES)
def test_oracle_equals_reference_build(name):
    rng = np.random.default_rng(PL.SHAPES.index(name) + 40)
    loops = PL.loops(rng, (name, name))
    x, y, z = PL.atoms(loops, rng, nrand=300, per_loop=150)
    ref = np.stack(oracle.ref_kroupa_cpu_line(x, y, z, loops, PL.BOX)[:3])
    got = np.stack(oracle.line_sum(x, y, z, loops, PL.BOX, 1))
    assert np.abs(got - ref).max() < 1e-12


@pytest.mark.skipif(not oracle.ref_available(), reason="reference build not present")
def test_barnett_loop_equals_reference_inside_the_cone():
    rng = np.random.default_rng(5)
    V, apex, n, _, _, R = PL.shape("cone", rng)
    b = PL.burgers(rng)
    P_in, P_out = PL.inside_cone(V, apex, n, R, rng, 40)
    for p in np.concatenate([P_in, P_out]):
        a = oracle.barnett_loop(p, V, apex, b)
        r = oracle.barnett_loop(p, V, apex, b, use_ref=True)
        assert np.abs(a - r).max() < 1e-13


def test_moving_the_apex_adds_b_inside_the_cone():
    """Cut fanned from the cone's tip or from the centre of its base: the line
    terms are the same, and the two solid angles differ by 4 pi inside the
    solid cone (where the tip's fan subtends more than 2 pi) and by 0 outside."""
    rng = np.random.default_rng(6)
    V, apex, n, _, _, R = PL.shape("cone", rng)
    base = V.mean(0)
    b = PL.burgers(rng)
    P_in, P_out = PL.inside_cone(V, apex, n, R, rng, 60)
    d_in = np.array([oracle.barnett_loop(p, V, apex, b) - oracle.barnett_loop(p, V, base, b)
                     for p in P_in])
    d_out = np.array([oracle.barnett_loop(p, V, apex, b) - oracle.barnett_loop(p, V, base, b)
                      for p in P_out])
    sgn = np.sign(d_in[0] @ b)
    assert np.abs(d_in - sgn * b).max() < 1e-9
    assert np.abs(d_out).max() < 1e-9


@pytest.mark.parametrize("name,w", [("pentagram", 2), ("figure8", 1)])
def test_jump_across_a_flat_cut_is_the_winding_number(name, w):
    """A flat contour's jump across its plane is -w b with w its winding number
    there: 2 at the pentagram's centre (1 in its tips), +-1 in the two lobes
    of a figure eight of opposite senses."""
    rng = np.random.default_rng(7)
    V, apex, n, e1, e2, R = PL.shape(name, rng)
    b = PL.burgers(rng)
    eps = 1e-7

    def jump(p):
        return oracle.barnett_loop(p + eps * n, V, apex, b) - oracle.barnett_loop(p - eps * n, V,
                                                                                  apex, b)
    c = V.mean(0) if name == "pentagram" else apex - 0.2 * R * e2
    if name == "pentagram":
        # centre: winding 2 (the sense of the vertex order sets the sign)
        j0 = jump(c)
        s = -np.sign(j0 @ b)
        assert np.abs(j0 + 2 * s * b).max() < 1e-5
        tip = c + 0.8 * R * e1                       # inside the tip at vertex 0
        assert np.abs(jump(tip) + s * b).max() < 1e-5
        assert np.abs(jump(c + 1.2 * R * e1)).max() < 1e-5
    else:
        l1, l2 = jump(c + 0.5 * R * e1), jump(c - 0.5 * R * e1)
        assert np.abs(l1 + l2).max() < 1e-5          # opposite senses
        assert abs(abs(l1 @ b) - b @ b) < 1e-5 * (b @ b)
        assert np.abs(jump(c + 2.0 * R * e2)).max() < 1e-5


def _fan_quadrature(p, V, c, b, sub=24):
    """Kroupa kernel summed over the fan triangles (c, V[k], V[k+1]), each with
    its own unit normal (V[k] - c) x (V[k+1] - c) (the orientation that the
    flat loops' n = N / |N| takes), Dunavant degree 5 on sub^2 sub-triangles."""
    bm = float(np.linalg.norm(b))
    m = b / bm
    u = np.zeros(3)
    N = len(V)
    for k in range(N):
        a1, a2 = V[k] - c, V[(k + 1) % N] - c
        nk = np.cross(a1, a2)
        area = 0.5 * np.linalg.norm(nk)
        if area == 0.0:
            continue
        nk = nk / (2 * area)
        e1, e2 = a1 / sub, a2 / sub
        tri = []
        for i in range(sub):
            for j in range(sub - i):
                tri.append((c + i * e1 + j * e2, c + (i + 1) * e1 + j * e2, c + i * e1 + (j + 1) * e2))
                if i + j < sub - 1:
                    tri.append((c + (i + 1) * e1 + j * e2, c + (i + 1) * e1 + (j + 1) * e2,
                                c + i * e1 + (j + 1) * e2))
        T = np.array(tri)
        X = np.einsum("gk,tkd->tgd", _DB, T).reshape(-1, 3)
        W = np.tile(_DW * area / sub ** 2, len(tri))
        u += _kroupa_field(p, X, W, (bm, bm, bm, *m, *nk, 0.0))
    return u


@pytest.mark.parametrize("name", ["saddle", "crown", "figure8_3d", "apex_off"])
def test_line_integral_vs_fan_quadrature(name):
    rng = np.random.default_rng(8)
    V, apex, n, e1, e2, R = PL.shape(name, rng, R=8.0, nseg=24)
    b = PL.burgers(rng)
    lp = PL.loop_data(V, apex, b, n)
    pts = PL.fan_points(lp, rng, per_loop=6, dmin=1.5, dmax=3.0)[:6]
    pts = np.concatenate([pts, apex + rng.normal(size=(3, 3)) * 3 * R])
    for p in pts:
        ul = oracle.barnett_loop(p, V, apex, b)
        uq = _fan_quadrature(p, V, apex, b)
        assert np.abs(ul - uq).max() < 2e-6, (name, p, ul, uq)
