"""Host side of the cubature form: the 9-point rule the library builds for a
flat cut (weights, ring radius, eligibility) and the selection of far images,
checked against closed forms and an independent numpy construction."""
import numpy as np
import pytest

from disco_amd import backend as B
from disco_amd import cubature as C
from disco_amd import geometry as G
from disco_amd import surrogate as S


def _polygon(nv, R, rot=0.3, c=(0.0, 0.0, 0.0), n=(0.0, 0.0, 1.0), e1=(1.0, 0.0, 0.0)):
    n, e1 = np.array(n, float), np.array(e1, float)
    e2 = np.cross(n, e1)
    th = rot + 2 * np.pi * np.arange(nv) / nv
    V = np.array(c) + R * (np.cos(th)[:, None] * e1 + np.sin(th)[:, None] * e2)
    b = 2.5 * e1
    return B.LoopData(phys=(2.5, 2.5, 2.5, *e1, *n, 0.0), line=True, vx=V[:, 0].copy(),
                      vy=V[:, 1].copy(), vz=V[:, 2].copy(), cx=c[0], cy=c[1], cz=c[2])


def _rect(a, b):
    V = np.array([[a, b, 0], [-a, b, 0], [-a, -b, 0], [a, -b, 0]], float)
    return B.LoopData(phys=(2.5, 2.5, 2.5, 1, 0, 0, 0, 0, 1, 0.0), line=True, vx=V[:, 0].copy(),
                      vy=V[:, 1].copy(), vz=V[:, 2].copy())


def test_disc_limit_is_radons_rule():
    """A fine regular polygon tends to the disc: rho^2 = 2R^2/3, w0 = A/4, w1 = 3A/32
    (I2 = pi R^4 / 2, I4 = pi R^6 / 3)."""
    R = 17.0
    c = B.line_cubature([_polygon(4096, R)])[0]
    A = 0.5 * 4096 * R * R * np.sin(2 * np.pi / 4096)         # the polygon's area
    assert c[0] == 1.0
    assert abs(c[1] - A / 4) < 1e-4 * A and abs(c[2] - 3 * A / 32) < 1e-4 * A
    assert abs(c[3] - np.sqrt(2.0 / 3.0) * R) < 1e-4 * R
    assert abs(c[4] - R) < 1e-5 * R


@pytest.mark.parametrize("nv,ok", [(3, 0), (4, 0), (5, 0), (6, 1), (7, 1), (8, 1), (48, 1)])
def test_eligible_only_where_moments_match(nv, ok):
    """Regular n-gons have moments isotropic through rank n - 1: the 9-point rule
    (isotropic through rank 7) matches every moment of degree <= 5 only for n >= 6."""
    c = B.line_cubature([_polygon(nv, 10.0, rot=0.7)])[0]
    assert c[0] == ok


def test_anisotropic_cuts_stay_exact():
    c = B.line_cubature([_rect(10, 5), _rect(10, 10), _polygon(6, 10.0)])
    assert list(c[:, 0]) == [0.0, 0.0, 1.0]


def test_rule_matches_the_numpy_construction():
    """Weights and ring radius equal surrogate.rule7_loopons_many with a ring of
    eight (moments by numpy fan triangles with Dunavant's rule) on the 4M-atom
    workload's loops, in any orientation."""
    s = G.config_c2(seed=5, nloops=20)
    c = B.line_cubature(s.loops)
    r = S.rule7_loopons_many(s.loops, ring=8)
    for k, lp in enumerate(s.loops):
        cen, ring = r[2 * k], r[2 * k + 1]
        rho = np.linalg.norm([ring.x[0] - lp.cx, ring.y[0] - lp.cy, ring.z[0] - lp.cz])
        assert c[k, 0] == 1.0
        assert abs(c[k, 1] - cen.phys[9]) < 1e-6 * cen.phys[9]
        assert abs(c[k, 2] - ring.phys[9]) < 1e-6 * ring.phys[9]
        assert abs(c[k, 3] - rho) < 1e-6 * rho


def test_far_image_selection():
    """An image goes to the grid when its centre lies at least max(k R, R + dmin,
    (GRID_ERR |b| R^2 h^p / tol)^(1/(p+2))) from the atom box (p-node stencils);
    bit 9 (i + 1) + 3 (j + 1) + (k + 1)."""
    lp = _polygon(48, 10.0, c=(40.0, 0.0, 0.0))
    box = (100.0, 100.0, 100.0)
    lo, hi = (-50.0, -50.0, -50.0), (50.0, 50.0, 50.0)
    h, tol, k = 2.0, 1e-7, 6.0
    masks, far = C.far_images([lp], box, lo, hi, h, tol, k, dmin=30.0, order=4)
    got = {tuple(int(round(t / 100.0)) for t in off) for _, off in far}
    R = C.B.line_cubature([lp])[0, 4]
    dmin = max(k * R, R + 30.0, (C.GRID_ERR[4] * 2.5 * R * R * h ** 4 / tol) ** (1 / 6))
    want = set()
    for i in (-1, 0, 1):
        for j in (-1, 0, 1):
            for kk in (-1, 0, 1):
                if i == j == kk == 0:
                    continue
                ci = np.array([40.0 + 100 * i, 100.0 * j, 100.0 * kk])
                gap = np.maximum(np.maximum(np.array(lo) - ci, ci - np.array(hi)), 0)
                if np.linalg.norm(gap) >= dmin:
                    want.add((i, j, kk))
    assert got == want and (1, 0, 0) in got and (-1, 0, 0) not in got
    bits = {b for b in range(27) if (int(masks[0]) >> b) & 1}
    assert bits == {9 * (i + 1) + 3 * (j + 1) + (kk + 1) for i, j, kk in want}
    # a coarser grid (larger interpolation error) takes fewer images
    m_coarse, f_coarse = C.far_images([lp], box, lo, hi, 20.0, tol, k, dmin=30.0, order=4)
    assert len(f_coarse) < len(far)
    # on a grid of spacing 5 a 6-node stencil errs less than the 4-node one: more images
    m4, f4 = C.far_images([lp], box, lo, hi, 5.0, tol, k, dmin=30.0, order=4)
    m6, f6 = C.far_images([lp], box, lo, hi, 5.0, tol, k, dmin=30.0, order=6)
    assert len(f6) > len(f4)
    # loops the rule does not represent are never taken out
    m2, f2 = C.far_images([_rect(5, 3)], box, lo, hi, h, tol, k)
    assert int(m2[0]) == 0 and not f2


# ------------------------------------------------------------ cuts that are not flat discs
def _far_points(c, R, rng, dist, count=40):
    d = rng.normal(size=(count, 3))
    d /= np.linalg.norm(d, axis=1)[:, None]
    return c + dist * R * d


@pytest.mark.parametrize("name", ["saddle", "crown", "cone", "figure8", "figure8_3d", "apex_off"])
def test_fan_loopons_carry_the_far_field_of_any_polyline(name):
    """Radon's rule on every fan triangle, each with its own normal, against the
    oracle's line integral (Code/barnett.hpp:44-107) of the same polyline: the
    error falls as (R/r)^6 of |b| or faster (degree 5 on each triangle), and
    the centred ring rule would not stand for these cuts (ring_rule_ok)."""
    import oracle
    import polylines as PL
    rng = np.random.default_rng(PL.SHAPES.index(name) + 80)
    V, apex, n, _, _, R = PL.shape(name, rng, nseg=24)
    b = PL.burgers(rng)
    lp = PL.loop_data(V, apex, b, n)
    assert not S.ring_rule_ok(lp)
    fan = S.fan_loopons(lp)
    Rf = float(np.linalg.norm(V - apex, axis=1).max())
    box = (1e6, 1e6, 1e6)
    for dist in (6.0, 12.0):
        P = _far_points(V.mean(0), Rf, rng, dist)
        x, y, z = (P[:, k].copy() for k in range(3))
        ul = np.stack(oracle.line_sum(x, y, z, [lp], box, 0))
        uf = np.stack(oracle.loopon_sum(x, y, z, fan, box, 0))
        err = np.abs(ul - uf).max()
        assert err < 0.05 * np.linalg.norm(b) * dist ** -6, (dist, err)


def test_ring_rule_kept_for_flat_cuts_about_the_apex():
    """Every DISCO loop (flat, star-shaped about its centre) keeps the ring rule;
    per-loop lists hold [centre, ring] for those and the fan rule for the rest."""
    import polylines as PL
    s = G.config_c2(seed=5, nloops=10)
    assert all(S.ring_rule_ok(lp) for lp in s.loops)
    rng = np.random.default_rng(3)
    odd = PL.loops(rng, ("saddle", "figure8"))
    per = S.rule_loopons_per_loop(s.loops[:3] + odd, ring=8)
    assert [len(p) for p in per[:3]] == [2, 2, 2]
    assert len(per[3]) == 3 * 40 and len(per[4]) == 3 * 40
    flat = S.rule7_loopons_many(s.loops[:3], ring=8)
    for k in range(3):
        assert per[k][0].phys == flat[2 * k].phys and np.array_equal(per[k][1].x, flat[2 * k + 1].x)
