"""Closed polylines that are not simple flat polygons, for the parity tests.

Code/barnett.hpp:44-107 sums any closed polyline with its cut fanned from the
apex c, so the solid angle of the fan can reach +-4 pi (a cone seen from
inside, a star polygon winding twice about its centre).  Shapes, in a
150 x 160 x 170 A box:

  saddle      circle of radius R with heights 0.3 R cos 2t off its plane
  crown       heights 0.2 R cos 5t
  cone        flat circle 2.5 R above the apex (the fan is a narrow cone)
  pentagram   the star polygon {5/2}, flat, winding number 2 at the centre
  figure8     two flat lobes of opposite sense (no vector area)
  figure8_3d  two lobes in perpendicular planes
  apex_off    flat circle, apex 0.4 R off its plane
"""
import numpy as np

from disco_amd import backend as B

BOX = (150.0, 160.0, 170.0)
SHAPES = ("saddle", "crown", "cone", "pentagram", "figure8", "figure8_3d", "apex_off")


def frame(rng):
    n = rng.normal(size=3)
    n /= np.linalg.norm(n)
    a = np.array([1.0, 0, 0]) if abs(n[0]) < 0.9 else np.array([0, 1.0, 0])
    e1 = np.cross(n, a)
    e1 /= np.linalg.norm(e1)
    return n, e1, np.cross(n, e1)


def shape(name, rng, c=None, R=None, nseg=40):
    """(vertices (nv x 3), apex c, plane normal n, e1, e2, R)."""
    n, e1, e2 = frame(rng)
    c = rng.uniform(50, 100, 3) if c is None else np.asarray(c, float)
    R = float(rng.uniform(10, 16)) if R is None else float(R)
    t = 2 * np.pi * np.arange(nseg) / nseg
    ring = np.cos(t)[:, None] * e1[None] + np.sin(t)[:, None] * e2[None]
    apex = c
    if name == "saddle":
        V = c + R * ring + (0.3 * R * np.cos(2 * t))[:, None] * n
    elif name == "crown":
        V = c + R * ring + (0.2 * R * np.cos(5 * t))[:, None] * n
    elif name == "cone":
        V = c + 2.5 * R * n + R * ring
    elif name == "pentagram":
        a = 4 * np.pi * np.arange(5) / 5
        V = c + R * (np.cos(a)[:, None] * e1 + np.sin(a)[:, None] * e2)
    elif name == "figure8":
        h = nseg // 2
        s = 2 * np.pi * np.arange(h) / h
        lobe1 = (c + R / 2 * e1) + R / 2 * (np.cos(np.pi + s)[:, None] * e1 +
                                             np.sin(np.pi + s)[:, None] * e2)
        lobe2 = (c - R / 2 * e1) + R / 2 * (np.cos(-s)[:, None] * e1 + np.sin(-s)[:, None] * e2)
        V = np.concatenate([lobe1, lobe2])
        apex = c + 0.2 * R * e2
    elif name == "figure8_3d":
        h = nseg // 2
        s = 2 * np.pi * np.arange(h) / h
        lobe1 = (c + R / 2 * e1) + R / 2 * (np.cos(np.pi + s)[:, None] * e1 +
                                             np.sin(np.pi + s)[:, None] * e2)
        lobe2 = (c - R / 2 * e1) + R / 2 * (np.cos(s)[:, None] * e1 + np.sin(s)[:, None] * n)
        V = np.concatenate([lobe1, lobe2])
        apex = c + 0.1 * R * (e2 + n)
    elif name == "apex_off":
        V = c + R * ring
        apex = c + 0.4 * R * n
    else:
        raise ValueError(name)
    return V, apex, n, e1, e2, R


def loop_data(V, apex, b, n):
    bm = float(np.linalg.norm(b))
    return B.LoopData(phys=(bm, bm, bm, *(np.asarray(b) / bm), *n, 0.0), line=True,
                      vx=V[:, 0].copy(), vy=V[:, 1].copy(), vz=V[:, 2].copy(),
                      cx=float(apex[0]), cy=float(apex[1]), cz=float(apex[2]))


def burgers(rng):
    b = rng.normal(size=3)
    return 2.5 * b / np.linalg.norm(b)


def loops(rng, names=SHAPES):
    out = []
    for nm in names:
        V, apex, n, _, _, _ = shape(nm, rng)
        out.append(loop_data(V, apex, burgers(rng), n))
    return out


def fan_points(lp, rng, per_loop=200, dmin=0.01, dmax=4.0):
    """Points near the fan triangles (apex, V[k-1], V[k]), off them by dmin to
    dmax along each triangle's normal, plus points near the vertices."""
    V = np.stack([lp.vx, lp.vy, lp.vz], 1)
    c = np.array([lp.cx, lp.cy, lp.cz])
    A, Bv = np.roll(V, 1, axis=0), V
    k = rng.integers(0, len(V), per_loop)
    u, v = rng.random(per_loop), rng.random(per_loop)
    sw = u + v > 1
    u[sw], v[sw] = 1 - u[sw], 1 - v[sw]
    P = c + u[:, None] * (A[k] - c) + v[:, None] * (Bv[k] - c)
    N = np.cross(A[k] - c, Bv[k] - c)
    N /= np.linalg.norm(N, axis=1)[:, None]
    d = rng.uniform(dmin, dmax, per_loop) * np.where(rng.random(per_loop) < 0.5, 1, -1)
    P = P + d[:, None] * N
    j = rng.integers(0, len(V), per_loop // 2)
    Q = V[j] + rng.normal(size=(per_loop // 2, 3)) * 2.0
    return np.concatenate([P, Q])


def atoms(loop_list, rng, nrand=800, per_loop=200):
    pts = [rng.uniform(0, 1, (nrand, 3)) * np.array(BOX)]
    pts += [fan_points(lp, rng, per_loop) for lp in loop_list]
    p = np.concatenate(pts)
    return p[:, 0].copy(), p[:, 1].copy(), p[:, 2].copy()


def inside_cone(V, apex, n, R, rng, count=200):
    """Points strictly inside the cone (apex, polygon V) and outside it, with a
    flag: inside ones at heights 0.1 .. 0.9 of the cone, within 0.8 of its
    radius there."""
    H = float((V[0] - apex) @ n)
    e1 = np.cross(n, [1.0, 0, 0]) if abs(n[0]) < 0.9 else np.cross(n, [0, 1.0, 0])
    e1 /= np.linalg.norm(e1)
    e2 = np.cross(n, e1)
    z = rng.uniform(0.1, 0.9, count) * H
    rr = rng.uniform(0, 0.8, count) * R * z / H * np.cos(np.pi / len(V))
    ph = rng.uniform(0, 2 * np.pi, count)
    P_in = apex + z[:, None] * n + rr[:, None] * (np.cos(ph)[:, None] * e1 +
                                                 np.sin(ph)[:, None] * e2)
    z = rng.uniform(-0.5, 1.5, count) * H
    rr = (1.2 * R * np.clip(z / H, 0.1, None) + rng.uniform(0.5, 10, count))
    P_out = apex + z[:, None] * n + rr[:, None] * (np.cos(ph)[:, None] * e1 +
                                                   np.sin(ph)[:, None] * e2)
    return P_in, P_out
