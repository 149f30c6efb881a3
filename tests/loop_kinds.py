"""Loops of each FP32 far-pass form, for the GPU parity tests.

glide: b in the loop plane; prism: b along the normal; mixed: b with both
parts; bent: vertices off the plane by up to 3e-7 R (above the float
resolution 2^-24 R of the in-plane coordinates, below the 1e-6 R planarity
limit), summed in global axes.  Circles of 24-48 segments in
random orientations inside a 150 x 160 x 170 A box."""
import numpy as np

from disco_amd import backend as B

BOX = (150.0, 160.0, 170.0)


def _frame(rng):
    n = rng.normal(size=3)
    n /= np.linalg.norm(n)
    a = np.array([1.0, 0, 0]) if abs(n[0]) < 0.9 else np.array([0, 1.0, 0])
    e1 = np.cross(n, a)
    e1 /= np.linalg.norm(e1)
    return n, e1, np.cross(n, e1)


def kind_loops(kind, rng, count=3):
    loops = []
    for k in range(count):
        n, e1, e2 = _frame(rng)
        c = rng.uniform(30, 120, 3)
        nseg = 24 + 12 * k
        t = 2 * np.pi * np.arange(nseg) / nseg
        R = rng.uniform(10, 35)
        V = c[None] + R * (np.cos(t)[:, None] * e1[None] + np.sin(t)[:, None] * e2[None])
        if kind == "glide":
            b = 2.5 * (0.6 * e1 + 0.8 * e2)
        elif kind == "prism":
            b = 2.5 * n
        elif kind == "mixed":
            b = 1.5 * n + 2.0 * e1 - 0.7 * e2
        else:   # bent
            b = 2.5 * e1
            V = V + 3e-7 * R * np.sin(np.arange(nseg))[:, None] * n[None]
        nb = np.linalg.norm(b)
        loops.append(B.LoopData(phys=(nb, nb, nb, *(b / nb), *n, 0.0), line=True, vx=V[:, 0],
                                vy=V[:, 1], vz=V[:, 2], cx=c[0], cy=c[1], cz=c[2]))
    return loops


def atoms_near(loops, rng, nrand=1500, per_loop=300):
    """Random atoms plus atoms within 10 A of each loop's plane and rim."""
    pts = [rng.uniform(0, 1, (nrand, 3)) * np.array(BOX)]
    for lp in loops:
        c = np.array([lp.cx, lp.cy, lp.cz])
        n = np.array(lp.phys[6:9])
        V = np.stack([lp.vx, lp.vy, lp.vz], 1)
        R = np.linalg.norm(V - c, axis=1).max()
        e1 = (V[0] - c) / np.linalg.norm(V[0] - c)
        e2 = np.cross(n, e1)
        r = rng.uniform(0, R + 10, per_loop)
        ph = rng.uniform(0, 2 * np.pi, per_loop)
        h = rng.uniform(-10, 10, per_loop)
        pts.append(c + r[:, None] * (np.cos(ph)[:, None] * e1 + np.sin(ph)[:, None] * e2)
                   + h[:, None] * n)
    p = np.concatenate(pts)
    return p[:, 0].copy(), p[:, 1].copy(), p[:, 2].copy()
