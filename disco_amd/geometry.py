"""Host-side producer of the backend's inputs for the synthetic workloads.

This restates only what the workloads need of the reference's geometry, in
numpy, with the reference's operation order so that the inputs match it:

* ``lattice``            -- ``generateLattice`` (Code/lattice.hpp:35-51):
                            cells (i, j, k) with k fastest, then the basis sites
* ``plane_spacing``      -- ``latticePlaneSpacing`` (Code/lattice.hpp:146-153)
* ``snap_to_mid_plane``  -- ``snapToMidPlane`` (Code/loop.cpp:342-374)
* ``default_axes``       -- ``defaultAxes`` (Code/loop.cpp:40-57)
* ``regular_polygon``    -- ``regularPolygon`` (Code/loop.cpp:585-608)
* ``Loop.polyline``      -- ``loop::polyline`` (Code/loop.cpp:536-580) for sharp
                            shapes and circles (no corner rounding)
* ``Loop.loopons``       -- ``loop::generateLoopons`` / ``insideUV``
                            (Code/loop.cpp:253-316), sharp shapes
* ``Loop.loop_data``     -- the driver's ``LoopData`` assembly (Code/main.cpp:689-711)

Scalar trigonometry goes through Python's ``math`` (the C library), as the
reference's does, so circle and polygon vertices agree with it to the last bit
where the compiler does not contract products into FMAs.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

from .backend import LoopData

# ----------------------------------------------------------------- lattices
_BASIS = {
    "BCC": ((0.0, 0.0, 0.0), (0.5, 0.5, 0.5)),
    "FCC": ((0.0, 0.0, 0.0), (0.5, 0.5, 0.0), (0.5, 0.0, 0.5), (0.0, 0.5, 0.5)),
}


def lattice(kind: str, nx: int, ny: int, nz: int, a0: float):
    """Perfect cubic lattice centred on the origin, in the reference's atom order.

    Each coordinate is ``i*a0 + f*a0 - h`` evaluated left to right, as in
    Code/lattice.hpp:47, so positions are bit-identical to the reference's.
    """
    basis = _BASIS[kind]
    nb = len(basis)
    hx, hy, hz = nx * a0 / 2.0, ny * a0 / 2.0, nz * a0 / 2.0
    i = np.arange(nx, dtype=np.float64)[:, None, None, None]
    j = np.arange(ny, dtype=np.float64)[None, :, None, None]
    k = np.arange(nz, dtype=np.float64)[None, None, :, None]
    f = np.asarray(basis, dtype=np.float64)
    shape = (nx, ny, nz, nb)
    x = np.broadcast_to((i * a0 + f[:, 0] * a0) - hx, shape).reshape(-1)
    y = np.broadcast_to((j * a0 + f[:, 1] * a0) - hy, shape).reshape(-1)
    z = np.broadcast_to((k * a0 + f[:, 2] * a0) - hz, shape).reshape(-1)
    return np.ascontiguousarray(x), np.ascontiguousarray(y), np.ascontiguousarray(z)


def plane_spacing(kind: str, h, a0: float) -> float:
    """Spacing of the (hkl) planes, h coprime (Code/lattice.hpp:146-153)."""
    norm = math.sqrt(float(h[0] * h[0] + h[1] * h[1] + h[2] * h[2]))
    if kind == "FCC":
        full = (abs(h[0]) % 2) and (abs(h[1]) % 2) and (abs(h[2]) % 2)
    else:
        full = abs(h[0] + h[1] + h[2]) % 2 == 0
    return a0 / norm if full else a0 / (2.0 * norm)


def snap_to_mid_plane(center, n, a0: float, nx: int, ny: int, nz: int, kind: str):
    """Move ``center`` along n onto the nearest mid-plane between (hkl) planes.

    Code/loop.cpp:342-374; returns (new centre, plane spacing).
    """
    h = [int(round(v)) for v in n]
    if any(abs(n[i] - h[i]) > 1e-9 for i in range(3)):
        raise ValueError("loop normal n must be integer Miller indices")
    g = 0
    for v in h:
        a = abs(v)
        while a:
            g, a = a, g % a
    if g == 0:
        raise ValueError("loop normal n is zero")
    h = [v // g if v >= 0 else -((-v) // g) for v in h]
    norm = math.sqrt(float(h[0] * h[0] + h[1] * h[1] + h[2] * h[2]))
    nh = [h[0] / norm, h[1] / norm, h[2] / norm]
    spacing = plane_spacing(kind, h, a0)
    o = [-0.5 * a0 * nx, -0.5 * a0 * ny, -0.5 * a0 * nz]
    s0 = o[0] * nh[0] + o[1] * nh[1] + o[2] * nh[2]
    s = center[0] * nh[0] + center[1] * nh[1] + center[2] * nh[2]
    s_mid = s0 + (math.floor((s - s0) / spacing + 1e-9) + 0.5) * spacing
    return [center[i] + (s_mid - s) * nh[i] for i in range(3)], spacing


def _unit(v):
    l = math.sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2])
    return [v[0] / l, v[1] / l, v[2] / l]


def _cross(a, b):
    return [a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]]


def default_axes(m, n):
    """In-plane grid axes (u, v) for unit m, n (Code/loop.cpp:40-57)."""
    mn = m[0] * n[0] + m[1] * n[1] + m[2] * n[2]
    u = [m[0] - mn * n[0], m[1] - mn * n[1], m[2] - mn * n[2]]
    l = math.sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2])
    if l < 1e-6:
        k = 0
        for i in (1, 2):
            if abs(n[i]) < abs(n[k]):
                k = i
        e = [0.0, 0.0, 0.0]
        e[k] = 1.0
        u = _cross(n, e)
        l = math.sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2])
    u = [x / l for x in u]
    v = _cross(u, n)
    return u, v


def regular_polygon(sides: int, R: float, m, n):
    """Regular polygon vertices (offsets from the centre), Code/loop.cpp:585-608."""
    if sides < 3:
        raise ValueError("a polygon loop needs Sides >= 3")
    u, v = default_axes(_unit(m), _unit(n))
    out = []
    for k in range(sides):
        th = 2.0 * math.pi * k / sides
        out.append([R * (math.cos(th) * u[i] + math.sin(th) * v[i]) for i in range(3)])
    return out


# ----------------------------------------------------------------- loops
@dataclass
class Loop:
    """A planar dislocation loop (sharp shapes and circles).

    shape: "circle" (radius), "polygon" (vertices: in-plane offsets),
    "rect" (half_m, half_p), "parallelogram" (half1, half2, dir1, dir2).
    ``m`` is the Burgers direction in units of a0/2 (|b| = a0 |m| / 2,
    Code/loop.cpp:25-29), ``n`` the plane normal (Miller indices).
    """
    shape: str
    center: list
    m: list
    n: list
    a0: float
    radius: float = 0.0
    vertices: list = field(default_factory=list)
    half_m: float = 0.0
    half_p: float = 0.0
    dir1: list = field(default_factory=list)
    dir2: list = field(default_factory=list)
    corner_r: float = 0.0          # rect / polygon corner radius (native producer only)
    axis: list = field(default_factory=list)   # rect Side / polygon grid axis (native only)

    def __post_init__(self):
        normm = math.sqrt(sum(t * t for t in self.m))
        self.mh = [t / normm for t in self.m]
        self.nh = _unit(self.n)
        self.bmag = self.a0 * 0.5 * normm            # compress = 1
        if self.shape == "parallelogram":
            self.u, self.v = _unit(self.dir1), _unit(self.dir2)
        else:
            self.u, self.v = default_axes(self.mh, self.nh)
        if self.shape == "polygon":
            pu = [x[0] * self.u[0] + x[1] * self.u[1] + x[2] * self.u[2] for x in self.vertices]
            pv = [x[0] * self.v[0] + x[1] * self.v[1] + x[2] * self.v[2] for x in self.vertices]
            a2 = sum(pu[i] * pv[(i + 1) % len(pu)] - pu[(i + 1) % len(pu)] * pv[i]
                     for i in range(len(pu)))
            if a2 < 0.0:
                pu.reverse()
                pv.reverse()
            self.pu, self.pv = pu, pv
            self.half_m = max(abs(t) for t in pu)
            self.half_p = max(abs(t) for t in pv)

    # --- contour
    def _plane_coords(self, nseg=None, eps=None):
        if self.shape == "circle":
            if nseg is None:
                c = 1.0 - eps / self.radius
                step = math.pi if c <= -1.0 else 2.0 * math.acos(max(c, -1.0))
                nseg = max(8, max(1, int(math.ceil(2.0 * math.pi / step - 1e-12))))
            rp = self.radius * math.sqrt(2.0 * math.pi / (nseg * math.sin(2.0 * math.pi / nseg)))
            a = [rp * math.cos(2.0 * math.pi * k / nseg) for k in range(nseg)]
            b = [rp * math.sin(2.0 * math.pi * k / nseg) for k in range(nseg)]
            return a, b
        if self.shape in ("rect", "parallelogram"):
            return ([self.half_m, -self.half_m, -self.half_m, self.half_m],
                    [self.half_p, self.half_p, -self.half_p, -self.half_p])
        a, b = [], []
        for pu, pv in zip(self.pu, self.pv):
            if a and abs(a[-1] - pu) < 1e-12 and abs(b[-1] - pv) < 1e-12:
                continue
            a.append(pu)
            b.append(pv)
        if len(a) > 1 and abs(a[-1] - a[0]) < 1e-12 and abs(b[-1] - b[0]) < 1e-12:
            a.pop()
            b.pop()
        return a, b

    def polyline(self, eps: float = 0.01, nseg: int | None = None):
        """Closed contour, absolute positions, counter-clockwise about n.

        Code/loop.cpp:536-580 (circle: ``nseg`` chords of the area-preserving
        radius, or as many as sagitta ``eps`` needs).
        """
        a, b = self._plane_coords(nseg=nseg, eps=eps)
        u, v, c = self.u, self.v, self.center
        out = [[(u[i] * a[k] + v[i] * b[k]) + c[i] for i in range(3)] for k in range(len(a))]
        s = 0.0
        N = len(out)
        for k in range(N):
            p, q = out[k], out[(k + 1) % N]
            r1 = [p[i] - c[i] for i in range(3)]
            r2 = [q[i] - c[i] for i in range(3)]
            cr = _cross(r1, r2)
            s += cr[0] * self.nh[0] + cr[1] * self.nh[1] + cr[2] * self.nh[2]
        if s < 0.0:
            out.reverse()
        return np.asarray(out, dtype=np.float64)

    def subdivided_polyline(self, nseg_total: int):
        """Sharp polygon contour with each side cut into collinear pieces so that
        the polyline has ``nseg_total`` segments (the field is unchanged)."""
        P = self.polyline()
        ns = len(P)
        per = [nseg_total // ns + (1 if i < nseg_total % ns else 0) for i in range(ns)]
        out = []
        for i in range(ns):
            A, B = P[i], P[(i + 1) % ns]
            for j in range(per[i]):
                t = j / per[i]
                out.append(A + t * (B - A))
        return np.asarray(out, dtype=np.float64)

    # --- loopons
    def inside_uv(self, u, v):
        eps = 1e-9
        if self.shape == "circle":
            return np.sqrt(u * u + v * v) < self.radius - eps
        if self.shape in ("rect", "parallelogram"):
            du = np.maximum(np.abs(u) - max(self.half_m, 0.0), 0.0)
            dv = np.maximum(np.abs(v) - max(self.half_p, 0.0), 0.0)
            return du * du + dv * dv <= eps
        N = len(self.pu)
        inside = np.ones(np.shape(u), dtype=bool)
        for i in range(N):
            j = (i + 1) % N
            eu, ev = self.pu[j] - self.pu[i], self.pv[j] - self.pv[i]
            L = math.sqrt(eu * eu + ev * ev)
            nu, nv = -ev / L, eu / L
            inside &= nu * (u - self.pu[i]) + nv * (v - self.pv[i]) >= -eps
        return inside

    def surface(self, da: float) -> float:
        ds = da * da
        if self.shape == "parallelogram":
            cr = _cross(self.u, self.v)
            ds *= math.sqrt(cr[0] ** 2 + cr[1] ** 2 + cr[2] ** 2)
        return ds

    def loopons(self, da: float):
        """Loopon grid u = i da, v = j da (i outer, j inner), Code/loop.cpp:253-273."""
        ext_m = self.radius if self.shape == "circle" else self.half_m
        ext_p = self.radius if self.shape == "circle" else self.half_p
        sm = int(math.ceil(ext_m / da)) + 1
        sp = int(math.ceil(ext_p / da)) + 1
        xi = np.arange(-sm, sm + 1, dtype=np.float64)
        yi = np.arange(-sp, sp + 1, dtype=np.float64)
        U = np.repeat(xi * da, yi.size)
        V = np.tile(yi * da, xi.size)
        keep = self.inside_uv(U, V)
        U, V = U[keep], V[keep]
        pts = [(self.u[i] * U + self.v[i] * V) + self.center[i] for i in range(3)]
        return np.stack(pts, axis=1)

    # --- backend input
    def phys(self, da: float = 0.0):
        b = self.bmag
        return (b, b, b, *self.mh, *self.nh, self.surface(da) if da > 0 else 0.0)

    def burgers(self):
        return [self.bmag * t for t in self.mh]

    def line_data(self, eps: float = 0.01, nseg: int | None = None, subdivide: int | None = None):
        """LoopData for Method: line (Code/main.cpp:689-711)."""
        P = self.subdivided_polyline(subdivide) if subdivide else self.polyline(eps, nseg)
        return LoopData(phys=self.phys(), line=True, vx=P[:, 0].copy(), vy=P[:, 1].copy(),
                        vz=P[:, 2].copy(), cx=self.center[0], cy=self.center[1],
                        cz=self.center[2])

    def loopon_data(self, da: float):
        """LoopData for Method: loopons."""
        P = self.loopons(da)
        return LoopData(phys=self.phys(da), line=False, x=P[:, 0].copy(), y=P[:, 1].copy(),
                        z=P[:, 2].copy())


# ----------------------------------------------------------------- samples
@dataclass
class Sample:
    """Atoms (SoA, FP64) and loops of one workload."""
    name: str
    x: np.ndarray
    y: np.ndarray
    z: np.ndarray
    loops: list
    box: tuple
    img: int
    meta: dict = field(default_factory=dict)

    @property
    def n(self):
        return int(self.x.shape[0])

    def segments(self):
        return int(sum(len(l.vx) if l.line else len(l.x) for l in self.loops))

    def evals(self):
        """atom x segment (or loopon) x image evaluations of one pass."""
        return self.n * self.segments() * (2 * self.img + 1) ** 3

    def subsample(self, idx):
        idx = np.asarray(idx)
        return Sample(self.name + "-sub", self.x[idx].copy(), self.y[idx].copy(),
                      self.z[idx].copy(), self.loops, self.box, self.img, dict(self.meta))


FCC_111 = [(1, 1, 1), (-1, 1, 1), (1, -1, 1), (1, 1, -1)]
BCC_110 = [(1, 1, 0), (1, -1, 0), (1, 0, 1), (1, 0, -1), (0, 1, 1), (0, 1, -1)]
_FCC_B = [(1, 1, 0), (1, -1, 0), (1, 0, 1), (1, 0, -1), (0, 1, 1), (0, 1, -1)]
_BCC_B = [(1, 1, 1), (-1, 1, 1), (1, -1, 1), (1, 1, -1)]


def _in_plane(bs, n):
    return [b for b in bs if b[0] * n[0] + b[1] * n[1] + b[2] * n[2] == 0]


def config_c1():
    """BCC a0 = 2.8553, 40^3 cells (128,000 atoms); two 64-gon glide loops,
    R = 30, b = a/2[111] on (1-10), centres at -L/4 and +L/4 of the diagonal
    (snapped), images +-1."""
    a0, N = 2.8553, 40
    x, y, z = lattice("BCC", N, N, N, a0)
    L = N * a0
    m, n = [1, 1, 1], [1, -1, 0]
    loops = []
    for s in (-0.25, 0.25):
        c, _ = snap_to_mid_plane([s * L] * 3, n, a0, N, N, N, "BCC")
        lp = Loop("polygon", c, m, n, a0, vertices=regular_polygon(64, 30.0, m, n))
        loops.append(lp.line_data())
    return Sample("C1", x, y, z, loops, (L, L, L), 1, {"lattice": "BCC", "a0": a0, "cells": N})


def _fcc_loop_set(rng, nloops, a0, N):
    L = N * a0
    shapes = np.array(["circle"] * int(round(0.6 * nloops)) +
                      ["hexagon"] * (nloops - int(round(0.6 * nloops))))
    rng.shuffle(shapes)
    loops, geo = [], []
    for l in range(nloops):
        R = float(rng.uniform(10.0, 60.0))
        c0 = [float(t) for t in rng.uniform(-L / 2, L / 2, 3)]
        n = FCC_111[int(rng.integers(4))]
        m = _in_plane(_FCC_B, n)[int(rng.integers(3))]
        c, _ = snap_to_mid_plane(c0, n, a0, N, N, N, "FCC")
        if shapes[l] == "circle":
            lp = Loop("circle", c, list(m), list(n), a0, radius=R)
            loops.append(lp.line_data(nseg=48))
            geo.append((0, R, c, lp.nh, lp.burgers()))
        else:
            lp = Loop("polygon", c, list(m), list(n), a0, vertices=regular_polygon(6, R, m, n))
            loops.append(lp.line_data())
            geo.append((1, R, c, lp.nh, lp.burgers()))
    return loops, geo


def config_c2(seed: int = 20241, nloops: int = 50):
    """FCC a0 = 3.615, 100^3 cells (4,000,000 atoms); 50 glide loops, radius
    U(10, 60), centres U(box) snapped, n one of the 4 {111}, b one of the 3
    in-plane a/2<110>; 60 % circles (48 segments), 40 % hexagons; images +-2."""
    a0, N = 3.615, 100
    rng = np.random.default_rng(seed)
    loops, geo = _fcc_loop_set(rng, nloops, a0, N)
    x, y, z = lattice("FCC", N, N, N, a0)
    L = N * a0
    return Sample("C2", x, y, z, loops, (L, L, L), 2,
                  {"lattice": "FCC", "a0": a0, "cells": N, "seed": seed, "geo": geo})


def config_c3(seed: int = 20243, nloops: int = 20, da: float = 1.585):
    """BCC 60^3 (432,000 atoms); 20 circular loops R = 40 on random {110},
    b one of the 2 in-plane a/2<111>, loopon step da (2,001 loopons each),
    images +-1."""
    a0, N = 2.8553, 60
    rng = np.random.default_rng(seed)
    L = N * a0
    loops = []
    for _ in range(nloops):
        c0 = [float(t) for t in rng.uniform(-L / 2, L / 2, 3)]
        n = BCC_110[int(rng.integers(6))]
        m = _in_plane(_BCC_B, n)[int(rng.integers(2))]
        c, _ = snap_to_mid_plane(c0, n, a0, N, N, N, "BCC")
        loops.append(Loop("circle", c, list(m), list(n), a0, radius=40.0).loopon_data(da))
    x, y, z = lattice("BCC", N, N, N, a0)
    return Sample("C3", x, y, z, loops, (L, L, L), 1,
                  {"lattice": "BCC", "a0": a0, "cells": N, "seed": seed, "da": da})


def _c4_draw(rng, a0: float):
    """Random draws of one batched-workload sample: the side N (cells) and its
    loops as (Loop, nseg for a circle, subdivide for a sharp shape)."""
    N = int(rng.integers(20, 61))
    L = N * a0
    nl = int(rng.integers(1, 17))
    loops = []
    for _ in range(nl):
        nseg = int(rng.integers(16, 129))
        size = min(float(rng.uniform(8.0, 40.0)), L / 3.0)
        n = BCC_110[int(rng.integers(6))]
        m = _in_plane(_BCC_B, n)[int(rng.integers(2))]
        c0 = [float(t) for t in rng.uniform(-L / 2, L / 2, 3)]
        c, _ = snap_to_mid_plane(c0, n, a0, N, N, N, "BCC")
        kind = int(rng.integers(3))
        if kind == 0:
            loops.append((Loop("circle", c, list(m), list(n), a0, radius=size), nseg, None))
        elif kind == 1:
            lp = Loop("rect", c, list(m), list(n), a0, half_m=size,
                      half_p=size * float(rng.uniform(0.5, 1.0)))
            loops.append((lp, None, max(nseg, 4)))
        else:
            u, _v = default_axes(_unit(m), _unit(n))
            w = _cross(_unit(n), u)
            ang = float(rng.uniform(0.35, 1.2))
            d2 = [math.cos(ang) * u[i] + math.sin(ang) * w[i] for i in range(3)]
            lp = Loop("parallelogram", c, list(m), list(n), a0, half_m=size,
                      half_p=size * float(rng.uniform(0.5, 1.0)), dir1=u, dir2=d2)
            loops.append((lp, None, max(nseg, 4)))
    return N, loops


def config_c4_sample(rng, a0: float = 2.8553):
    """One sample of the batched workload: BCC side U{20..60} cells, 1-16 loops
    of 16-128 segments (circles, rectangles, parallelograms), size U(8, 40)
    capped at a third of the box, images +-1."""
    N, draws = _c4_draw(rng, a0)
    loops = [lp.line_data(nseg=ns) if ns else lp.line_data(subdivide=sd) for lp, ns, sd in draws]
    x, y, z = lattice("BCC", N, N, N, a0)
    L = N * a0
    return Sample("C4", x, y, z, loops, (L, L, L), 1, {"lattice": "BCC", "a0": a0, "cells": N})


@dataclass
class SampleSpec:
    """A sample whose atoms are generated on the device: its lattice, loops, box."""
    name: str
    n: int
    loops: list
    box: tuple
    img: int
    meta: dict = field(default_factory=dict)


def config_c4_specs(seed: int = 20244, nsamples: int = 256, a0: float = 2.8553):
    """The batched workload (same draws as config_c4) with the loop contours built
    by the C++ producer in one call and no host atoms: each sample's lattice is
    generated on the device (backend.lattice_torch)."""
    rng = np.random.default_rng(seed)
    draws = [_c4_draw(rng, a0) for _ in range(nsamples)]
    flat = [d for _, ds in draws for d in ds]
    V, off = native_polylines([lp for lp, _, _ in flat], 0.01, [ns for _, ns, _ in flat],
                              [sd for _, _, sd in flat])
    out, k = [], 0
    for N, ds in draws:
        loops = []
        for lp, _, _ in ds:
            P = V[off[k]:off[k + 1]]
            loops.append(LoopData(phys=lp.phys(), line=True, vx=P[:, 0].copy(),
                                  vy=P[:, 1].copy(), vz=P[:, 2].copy(), cx=lp.center[0],
                                  cy=lp.center[1], cz=lp.center[2]))
            k += 1
        L = N * a0
        out.append(SampleSpec("C4", 2 * N * N * N, loops, (L, L, L), 1,
                              {"lattice": "BCC", "a0": a0, "cells": N}))
    return out


def config_c4(seed: int = 20244, nsamples: int = 256):
    rng = np.random.default_rng(seed)
    return [config_c4_sample(rng) for _ in range(nsamples)]


# ----------------------------------------------------------------- native producer
# The same producer in C++ (include/disco_geometry.h), for batched sample
# generation: thousands of loops per batch without a Python loop per vertex.
import ctypes as _ct


class _Shape(_ct.Structure):
    _fields_ = [("shape", _ct.c_int), ("center", _ct.c_double * 3), ("m", _ct.c_double * 3),
                ("n", _ct.c_double * 3), ("a0", _ct.c_double), ("radius", _ct.c_double),
                ("half_m", _ct.c_double), ("half_p", _ct.c_double),
                ("dir1", _ct.c_double * 3), ("dir2", _ct.c_double * 3), ("nvert", _ct.c_int),
                ("vert", _ct.POINTER(_ct.c_double)), ("corner_r", _ct.c_double),
                ("axis", _ct.c_double * 3)]


_MAXV = 128


class _LoopState(_ct.Structure):
    """disco_loop_state: a loop as the reference's constructors leave it."""
    _fields_ = [("shape", _ct.c_int), ("nvert", _ct.c_int), ("center", _ct.c_double * 3),
                ("m", _ct.c_double * 3), ("n", _ct.c_double * 3), ("u", _ct.c_double * 3),
                ("v", _ct.c_double * 3), ("bmag", _ct.c_double), ("radius", _ct.c_double),
                ("half_m", _ct.c_double), ("half_p", _ct.c_double), ("r_corner", _ct.c_double)] + \
               [(k, _ct.c_double * _MAXV) for k in ("pu", "pv", "qu", "qv", "nu", "nv")]


_SHAPES = {"circle": 0, "polygon": 1, "rect": 2, "parallelogram": 3}
_native = None


def _nlib():
    global _native
    if _native is None:
        from .backend import lib
        L = lib()
        dp, vp, i, d = _ct.POINTER(_ct.c_double), _ct.c_void_p, _ct.c_int, _ct.c_double
        L.disco_geo_regular_polygon.argtypes = [i, d, dp, dp, dp]
        L.disco_geo_snap_mid_plane.argtypes = [dp, dp, d, i, i, i, i, dp, dp]
        L.disco_geo_polyline.argtypes = [_ct.POINTER(_Shape), d, i, i, dp, dp, dp, i]
        L.disco_geo_loopons.argtypes = [_ct.POINTER(_Shape), d, dp, dp, dp, i]
        L.disco_geo_phys.argtypes = [_ct.POINTER(_Shape), d, dp]
        L.disco_geo_polylines.argtypes = [_ct.POINTER(_Shape), i, d, _ct.POINTER(i),
                                          _ct.POINTER(i), _ct.POINTER(_ct.c_longlong), dp, dp,
                                          dp, _ct.c_longlong]
        L.disco_geo_polylines.restype = _ct.c_longlong
        L.disco_geo_loop_state.argtypes = [_ct.POINTER(_Shape), _ct.POINTER(_LoopState)]
        L.disco_geo_classify_burgers.argtypes = [i, dp, dp]
        L.disco_geo_frank_copy_shift.argtypes = [dp, dp, d, dp]
        L.disco_geo_copy_shifts.argtypes = [dp, i, dp, vp, d, dp]
        _native = L
    return _native


def _dptr(a):
    return a.ctypes.data_as(_ct.POINTER(_ct.c_double))


def native_shape(lp: Loop):
    """The C struct of a loop (keeps the polygon vertices alive with it)."""
    s = _Shape()
    s.shape = _SHAPES[lp.shape]
    s.center[:] = [float(t) for t in lp.center]
    s.m[:] = [float(t) for t in lp.m]
    s.n[:] = [float(t) for t in lp.n]
    s.a0, s.radius, s.half_m, s.half_p = lp.a0, lp.radius, lp.half_m, lp.half_p
    if lp.shape == "parallelogram":
        s.dir1[:] = [float(t) for t in lp.dir1]
        s.dir2[:] = [float(t) for t in lp.dir2]
    s.corner_r = float(lp.corner_r)
    if lp.axis:
        s.axis[:] = [float(t) for t in lp.axis]
    keep = None
    if lp.shape == "polygon":
        keep = np.ascontiguousarray(np.asarray(lp.vertices, dtype=np.float64).reshape(-1))
        s.nvert = len(lp.vertices)
        s.vert = _dptr(keep)
    return s, keep


def _call_points(fn, *args):
    n = fn(*args, None, None, None, 0)
    if n < 0:
        raise ValueError("invalid loop shape")
    x, y, z = np.zeros(n), np.zeros(n), np.zeros(n)
    fn(*args, _dptr(x), _dptr(y), _dptr(z), n)
    return np.stack([x, y, z], 1)


def native_polyline(lp: Loop, eps: float = 0.01, nseg: int | None = None,
                    subdivide: int | None = None):
    s, keep = native_shape(lp)
    return _call_points(_nlib().disco_geo_polyline, _ct.byref(s), float(eps), int(nseg or 0),
                        int(subdivide or 0))


def native_loopons(lp: Loop, da: float):
    s, keep = native_shape(lp)
    return _call_points(_nlib().disco_geo_loopons, _ct.byref(s), float(da))


def native_phys(lp: Loop, da: float = 0.0):
    s, keep = native_shape(lp)
    out = np.zeros(10)
    if _nlib().disco_geo_phys(_ct.byref(s), float(da), _dptr(out)) < 0:
        raise ValueError("invalid loop shape")
    return tuple(out)


def native_loop_state(lp: Loop) -> _LoopState:
    """The loop's state as the reference's constructors leave it (Code/loop.cpp:13-244)."""
    s, keep = native_shape(lp)
    st = _LoopState()
    if _nlib().disco_geo_loop_state(_ct.byref(s), _ct.byref(st)) < 0:
        raise ValueError("invalid loop shape")
    return st


def native_regular_polygon(sides: int, R: float, m, n):
    out = np.zeros(3 * sides)
    mm, nn = np.asarray(m, dtype=np.float64), np.asarray(n, dtype=np.float64)
    if _nlib().disco_geo_regular_polygon(sides, float(R), _dptr(mm), _dptr(nn), _dptr(out)) < 0:
        raise ValueError("a polygon loop needs Sides >= 3")
    return out.reshape(sides, 3)


def native_snap_to_mid_plane(center, n, a0, nx, ny, nz, kind):
    c, nn = np.asarray(center, dtype=np.float64), np.asarray(n, dtype=np.float64)
    out, sp = np.zeros(3), np.zeros(1)
    if _nlib().disco_geo_snap_mid_plane(_dptr(c), _dptr(nn), float(a0), nx, ny, nz,
                                        1 if kind == "FCC" else 0, _dptr(out), _dptr(sp)) < 0:
        raise ValueError("loop normal n must be non-zero integer Miller indices")
    return list(out), float(sp[0])


def native_polylines(loops, eps: float = 0.01, nseg=None, subdivide=None):
    """Many loops' polylines in one call: (vertices (N x 3), offsets (nloops + 1))."""
    shapes = (_Shape * max(1, len(loops)))()
    keep = []
    for k, lp in enumerate(loops):
        s, kp = native_shape(lp)
        shapes[k] = s
        keep.append(kp)
    nl = len(loops)
    ns = (_ct.c_int * max(1, nl))(*[int(v or 0) for v in (nseg or [0] * nl)])
    sd = (_ct.c_int * max(1, nl))(*[int(v or 0) for v in (subdivide or [0] * nl)])
    off = np.zeros(nl + 1, dtype=np.int64)
    offp = off.ctypes.data_as(_ct.POINTER(_ct.c_longlong))
    L = _nlib()
    total = L.disco_geo_polylines(shapes, nl, float(eps), ns, sd, offp, None, None, None, 0)
    if total < 0:
        raise ValueError("invalid loop shape")
    x, y, z = np.zeros(total), np.zeros(total), np.zeros(total)
    L.disco_geo_polylines(shapes, nl, float(eps), ns, sd, offp, _dptr(x), _dptr(y), _dptr(z),
                          total)
    return np.stack([x, y, z], 1), off


BURGERS = {0: "perfect", 1: "Frank", 2: "Shockley", 3: "other"}


def _v3(v):
    return np.ascontiguousarray(np.asarray(v, dtype=np.float64).reshape(3))


def native_classify_burgers(kind: str, m, n) -> int:
    """classifyBurgers (Code/lattice.hpp:189-212) of b = (a0/2) m on the plane n,
    crystal frame: 0 perfect, 1 Frank, 2 Shockley, 3 other."""
    mm, nn = _v3(m), _v3(n)
    return _nlib().disco_geo_classify_burgers(1 if kind == "FCC" else 0, _dptr(mm), _dptr(nn))


def native_frank_copy_shift(v, nhat, a0: float):
    """frankCopyShift (Code/lattice.hpp:239-256): v plus the in-plane a0/6<112>
    that puts an inserted {111} layer on the third stacking position."""
    vv, nn, out = _v3(v), _v3(nhat), np.zeros(3)
    _nlib().disco_geo_frank_copy_shift(_dptr(vv), _dptr(nn), float(a0), _dptr(out))
    return out


def native_copy_shifts(b, frank: bool, n, R, a0: float):
    """The copy shifts of an interstitial loop (Code/main.cpp:633-650): a (2, 3)
    array, row 0 for the atoms on the +n side of the cut.  b = |b| m and n (unit)
    in the lab frame; R the grain's 3x3 orientation (rows: crystal directions of
    the box axes) or None."""
    bb, nn, out = _v3(b), _v3(n), np.zeros(6)
    Rm = None if R is None else np.ascontiguousarray(np.asarray(R, dtype=np.float64).reshape(9))
    _nlib().disco_geo_copy_shifts(_dptr(bb), int(bool(frank)), _dptr(nn),
                                  None if Rm is None else Rm.ctypes.data, float(a0), _dptr(out))
    return out.reshape(2, 3)
