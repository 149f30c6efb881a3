"""Cubature form of the periodic image sum (host side, on the C ABI).

The plain sum (Code/backend_cpu.cpp:38-99, generalised to images [-img, img]^3)
evaluates the closed-form line integral (Code/barnett.hpp:44-107) of every
loop at every image.  Far from a loop image its field is that of the Kroupa
term (Code/kroupa.hpp:52-73) integrated over the cut, which a 9-point cubature
(the centre and a ring of eight matching the cut's area and its second and
fourth radial moments) reproduces to O((R/r)^8) of |b|: measured against the
FP64 line integral on the CPU, at most about 0.15 (R/r)^8 A for |b| = 2.56 A
(circles and regular hexagons).  The cubature form is

* the +-1 shell by an FP32C line plan whose far pass takes, per lane and loop
  image, the exact field within k_exact cut radii of the image's centre, the
  cubature beyond k_rule radii and a smooth blend between
  (``disco_hip_plan_set_cubature``); the near slab stays FP64 exact, and loops
  whose cut the rule does not match (rectangles, parallelograms, ...) stay
  exact;
* the far shell (images with max(|i|, |j|, |k|) in (1, img]), at least one box
  length from every atom, by the same cubature summed by the loopon kernels at
  the nodes of a grid of spacing ``grid_h`` over the atoms and interpolated to
  them by 6-node Lagrange (``disco_hip_grid_interp``); so are the +-1 images so
  far from the atoms of an octant of the atom box that the cubature alone is
  accurate at all of them and a finer grid over that octant interpolates
  their field to a stated tolerance (``far_images``; the line plan leaves them
  out for those atoms only, ``Plan.set_image_skip`` with the octant planes).
  A loop whose cut is not a flat region about its apex (off its plane, lobes
  of opposite sense, ...) is summed there as Radon's rule on each of its fan
  triangles instead (``surrogate.fan_loopons``), and its +-1 images stay in
  the line plan.

Its error against the plain sum is measured, never assumed: the bench and the
tests compare it with the FP64 plain sum.
"""
from __future__ import annotations

import itertools

import numpy as np

from . import backend as B
from . import surrogate as S


# Interpolation error model of a loop image's field on a grid of spacing h by
# tensor-product Lagrange on p nodes per axis, at distance r from the image's
# centre: the field is about 0.31 |b| (R/r)^2 there, its p-th derivatives
# about (p + 1)! times that over r^p, and the p-node interpolant errs by at
# most max|prod (t - t_i)| / p! h^p f^(p) per axis (9/16 and 225/64 for p = 4
# and 6): error <= GRID_ERR[p] |b| R^2 h^p / r^(p+2) over three axes (for p = 4
# an upper bound by a factor of about 2.5 on the 4M-atom samples, measured on
# the GPU).
GRID_ERR = {4: 3 * 0.31 * 120 * (9 / 16) / 24, 6: 3 * 0.31 * 5040 * (225 / 64) / 720}


def far_images(loops, box, lo, hi, h, tol, k_exact, dmin=0.0, info=None, order=6, k_band=None,
               band=None):
    """Images of the +-1 shell to sum on a grid of spacing h instead of at the
    atoms (loops whose cut the cubature represents only): those whose centre
    lies at least max(k_exact R, R + dmin, (GRID_ERR |b| R^2 h^p / tol)^(1/(p+2)))
    (p = order) from the box [lo, hi] holding the atoms, so that the cubature
    alone is accurate at every atom and the grid's interpolation error is below
    tol.  With k_band (and band = B.line_band(loops)), an image of a loop the
    two-ring cubature represents also goes to the grid when its centre lies at
    least max(k_band R + 3 h, R + dmin, the interpolation bound) from the box:
    every node an atom's stencil reaches (within 3 h of the box) is then at
    least k_band R from the image's centre, where that cubature is accurate.
    Returns the skip mask of each loop (bit 9 (i + 1) + 3 (j + 1) + (k + 1))
    and the list of (loop index, offset, two_ring) of those images, offset =
    (i Lx, j Ly, k Lz) (the image of a loop centred at c sits at c + offset),
    two_ring true when the grid must sum the two-ring cubature there."""
    info = B.line_cubature(loops) if info is None else info
    lo, hi = np.asarray(lo, float), np.asarray(hi, float)
    masks = np.zeros(len(loops), np.uint32)
    out = []
    for l, lp in enumerate(loops):
        if not info[l, 0]:
            continue
        c = np.array([lp.cx, lp.cy, lp.cz])
        R = float(info[l, 4])
        bm = float(np.linalg.norm(lp.burgers()))
        interp = (GRID_ERR[order] * bm * R * R * h ** order / tol) ** (1 / (order + 2))
        dmin_l = max(k_exact * R, R + dmin, interp)
        dmin_b = np.inf
        if k_band is not None and band is not None and band[l, 0]:
            dmin_b = max(k_band * R + 3.0 * h, R + dmin, interp)
        for i, j, k in itertools.product((-1, 0, 1), repeat=3):
            if i == j == k == 0:
                continue
            off = np.array([i * box[0], j * box[1], k * box[2]])
            ci = c + off
            gap = np.maximum(np.maximum(lo - ci, ci - hi), 0.0)
            d = float(np.sqrt((gap * gap).sum()))
            if d >= dmin_l or d >= dmin_b:
                masks[l] |= np.uint32(1 << (9 * (i + 1) + 3 * (j + 1) + (k + 1)))
                out.append((l, off, d < dmin_l))
    return masks, out


def band_loopons(loops, band):
    """Per loop, the loopons of its two-ring cubature ([centre, ring 1, ring 2];
    None where band[l, 0] is 0): weights and radii from B.line_band, ring 1 at
    angles j pi/6 from the in-plane direction of the loop's first vertex, ring 2
    at j pi/6 + pi/12 (the rule is isotropic through rank 11, so the rings'
    orientation enters at degree 12 only)."""
    out = [None] * len(loops)
    th = 2.0 * np.pi * np.arange(12) / 12
    for l, lp in enumerate(loops):
        if not band[l, 0]:
            continue
        c = np.array([lp.cx, lp.cy, lp.cz])
        V = np.stack([lp.vx, lp.vy, lp.vz], 1) - c
        N = np.cross(V, np.roll(V, -1, axis=0)).sum(0)
        n = N / np.linalg.norm(N)
        e1 = V[0] - (V[0] @ n) * n
        e1 /= np.linalg.norm(e1)
        e2 = np.cross(n, e1)
        b = np.array(lp.burgers())
        bm = float(np.linalg.norm(b))
        m = b / bm
        parts = []
        for rho, w, ph in ((0.0, band[l, 1], None), (band[l, 4], band[l, 2], 0.0),
                           (band[l, 5], band[l, 3], np.pi / 12)):
            if ph is None:
                P = c[None]
            else:
                P = c + rho * (np.cos(th + ph)[:, None] * e1 + np.sin(th + ph)[:, None] * e2)
            parts.append(B.LoopData(phys=(bm, bm, bm, *m, *n, float(w)), line=False,
                                    x=P[:, 0].copy(), y=P[:, 1].copy(), z=P[:, 2].copy()))
        out[l] = parts
    return out


class _Grid:
    """Loopon plans summed at the nodes of a uniform grid over the atoms and
    interpolated to them (tensor-product Lagrange on `order` nodes per axis).
    With idx (a device index tensor), the grid covers the atoms idx of the
    arrays run() is given, and adds to their u only."""

    def __init__(self, ctx, x, y, z, box, h, parts, groups=None, order=6, idx=None):
        import torch
        self.ctx, self.h, self.order = ctx, float(h), int(order)
        self.idx = idx
        if idx is not None:
            x, y, z = x[idx].contiguous(), y[idx].contiguous(), z[idx].contiguous()
            self.sub = (x, y, z)
            self.su = [torch.zeros_like(x) for _ in range(3)]
        self.origin, dims, self.nodes = S.grid_nodes(x, y, z, self.h, margin=self.order // 2 - 1)
        tiles = (self.nodes[0].numel() + 255) // 256
        self.plans = []
        for lps, img, inner in parts:
            want = groups if groups else -(-8192 // tiles)
            p = B.Plan(ctx, lps, box, img, B.PREC_FP32C, kind="loopon")
            if inner >= 0:
                S.set_inner(p, inner)
            S.set_split(p, max(1, min(want, len(lps), 256)))
            self.plans.append(p)
        self.shape = (dims[2], dims[1], dims[0])
        self.gu = [torch.zeros_like(self.nodes[0]) for _ in range(3)]

    def run(self, x, y, z, ux, uy, uz):
        for k, p in enumerate(self.plans):
            p.run_torch(*self.nodes, *self.gu, accumulate=k > 0)
        g = [t.view(self.shape) for t in self.gu]
        if self.idx is None:
            B.grid_interp_torch(self.ctx, x, y, z, ux, uy, uz, g, self.origin, self.h,
                                accumulate=True, order=self.order)
            return
        B.grid_interp_torch(self.ctx, *self.sub, *self.su, g, self.origin, self.h,
                            accumulate=False, order=self.order)
        for u, v in zip((ux, uy, uz), self.su):
            u.index_add_(0, self.idx, v)

    def free(self):
        for p in self.plans:
            p.free()
        self.plans = []


class CubatureSum:
    """u (+)= sum over loops x images [-img, img]^3 of line loops in cubature
    form, at device atoms x, y, z (FP64 torch tensors).  The plans and the grids
    are built once (for one atom set and loop set); run() launches them.

    near_tol (A): images of the +-1 shell far enough from every atom that the
    cubature alone is accurate there and a grid of spacing near_h (default
    L/32) interpolates their field to within near_tol (far_images) are taken
    out of the line plan and summed as cubatures at that grid's nodes (None:
    every +-1 image stays in the line plan).  The far shell (1, img] is summed
    as cubatures on a grid of spacing grid_h (default L/12), its images beyond
    shell 2 on one of spacing outer_h (default L/6; None: on the grid_h grid
    too).  The grids are interpolated by 6-node (quintic) stencils (order = 4:
    cubic, with defaults L/64 and L/24 and no outer grid).  Defaults: measured
    on the 4M-atom samples at images +-2 and +-4 (max |du| 1.0e-6 and 2.1e-6 A
    against the FP64 plain sum, every atom).

    k_band (None: off): lanes between k_band and k_exact cut radii of a loop
    image take the two-ring cubature (centre and two rings of twelve, error
    O((R/r)^12)) in place of the exact line integral, for the loops that rule
    represents (circles and regular polygons of twelve sides or more); the
    +-1 images at least k_band R + 3 near_h from the atoms (and beyond the
    interpolation bound) then go to the near_h grid with that cubature
    (far_images).

    octants (default True): the +-1 images are chosen per octant of the atoms
    (split at the middle of their bounding box): an image far enough from the
    atoms of one octant goes to that octant's near_h grid and is left out of
    the line plan for those atoms only (Plan.set_image_skip with split), so
    far fewer (atom, loop image) pairs stay exact; False: one grid, images far
    from every atom."""

    def __init__(self, ctx, x, y, z, loops, box, img, k_exact=5.0, k_rule=None,
                 grid_h=None, grid_groups=None, near_tol=3e-7, near_h=None, order=6,
                 k_band=None, octants=True, outer_h=-1.0):
        self.ctx = ctx
        self.img = int(img)
        self.near = B.Plan(ctx, loops, box, min(1, self.img), B.PREC_FP32C)
        self.eligible = self.near.set_cubature(k_exact, k_rule)
        self.band = self.near.set_band(k_band) if k_band else 0
        self.nloops = len(loops)
        self.grids = []
        self.skipped = 0
        self.skipped_band = 0
        hf = float(grid_h) if grid_h else min(box) / (12.0 if order == 6 else 24.0)
        parts = {}                 # grid spacing -> [(loopon loops, img, inner)]
        base = None
        near_grids = []            # (atom index or None, parts) of the near_h grids
        if near_tol is not None and self.img >= 1:
            import torch
            hn = float(near_h) if near_h else min(box) / (32.0 if order == 6 else 64.0)
            lo = [float(t.min()) for t in (x, y, z)]
            hi = [float(t.max()) for t in (x, y, z)]
            info = B.line_cubature(loops)
            for l, lp in enumerate(loops):      # the grid sums the ring rule's cubature
                if info[l, 0] and not S.ring_rule_ok(lp):
                    info[l, 0] = 0.0
            band = None
            if k_band:
                band = B.line_band(loops)
                band[:, 0] *= info[:, 0]        # the grid needs the ring rule's loops
            kx = max(float(k_exact), float(k_rule or 0.0))
            # the atoms of each octant (all atoms in one set without octants);
            # the kernels take the same comparisons against the same planes
            split = [0.5 * (a + b) for a, b in zip(lo, hi)]
            if octants:
                code = ((x >= split[0]).int() + 2 * (y >= split[1]).int()
                        + 4 * (z >= split[2]).int())
                sets = [torch.nonzero(code == o).flatten() for o in range(8)]
            else:
                sets = [None]
            masks = np.zeros((len(loops), 8), np.uint32)
            base = S.rule_loopons_per_loop(loops, ring=8)
            two = band_loopons(loops, band) if band is not None else None
            for o, idx in enumerate(sets):
                if idx is not None and idx.numel() == 0:
                    continue
                sx, sy, sz = (t if idx is None else t[idx] for t in (x, y, z))
                olo = [float(t.min()) for t in (sx, sy, sz)]
                ohi = [float(t.max()) for t in (sx, sy, sz)]
                m, far = far_images(loops, box, olo, ohi, hn, float(near_tol), kx, info=info,
                                    order=order, k_band=k_band, band=band)
                if not far:
                    continue
                if idx is None:
                    masks[:] = m[:, None]
                else:
                    masks[:, o] = m
                copies = []
                for l, off, two_ring in far:
                    for part in (two[l] if two_ring else base[l]):
                        copies.append(B.LoopData(phys=part.phys, line=False, x=part.x + off[0],
                                                 y=part.y + off[1], z=part.z + off[2]))
                near_grids.append((idx, [(copies, 0, -1)]))
                self.skipped += len(far)
                self.skipped_band += sum(1 for f in far if f[2])
            if near_grids:
                if octants:
                    self.near.set_image_skip(masks, split)
                else:
                    self.near.set_image_skip(masks[:, 0])
        if self.img > 1:
            base = S.rule_loopons_per_loop(loops, ring=8) if base is None else base
            flat = [p for ps in base for p in ps]
            if outer_h is not None and outer_h < 0:
                outer_h = min(box) / 6.0 if order == 6 else None
            if self.img > 2 and outer_h:
                # shell 2 on the grid_h grid, shells 3..img on the coarser
                # outer_h grid (their field is smoother at the atoms)
                parts.setdefault(hf, []).append((flat, 2, 1))
                parts.setdefault(float(outer_h), []).append((flat, self.img, 2))
            else:
                parts.setdefault(hf, []).append((flat, self.img, 1))
        for idx, pl in near_grids:
            self.grids.append(_Grid(ctx, x, y, z, box, hn, pl, grid_groups, order, idx=idx))
        for h, pl in parts.items():
            self.grids.append(_Grid(ctx, x, y, z, box, h, pl, grid_groups, order))

    def run(self, x, y, z, ux, uy, uz, accumulate=False):
        self.near.run_torch(x, y, z, ux, uy, uz, accumulate=accumulate)
        for g in self.grids:
            g.run(x, y, z, ux, uy, uz)

    def free(self):
        import torch
        torch.cuda.synchronize()
        if self.near is not None:
            self.near.free()
        for g in self.grids:
            g.free()
        self.near = None
        self.grids = []


def cubature_sum_torch(ctx, x, y, z, loops, box, img, k_exact=5.0, k_rule=None, u=None,
                       accumulate=False, grid_h=None, near_tol=3e-7, near_h=None, k_band=None):
    """One-shot CubatureSum: returns u (three device tensors)."""
    import torch
    if u is None:
        u = [torch.zeros_like(x) for _ in range(3)]
    cs = CubatureSum(ctx, x, y, z, loops, box, img, k_exact, k_rule, grid_h, near_tol=near_tol,
                     near_h=near_h, k_band=k_band)
    cs.run(x, y, z, *u, accumulate=accumulate)
    cs.free()
    return u

