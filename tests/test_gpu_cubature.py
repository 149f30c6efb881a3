"""The cubature form of wide image sums on the GPU, against the FP64 plain sum
(itself pinned to the oracle by test_gpu_line.py): the 4M-atom sample at +-2 on
every atom, the 200-loop +-4 sample (images +-4, 729 per loop) at random atoms
and atoms next to the loop cuts, the image skips, loop forms the rule does not
represent, and argument errors."""
import ctypes

import numpy as np
import pytest

from loop_kinds import BOX, atoms_near, kind_loops
from disco_amd import backend as B
from disco_amd import cubature as C
from disco_amd import geometry as G

pytestmark = pytest.mark.gpu

TOL = 1e-5          # max |du| against the FP64 plain sum, Angstrom
K_EXACT, NEAR_TOL = 5.0, 3e-7


def _dev(*arrs):
    import torch
    return [torch.from_numpy(np.ascontiguousarray(a)).cuda() for a in arrs]


def _plain64(ctx, x, y, z, loops, box, img):
    import torch
    u = [torch.zeros_like(x) for _ in range(3)]
    p = B.Plan(ctx, loops, box, img, B.PREC_FP64)
    p.run_torch(x, y, z, *u)
    torch.cuda.synchronize()
    p.free()
    return u


def _err(a, b):
    return max(float((s - t).abs().max()) for s, t in zip(a, b))


@pytest.mark.parametrize("k_band", [None, 3.0])
def test_c2_every_atom(hip_ctx, k_band):
    """4,000,000 atoms, 50 loops, images +-2: the cubature form (and with the
    two-ring band from 3 R) within 1e-5 A of the FP64 plain sum on every atom."""
    s = G.config_c2()
    x, y, z = _dev(s.x, s.y, s.z)
    ref = _plain64(hip_ctx, x, y, z, s.loops, s.box, s.img)
    u = C.cubature_sum_torch(hip_ctx, x, y, z, s.loops, s.box, s.img, K_EXACT, near_tol=NEAR_TOL,
                             k_band=k_band)
    e = _err(u, ref)
    print(f"C2 cubature (band {k_band}) vs FP64 plain, all {s.n} atoms: {e:.3e} A")
    assert e < TOL


def _near_cut_atoms(loops, rng, per_loop, dist):
    pts = []
    for lp in loops:
        c = np.array([lp.cx, lp.cy, lp.cz])
        n = np.array(lp.phys[6:9])
        V = np.stack([lp.vx, lp.vy, lp.vz], 1)
        R = np.linalg.norm(V - c, axis=1).max()
        e1 = (V[0] - c) / np.linalg.norm(V[0] - c)
        e2 = np.cross(n, e1)
        r = rng.uniform(0, R + dist, per_loop)
        ph = rng.uniform(0, 2 * np.pi, per_loop)
        h = rng.uniform(-dist, dist, per_loop)
        pts.append(c + r[:, None] * (np.cos(ph)[:, None] * e1 + np.sin(ph)[:, None] * e2)
                   + h[:, None] * n)
    return np.concatenate(pts)


@pytest.mark.parametrize("k_band,octants", [(None, True), (3.0, True), (3.0, False)])
def test_c5_images_4(hip_ctx, k_band, octants):
    """200 loops drawn as the 4M-atom sample's, images +-4: 4,096 random atoms of
    the box and 4,000 atoms within 6 A of the loop cuts (the near slab, where the
    sum is exact in FP64) against the FP64 plain sum over all 729 images; the
    cubature form and its two-ring band from 3 R, with the +-1 images on a grid
    per octant of the atoms or on one grid."""
    s = G.config_c2(seed=20245, nloops=200)
    rng = np.random.default_rng(4)
    L = s.box[0]
    P = np.concatenate([rng.uniform(-L / 2, L / 2, (4096, 3)),
                        _near_cut_atoms(s.loops, rng, 20, 6.0)])
    P = np.clip(P, -L / 2, L / 2 - 1e-9)          # inside the periodic box
    x, y, z = _dev(P[:, 0], P[:, 1], P[:, 2])
    ref = _plain64(hip_ctx, x, y, z, s.loops, s.box, 4)
    cs = C.CubatureSum(hip_ctx, x, y, z, s.loops, s.box, 4, K_EXACT, near_tol=NEAR_TOL,
                       k_band=k_band, octants=octants)
    import torch
    u = [torch.zeros_like(x) for _ in range(3)]
    cs.run(x, y, z, *u)
    torch.cuda.synchronize()
    assert cs.eligible == 200 and cs.skipped > 0
    assert cs.band == (0 if k_band is None else sum(len(lp.vx) >= 12 for lp in s.loops))
    cs.free()
    e = _err(u, ref)
    print(f"C5 cubature (band {k_band}, octants {octants}) vs FP64 plain (+-4), {x.numel()} "
          f"atoms: {e:.3e} A")
    assert e < TOL


def test_image_skip_leaves_exactly_those_images_out(hip_ctx):
    """FP64: a plan with images skipped plus those images summed as translated
    single-image loops equals the plain sum."""
    import torch
    rng = np.random.default_rng(9)
    loops = kind_loops("glide", rng, 3) + kind_loops("prism", rng, 2)
    x, y, z = _dev(*atoms_near(loops, rng, nrand=600, per_loop=60))
    ref = _plain64(hip_ctx, x, y, z, loops, BOX, 1)
    masks = np.zeros(len(loops), np.uint32)
    extra = []
    for l, lp in enumerate(loops):
        for b in rng.choice(27, 9, replace=False):
            i, j, k = b // 9 - 1, (b // 3) % 3 - 1, b % 3 - 1
            masks[l] |= np.uint32(1 << int(b))
            off = (i * BOX[0], j * BOX[1], k * BOX[2])
            extra.append(B.LoopData(phys=lp.phys, line=True, vx=lp.vx + off[0],
                                    vy=lp.vy + off[1], vz=lp.vz + off[2], cx=lp.cx + off[0],
                                    cy=lp.cy + off[1], cz=lp.cz + off[2]))
    p = B.Plan(hip_ctx, loops, BOX, 1, B.PREC_FP64)
    hip_ctx.check(B.lib().disco_hip_plan_set_image_skip(
        p._h, masks.ctypes.data_as(ctypes.POINTER(ctypes.c_uint)), len(loops)))
    u = [torch.zeros_like(x) for _ in range(3)]
    p.run_torch(x, y, z, *u)
    q = B.Plan(hip_ctx, extra, BOX, 0, B.PREC_FP64)
    q.run_torch(x, y, z, *u, accumulate=True)
    torch.cuda.synchronize()
    e = _err(u, ref)
    assert e < 1e-10, e
    # clearing the masks restores the plain sum
    hip_ctx.check(B.lib().disco_hip_plan_set_image_skip(p._h, None, 0))
    v = [torch.zeros_like(x) for _ in range(3)]
    p.run_torch(x, y, z, *v)
    torch.cuda.synchronize()
    assert all(torch.equal(a, b) for a, b in zip(v, ref))
    p.free()
    q.free()


def test_image_skip_per_octant(hip_ctx):
    """FP64: a plan whose atoms leave images out by octant (a random mask per
    loop and octant, planes through the middle of the atoms), plus, for the atoms
    of each octant, the images that octant left out (a plan skipping the
    complement), equals the plain sum; every atom takes its own octant's mask."""
    import torch
    rng = np.random.default_rng(11)
    loops = kind_loops("glide", rng, 2) + kind_loops("mixed", rng, 1) + kind_loops("bent", rng, 1)
    x, y, z = _dev(*atoms_near(loops, rng, nrand=900, per_loop=60))
    ref = _plain64(hip_ctx, x, y, z, loops, BOX, 1)
    split = [float(t.median()) for t in (x, y, z)]
    masks = rng.integers(0, 1 << 27, size=(len(loops), 8), dtype=np.int64).astype(np.uint32)
    p = B.Plan(hip_ctx, loops, BOX, 1, B.PREC_FP64)
    p.set_image_skip(masks, split)
    u = [torch.zeros_like(x) for _ in range(3)]
    p.run_torch(x, y, z, *u)
    code = (x >= split[0]).int() + 2 * (y >= split[1]).int() + 4 * (z >= split[2]).int()
    full = (1 << 27) - 1
    for o in range(8):
        idx = torch.nonzero(code == o).flatten()
        assert idx.numel() > 0
        q = B.Plan(hip_ctx, loops, BOX, 1, B.PREC_FP64)
        q.set_image_skip(np.uint32(full) ^ masks[:, o])
        xo, yo, zo = (t[idx].contiguous() for t in (x, y, z))
        v = [torch.zeros_like(xo) for _ in range(3)]
        q.run_torch(xo, yo, zo, *v)
        for a, b in zip(u, v):
            a.index_add_(0, idx, b)
        q.free()
    torch.cuda.synchronize()
    e = _err(u, ref)
    assert e < 1e-10, e
    # clearing restores the plain sum bit for bit; a NaN plane is refused
    p.set_image_skip(None)
    w = [torch.zeros_like(x) for _ in range(3)]
    p.run_torch(x, y, z, *w)
    torch.cuda.synchronize()
    assert all(torch.equal(a, b) for a, b in zip(w, ref))
    with pytest.raises(RuntimeError):
        p.set_image_skip(masks, [float("nan"), 0.0, 0.0])
    with pytest.raises(ValueError):
        p.set_image_skip(masks)                     # octant masks without their planes
    with pytest.raises(ValueError):
        p.set_image_skip(masks[:, :4], split)       # not eight per loop
    p.free()


def test_image_skip_per_octant_fp32c_every_loop_form(hip_ctx):
    """FP32C far and near passes without the cubature, loops of every form
    (global axes, flat, glide, prismatic): images far from an octant's atoms left
    out for those atoms, plus those images summed for the octant's atoms by plans
    that keep only them, equals the plain FP32C sum (each image's FP32 field is
    the same in both; only the FP64 sums' order differs)."""
    import torch
    rng = np.random.default_rng(12)
    loops = (kind_loops("glide", rng, 2) + kind_loops("mixed", rng, 1) + kind_loops("prism", rng, 1)
             + kind_loops("bent", rng, 1))
    x, y, z = _dev(*atoms_near(loops, rng, nrand=900, per_loop=60))
    ref = [torch.zeros_like(x) for _ in range(3)]
    p0 = B.Plan(hip_ctx, loops, BOX, 1, B.PREC_FP32C)
    p0.run_torch(x, y, z, *ref)
    split = [float(t.median()) for t in (x, y, z)]
    code = (x >= split[0]).int() + 2 * (y >= split[1]).int() + 4 * (z >= split[2]).int()
    P = np.stack([t.cpu().numpy() for t in (x, y, z)], 1)
    oc = code.cpu().numpy()
    masks = np.zeros((len(loops), 8), np.uint32)
    for l, lp in enumerate(loops):
        c = np.array([lp.cx, lp.cy, lp.cz])
        R = np.linalg.norm(np.stack([lp.vx, lp.vy, lp.vz], 1) - c, axis=1).max()
        for o in range(8):
            lo, hi = P[oc == o].min(0), P[oc == o].max(0)
            for b in range(27):
                ci = c + np.array([b // 9 - 1, (b // 3) % 3 - 1, b % 3 - 1]) * np.array(BOX)
                gap = np.maximum(np.maximum(lo - ci, ci - hi), 0.0)
                if np.linalg.norm(gap) > R + 40.0:        # beyond every near slab
                    masks[l, o] |= np.uint32(1 << b)
    assert masks.any() and (masks != (1 << 27) - 1).any()    # some images left out, not all
    p = B.Plan(hip_ctx, loops, BOX, 1, B.PREC_FP32C)
    p.set_image_skip(masks, split)
    u = [torch.zeros_like(x) for _ in range(3)]
    p.run_torch(x, y, z, *u)
    full = (1 << 27) - 1
    for o in range(8):
        idx = torch.nonzero(code == o).flatten()
        q = B.Plan(hip_ctx, loops, BOX, 1, B.PREC_FP32C)
        q.set_image_skip(np.uint32(full) ^ masks[:, o])
        xo, yo, zo = (t[idx].contiguous() for t in (x, y, z))
        v = [torch.zeros_like(xo) for _ in range(3)]
        q.run_torch(xo, yo, zo, *v)
        for a, b in zip(u, v):
            a.index_add_(0, idx, b)
        q.free()
    torch.cuda.synchronize()
    e = _err(u, ref)
    assert e < 1e-11, e
    p.free()
    p0.free()


@pytest.mark.parametrize("k_rule,k_band", [(None, None), (8.0, None), (None, 3.0), (8.0, 3.0)])
def test_loop_forms_and_window(hip_ctx, k_rule, k_band):
    """Glide, prismatic and general-b circles (the rule represents them), loops
    off their plane (summed in global axes, exact) and rectangles (anisotropic:
    exact), images +-3 in a small box; the hard switch and a blend 5 R .. 8 R,
    each with and without the two-ring band from 3 R."""
    import torch
    rng = np.random.default_rng(21)
    loops = (kind_loops("glide", rng, 2) + kind_loops("prism", rng, 2) +
             kind_loops("mixed", rng, 2) + kind_loops("bent", rng, 1))
    V = np.array([[12, 7, 0], [-12, 7, 0], [-12, -7, 0], [12, -7, 0]], float) + [70, 80, 90]
    loops.append(B.LoopData(phys=(2.5, 2.5, 2.5, 1, 0, 0, 0, 0, 1, 0.0), line=True,
                            vx=V[:, 0].copy(), vy=V[:, 1].copy(), vz=V[:, 2].copy(), cx=70.0,
                            cy=80.0, cz=90.0))
    x, y, z = _dev(*atoms_near(loops, rng, nrand=3000, per_loop=100))
    ref = _plain64(hip_ctx, x, y, z, loops, BOX, 3)
    cs = C.CubatureSum(hip_ctx, x, y, z, loops, BOX, 3, K_EXACT, k_rule, near_tol=NEAR_TOL,
                       k_band=k_band)
    assert cs.eligible == 6                       # not the bent loop, not the rectangle
    assert cs.band == (0 if k_band is None else 6)
    u = [torch.zeros_like(x) for _ in range(3)]
    cs.run(x, y, z, *u)
    torch.cuda.synchronize()
    cs.free()
    e = _err(u, ref)
    print(f"loop forms, k_rule {k_rule}, band {k_band}: {e:.3e} A")
    assert e < TOL


def test_cubature_off_is_the_plain_fp32c_pass_and_errors(hip_ctx):
    import torch
    s = G.config_c1()
    idx = np.arange(0, s.n, 17)
    x, y, z = _dev(s.x[idx], s.y[idx], s.z[idx])
    p = B.Plan(hip_ctx, s.loops, s.box, 1, B.PREC_FP32C)
    a = [torch.zeros_like(x) for _ in range(3)]
    p.run_torch(x, y, z, *a)
    assert p.set_cubature(5.0) == 2
    assert p.set_cubature(0.0) == 0                # off again
    b = [torch.zeros_like(x) for _ in range(3)]
    p.run_torch(x, y, z, *b)
    torch.cuda.synchronize()
    assert all(torch.equal(s_, t) for s_, t in zip(a, b))
    for bad in ((1.5, 2.0), (5.0, 4.0), (5.0, float("inf"))):
        with pytest.raises(B.DiscoError):
            p.set_cubature(*bad)
    # the band form: off by default and when switched off again
    assert p.set_band(3.0) == 2 and p.set_band(0.0) == 0
    c = [torch.zeros_like(x) for _ in range(3)]
    p.run_torch(x, y, z, *c)
    torch.cuda.synchronize()
    assert all(torch.equal(s_, t) for s_, t in zip(a, c))
    for bad in (2.0, float("inf"), float("nan")):
        with pytest.raises(B.DiscoError):
            p.set_band(bad)
    p.free()
    p64 = B.Plan(hip_ctx, s.loops, s.box, 1, B.PREC_FP64)
    with pytest.raises(B.DiscoError):
        p64.set_cubature(5.0)
    with pytest.raises(B.DiscoError):
        p64.set_band(3.0)
    p64.free()


def _band_loop(kind, R, nseg, c, rng):
    """A regular polygon of the given kind whose vertex 0 lies 1e-7 R farther out
    than the others (so the packer's frame axis e1 points at it; the moments
    stay isotropic far below the rule's eligibility bound)."""
    from loop_kinds import _frame
    n, e1, e2 = _frame(rng)
    t = 2 * np.pi * np.arange(nseg) / nseg
    rr = np.full(nseg, R)
    rr[0] *= 1 + 1e-7
    V = c + rr[:, None] * (np.cos(t)[:, None] * e1 + np.sin(t)[:, None] * e2)
    b = {"glide": 2.5 * (0.6 * e1 + 0.8 * e2), "prism": 2.5 * n,
         "mixed": 1.5 * n + 2.0 * e1 - 0.7 * e2}[kind]
    nb = float(np.linalg.norm(b))
    lp = B.LoopData(phys=(nb, nb, nb, *(b / nb), *n, 0.0), line=True, vx=V[:, 0].copy(),
                    vy=V[:, 1].copy(), vz=V[:, 2].copy(), cx=c[0], cy=c[1], cz=c[2])
    return lp, n, e1, e2, b


@pytest.mark.parametrize("kind", ["glide", "prism", "mixed"])
def test_band_kernel_is_the_two_ring_cubature(hip_ctx, kind):
    """One loop image (images 0): lanes between 2.5 R and 5 R take the two-ring
    cubature -- equal to the FP64 oracle's sum of its 25 points as loopons to
    FP32 accuracy, and off the exact line integral by the rule's own error
    (about 1e-6 A just beyond 2.5 R) --, lanes nearer the loop the exact field."""
    import torch
    import oracle
    rng = np.random.default_rng({"glide": 31, "prism": 32, "mixed": 33}[kind])
    R, nseg = {"glide": (30.0, 48), "prism": (20.0, 24), "mixed": (25.0, 36)}[kind]
    c = np.array([500.0, 600.0, 700.0])
    lp, n, e1, e2, b = _band_loop(kind, R, nseg, c, rng)
    box = (2000.0, 2000.0, 2000.0)
    d = rng.normal(size=(2 * 1500, 3))
    d /= np.linalg.norm(d, axis=1)[:, None]
    r = np.concatenate([rng.uniform(1.6, 2.45, 1500), rng.uniform(2.55, 4.9, 1500)]) * R
    P = c + r[:, None] * d
    x, y, z = _dev(P[:, 0], P[:, 1], P[:, 2])
    p = B.Plan(hip_ctx, [lp], box, 0, B.PREC_FP32C)
    assert p.set_cubature(K_EXACT) == 1 and p.set_band(2.5) == 1
    u = [torch.zeros_like(x) for _ in range(3)]
    p.run_torch(x, y, z, *u)
    torch.cuda.synchronize()
    p.free()
    got = np.stack([t.cpu().numpy() for t in u], 1)
    ex = np.stack(oracle.line_sum(P[:, 0], P[:, 1], P[:, 2], [lp], box, 0), 1)
    info = B.line_band([lp])[0]
    nb = float(np.linalg.norm(b))
    pts = []
    for rho, w, npt, ph in ((0.0, info[1], 1, 0.0), (info[4], info[2], 12, 0.0),
                            (info[5], info[3], 12, np.pi / 12)):
        th = ph + 2 * np.pi * np.arange(npt) / npt
        Q = c + rho * (np.cos(th)[:, None] * e1 + np.sin(th)[:, None] * e2)
        pts.append(B.LoopData(phys=(nb, nb, nb, *(b / nb), *n, w), line=False, x=Q[:, 0].copy(),
                              y=Q[:, 1].copy(), z=Q[:, 2].copy()))
    rule = np.stack(oracle.loopon_sum(P[:, 0], P[:, 1], P[:, 2], pts, box, 0), 1)
    inner, outer = slice(0, 1500), slice(1500, None)
    e_in = np.abs(got[inner] - ex[inner]).max()
    e_rule = np.abs(got[outer] - rule[outer]).max()
    e_out = np.abs(got[outer] - ex[outer]).max()
    scale = np.abs(rule[outer]).max()
    print(f"{kind}: exact lanes vs line {e_in:.2e}, band lanes vs rule {e_rule:.2e} "
          f"(field up to {scale:.2e}), vs line {e_out:.2e} A")
    assert e_in < 5e-7
    assert e_rule < 1e-6 * scale + 1e-9
    assert 2e-7 < e_out < 3e-6


def test_polylines_off_the_plane_in_cubature_form(hip_ctx):
    """Closed polylines that are not flat discs (saddle, cone, figure eights,
    star, apex off the plane: polylines.py) beside circles, images +-3: their
    +-1 images stay exact in the line plan and their far shell is Radon's rule
    on each fan triangle (surrogate.fan_loopons); within 1e-5 A of the FP64
    plain sum."""
    import torch
    import polylines as PL
    rng = np.random.default_rng(23)
    loops = PL.loops(rng) + kind_loops("mixed", rng, 2)
    x, y, z = _dev(*PL.atoms(loops, rng, nrand=3000, per_loop=100))
    ref = _plain64(hip_ctx, x, y, z, loops, PL.BOX, 3)
    cs = C.CubatureSum(hip_ctx, x, y, z, loops, PL.BOX, 3, K_EXACT, near_tol=NEAR_TOL)
    assert cs.eligible == 2                       # the circles only
    u = [torch.zeros_like(x) for _ in range(3)]
    cs.run(x, y, z, *u)
    torch.cuda.synchronize()
    cs.free()
    e = _err(u, ref)
    print(f"polylines in cubature form, +-3: {e:.3e} A")
    assert e < TOL
