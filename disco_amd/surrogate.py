"""Learned far image shell of the periodic sum (host side, on the C ABI).

The field of the loops over images (i, j, k) with max(|i|,|j|,|k|) in (1, img]
is the far shell.  Its surrogate is

    far(p) = sum_loops [ base_far(p) + scale * MLP(features(p, loop)) ]

* ``base_far`` -- the loop's cut replaced by a few weighted loopons (the Kroupa
  term, Code/kroupa.hpp:52-73), summed exactly over the far shell by the
  loopon kernels (``Plan`` with ``inner = 1``).  Two rules:
  ``"dipole"``: one loopon of the loop's area at its centre (exact to first
  order in R/r); ``"rule7"``: the centre plus a ring of six, with weights and
  ring radius matching the cut's area and its second and fourth in-plane
  moments -- exact through degree 5 for cuts whose moments are isotropic up
  to rank 5 (circles, regular hexagons), so its far-shell error is
  O((R/r)^6);
* ``MLP`` -- 14 -> 128 -> 128 -> 128 -> 3 (hard swish), trained by the library's own
  GPU kernels (``disco_hip_surrogate_*``) on exact far shells generated by the
  line kernels, with ``scale = |b| (R / L)^4``.

The near shell (+-1) is always the exact line integral.  ``far_shell`` checks
the surrogate against the exact far shell on a validation subset of atoms and
falls back to the exact sum when the error exceeds the tolerance -- the plain
sum is both the fallback and the measure of its error.
"""
from __future__ import annotations

import ctypes
import math
import time
from dataclasses import dataclass

import numpy as np

from . import backend as B
from . import geometry as G

SHAPES = {"circle": 0, "rect": 1, "parallelogram": 2, "polygon": 3}
NIN, NOUT = 14, 3


class _SLoop(ctypes.Structure):
    _fields_ = [("c", ctypes.c_double * 3), ("n", ctypes.c_double * 3), ("b", ctypes.c_double * 3),
                ("radius", ctypes.c_double), ("shape", ctypes.c_int)]


_fp = ctypes.POINTER(ctypes.c_float)
_dp = ctypes.POINTER(ctypes.c_double)
_vp = ctypes.c_void_p
_bound = False


def _lib():
    global _bound
    L = B.lib()
    if not _bound:
        L.disco_hip_surrogate_num_params.restype = ctypes.c_int
        L.disco_hip_surrogate_features.argtypes = [_dp, ctypes.c_int, ctypes.POINTER(_SLoop), _dp, _fp]
        L.disco_hip_surrogate_scale.argtypes = [ctypes.POINTER(_SLoop), _dp]
        L.disco_hip_surrogate_scale.restype = ctypes.c_double
        L.disco_hip_surrogate_create.argtypes = [_vp, ctypes.c_uint, ctypes.POINTER(_vp)]
        L.disco_hip_surrogate_free.argtypes = [_vp]
        L.disco_hip_surrogate_get_params.argtypes = [_vp, _fp]
        L.disco_hip_surrogate_set_params.argtypes = [_vp, _fp]
        L.disco_hip_surrogate_forward.argtypes = [_vp, _fp, ctypes.c_int, _fp]
        L.disco_hip_surrogate_set_gain.argtypes = [_vp, ctypes.c_double]
        L.disco_hip_surrogate_get_gain.argtypes = [_vp]
        L.disco_hip_surrogate_get_gain.restype = ctypes.c_double
        L.disco_hip_surrogate_loss_grad.argtypes = [_vp, _fp, _fp, ctypes.c_int, _fp, _fp]
        L.disco_hip_surrogate_train.argtypes = [_vp, _fp, _fp, ctypes.c_int, ctypes.c_int,
                                                ctypes.c_int, ctypes.c_float, ctypes.c_uint, _fp]
        L.disco_hip_surrogate_apply.argtypes = [_vp, ctypes.c_int, _vp, _vp, _vp, _vp, _vp, _vp,
                                                ctypes.POINTER(_SLoop), ctypes.c_int, _dp,
                                                ctypes.c_int, _vp]
        L.disco_hip_plan_set_inner.argtypes = [_vp, ctypes.c_int]
        L.disco_hip_plan_set_split.argtypes = [_vp, ctypes.c_int]
        L.disco_hip_batch_set_inner.argtypes = [_vp, ctypes.c_int]
        _bound = True
    return L


def num_params() -> int:
    return _lib().disco_hip_surrogate_num_params()


# ------------------------------------------------------------------ loop descriptions
@dataclass
class LoopShapeInfo:
    """What the surrogate knows of a loop: centre, normal, b, radius, shape."""
    c: tuple
    n: tuple
    b: tuple
    radius: float
    shape: str

    def struct(self):
        s = _SLoop()
        s.c[:] = list(self.c)
        s.n[:] = list(self.n)
        s.b[:] = list(self.b)
        s.radius = float(self.radius)
        s.shape = SHAPES[self.shape]
        return s


def shape_info(lp: B.LoopData, radius: float, shape: str) -> LoopShapeInfo:
    """Surrogate description of a line loop (normal from its vector area)."""
    V = np.stack([lp.vx, lp.vy, lp.vz], 1) - np.array([lp.cx, lp.cy, lp.cz])
    N = np.cross(np.roll(V, 1, axis=0), V).sum(0)
    n = N / np.linalg.norm(N)
    return LoopShapeInfo((lp.cx, lp.cy, lp.cz), tuple(n), tuple(lp.burgers()), radius, shape)


def dipole_loopon(lp: B.LoopData) -> B.LoopData:
    """The loop as one loopon of its area at its centre (same b, normal from
    the polyline's vector area, m = b / |b|)."""
    V = np.stack([lp.vx, lp.vy, lp.vz], 1) - np.array([lp.cx, lp.cy, lp.cz])
    N = np.cross(np.roll(V, 1, axis=0), V).sum(0)
    area = 0.5 * np.linalg.norm(N)
    n = N / np.linalg.norm(N)
    b = np.array(lp.burgers())
    bm = float(np.linalg.norm(b))
    m = b / bm
    return B.LoopData(phys=(bm, bm, bm, *m, *n, area), line=False, x=np.array([lp.cx]),
                      y=np.array([lp.cy]), z=np.array([lp.cz]))


# Dunavant's 7-point degree-5 rule on a triangle (barycentric points, weights)
_DW = np.array([0.225, 0.1323941527, 0.1323941527, 0.1323941527,
                0.1259391805, 0.1259391805, 0.1259391805])
_A1, _B1, _A2, _B2 = 0.0597158717, 0.4701420641, 0.7974269853, 0.1012865073
_DB = np.array([[1 / 3, 1 / 3, 1 / 3], [_A1, _B1, _B1], [_B1, _A1, _B1], [_B1, _B1, _A1],
                [_A2, _B2, _B2], [_B2, _A2, _B2], [_B2, _B2, _A2]])


def _cross_rows(a, b):
    """Row-wise cross products of (k, 3) arrays (np.cross without its per-call
    axis handling, which dominates for the few rows of a loop)."""
    return np.stack([a[:, 1] * b[:, 2] - a[:, 2] * b[:, 1],
                     a[:, 2] * b[:, 0] - a[:, 0] * b[:, 2],
                     a[:, 0] * b[:, 1] - a[:, 1] * b[:, 0]], 1)


def area_moments(lp: B.LoopData):
    """Area A, normal n and in-plane moments I2 = int r^2 dA, I4 = int r^4 dA of
    the flat cut about its centre (fan triangles, exact degree-5 rule)."""
    c = np.array([lp.cx, lp.cy, lp.cz])
    V = np.stack([lp.vx, lp.vy, lp.vz], 1) - c
    # every fan triangle (centre, V_j, V_j+1) at once: its area and the rule's
    # points (the centre is the origin here)
    P0, P1 = V, np.roll(V, -1, axis=0)
    X = _cross_rows(P0, P1)
    N = X.sum(0)
    A = 0.5 * float(np.linalg.norm(N))
    ar = 0.5 * np.sqrt((X * X).sum(1))                                      # (nv,)
    pts = _DB[:, 1, None, None] * P0[None] + _DB[:, 2, None, None] * P1[None]   # (7, nv, 3)
    r2 = (pts ** 2).sum(2)                                                  # (7, nv)
    I2 = float(ar @ (_DW @ r2))
    I4 = float(ar @ (_DW @ (r2 * r2)))
    return A, N / np.linalg.norm(N), I2, I4


def rule7_loopons(lp: B.LoopData):
    """The cut as 7 weighted loopons: w0 at the centre, w1 at six points on a
    ring of radius rho, with w0 + 6 w1 = A, 6 w1 rho^2 = I2, 6 w1 rho^4 = I4
    (for a disc: rho^2 = 2R^2/3, w0 = A/4, w1 = A/8 -- Radon's degree-5 rule).
    Returns two loopon LoopData (centre, ring) sharing b, m, n."""
    return rule7_loopons_many([lp])


# Radon's 7-point degree-5 rule on a triangle, exact: barycentric classes
# (1/3, 1/3, 1/3), the 3 permutations of (a1, b1, b1), of (a2, b2, b2)
_S15 = float(np.sqrt(15.0))
_RADON = ((9 / 40, ((1 / 3, 1 / 3, 1 / 3),)),
          ((155 + _S15) / 1200, tuple(np.roll([(9 - 2 * _S15) / 21, (6 + _S15) / 21,
                                               (6 + _S15) / 21], k) for k in range(3))),
          ((155 - _S15) / 1200, tuple(np.roll([(9 + 2 * _S15) / 21, (6 - _S15) / 21,
                                               (6 - _S15) / 21], k) for k in range(3))))


def ring_rule_ok(lp: B.LoopData, tol=1e-7) -> bool:
    """The centred ring rules stand for the cut only when it is a flat region
    about the apex: every vertex within tol R of the plane of the vector area,
    every fan triangle (c, V_k, V_k+1) turning the way that area does
    (star-shaped about c)."""
    c = np.array([lp.cx, lp.cy, lp.cz])
    V = np.stack([lp.vx, lp.vy, lp.vz], 1) - c
    X = _cross_rows(V, np.roll(V, -1, axis=0))
    N = X.sum(0)
    nN = float(np.linalg.norm(N))
    if not nN > 0.0:
        return False
    n = N / nN
    R = float(np.sqrt((V * V).sum(1)).max())
    return bool(np.abs(V @ n).max() <= tol * R and np.all(X @ n > 0.0))


def fan_loopons(lp: B.LoopData):
    """The cut of any closed polyline -- the fan of triangles (c, V_k, V_k+1)
    that Code/barnett.hpp:44-107 takes -- as Radon's 7-point rule on each fan
    triangle with that triangle's own normal (V_k - c) x (V_k+1 - c) / |...|
    and area weights: three loopon LoopData per triangle, one per weight class.
    Exact through degree 5 on every triangle, for loops off their plane, with
    lobes of opposite sense or winding more than once."""
    c = np.array([lp.cx, lp.cy, lp.cz])
    V = np.stack([lp.vx, lp.vy, lp.vz], 1) - c
    V1 = np.roll(V, -1, axis=0)
    X = _cross_rows(V, V1)
    b = np.array(lp.burgers())
    bm = float(np.linalg.norm(b))
    m = b / bm
    out = []
    for k in range(len(V)):
        two_a = float(np.linalg.norm(X[k]))
        if not two_a > 0.0:
            continue
        nk = X[k] / two_a
        for w, bary in _RADON:
            P = np.array([c + bb[1] * V[k] + bb[2] * V1[k] for bb in bary])
            out.append(B.LoopData(phys=(bm, bm, bm, *m, *nk, w * 0.5 * two_a), line=False,
                                  x=P[:, 0].copy(), y=P[:, 1].copy(), z=P[:, 2].copy()))
    return out


def rule_loopons_per_loop(loops, ring=8):
    """Per loop, the loopons of its cut: [centre, ring] of the centred ring
    rule (rule7_loopons_many) where ring_rule_ok, else fan_loopons."""
    ok = [ring_rule_ok(lp) for lp in loops]
    idx = [k for k in range(len(loops)) if ok[k]]
    flat = rule7_loopons_many([loops[k] for k in idx], ring) if idx else []
    out = [None] * len(loops)
    for j, k in enumerate(idx):
        out[k] = flat[2 * j:2 * j + 2]
    for k in range(len(loops)):
        if not ok[k]:
            out[k] = fan_loopons(loops[k])
    return out


def rule7_loopons_many(loops, ring=6):
    """rule7_loopons of many loops, computed together for the loops of each
    vertex count: [centre, ring] per loop, in the loops' order.  ring = 8 gives
    the 9-point rule (w0 + 8 w1 = A, 8 w1 rho^2 = I2, 8 w1 rho^4 = I4; ring
    angles j pi/4), whose first angular error is at degree 8 instead of 6.
    For flat cuts about the apex (ring_rule_ok); base_loopons and
    rule_loopons_per_loop take fan_loopons for the others."""
    out = [None] * (2 * len(loops))
    by_nv = {}
    for i, lp in enumerate(loops):
        by_nv.setdefault(len(lp.vx), []).append(i)
    th = 2.0 * np.pi * np.arange(ring) / ring
    cs, sn = np.cos(th), np.sin(th)
    for nv, idx in by_nv.items():
        C = np.array([[loops[i].cx, loops[i].cy, loops[i].cz] for i in idx])        # (L, 3)
        V = np.stack([np.stack([loops[i].vx, loops[i].vy, loops[i].vz], 1)
                      for i in idx]) - C[:, None, :]                                  # (L, nv, 3)
        P1 = np.roll(V, -1, axis=1)
        X = np.stack([V[..., 1] * P1[..., 2] - V[..., 2] * P1[..., 1],
                      V[..., 2] * P1[..., 0] - V[..., 0] * P1[..., 2],
                      V[..., 0] * P1[..., 1] - V[..., 1] * P1[..., 0]], -1)           # (L, nv, 3)
        N = X.sum(1)
        nN = np.sqrt((N * N).sum(1))
        A = 0.5 * nN
        n = N / nN[:, None]
        ar = 0.5 * np.sqrt((X * X).sum(2))                                            # (L, nv)
        pts = (_DB[None, :, 1, None, None] * V[:, None] +
               _DB[None, :, 2, None, None] * P1[:, None])                             # (L, 7, nv, 3)
        r2 = (pts * pts).sum(3)                                                       # (L, 7, nv)
        I2 = np.einsum("lj,q,lqj->l", ar, _DW, r2)
        I4 = np.einsum("lj,q,lqj->l", ar, _DW, r2 * r2)
        rho = np.sqrt(I4 / I2)
        w1 = I2 * I2 / (ring * I4)
        w0 = A - ring * w1
        V0 = V[:, 0]
        e1 = V0 - (V0 * n).sum(1)[:, None] * n
        e1 /= np.sqrt((e1 * e1).sum(1))[:, None]
        e2 = np.stack([n[:, 1] * e1[:, 2] - n[:, 2] * e1[:, 1],
                       n[:, 2] * e1[:, 0] - n[:, 0] * e1[:, 2],
                       n[:, 0] * e1[:, 1] - n[:, 1] * e1[:, 0]], 1)
        pts_r = C[:, None] + rho[:, None, None] * (cs[None, :, None] * e1[:, None] +
                                                  sn[None, :, None] * e2[:, None])    # (L, ring, 3)
        for k, i in enumerate(idx):
            b = np.array(loops[i].burgers())
            bm = float(np.linalg.norm(b))
            m = b / bm
            nk = n[k]
            out[2 * i] = B.LoopData(phys=(bm, bm, bm, *m, *nk, float(w0[k])), line=False,
                                    x=C[k, :1].copy(), y=C[k, 1:2].copy(), z=C[k, 2:3].copy())
            out[2 * i + 1] = B.LoopData(phys=(bm, bm, bm, *m, *nk, float(w1[k])), line=False,
                                        x=pts_r[k, :, 0].copy(), y=pts_r[k, :, 1].copy(),
                                        z=pts_r[k, :, 2].copy())
    return out


def base_loopons(loops, kind: str):
    if kind == "dipole":
        return [dipole_loopon(l) for l in loops]
    if kind in ("rule7", "rule9"):
        per = rule_loopons_per_loop(loops, ring=6 if kind == "rule7" else 8)
        return [part for parts in per for part in parts]
    raise ValueError(f"unknown baseline {kind!r}")


def features(points, info: LoopShapeInfo, box) -> np.ndarray:
    """Network inputs of atoms `points` (n x 3) for one loop (n x 14 float32)."""
    p = np.ascontiguousarray(points, np.float64)
    out = np.zeros((p.shape[0], NIN), np.float32)
    L = (ctypes.c_double * 3)(*[float(t) for t in box])
    s = info.struct()
    rc = _lib().disco_hip_surrogate_features(p.ctypes.data_as(_dp), p.shape[0], ctypes.byref(s),
                                             L, out.ctypes.data_as(_fp))
    if rc:
        raise B.DiscoError(rc, "features")
    return out


def scale(info: LoopShapeInfo, box) -> float:
    L = (ctypes.c_double * 3)(*[float(t) for t in box])
    s = info.struct()
    return _lib().disco_hip_surrogate_scale(ctypes.byref(s), L)


# ------------------------------------------------------------------ the network
class Surrogate:
    """The MLP's parameters and optimizer state on the device."""

    def __init__(self, ctx: B.Context, seed: int = 1):
        self.ctx = ctx
        self._h = _vp()
        ctx.check(_lib().disco_hip_surrogate_create(ctx.handle, seed, ctypes.byref(self._h)))

    def params(self) -> np.ndarray:
        out = np.zeros(num_params(), np.float32)
        self.ctx.check(_lib().disco_hip_surrogate_get_params(self._h, out.ctypes.data_as(_fp)))
        return out

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        assert p.size == num_params()
        self.ctx.check(_lib().disco_hip_surrogate_set_params(self._h, p.ctypes.data_as(_fp)))

    def forward(self, feats) -> np.ndarray:
        f = np.ascontiguousarray(feats, np.float32)
        out = np.zeros((f.shape[0], NOUT), np.float32)
        self.ctx.check(_lib().disco_hip_surrogate_forward(self._h, f.ctypes.data_as(_fp),
                                                          f.shape[0], out.ctypes.data_as(_fp)))
        return out

    def loss_grad(self, feats, targets):
        f = np.ascontiguousarray(feats, np.float32)
        t = np.ascontiguousarray(targets, np.float32)
        loss = np.zeros(1, np.float32)
        g = np.zeros(num_params(), np.float32)
        self.ctx.check(_lib().disco_hip_surrogate_loss_grad(self._h, f.ctypes.data_as(_fp),
                                                            t.ctypes.data_as(_fp), f.shape[0],
                                                            loss.ctypes.data_as(_fp),
                                                            g.ctypes.data_as(_fp)))
        return float(loss[0]), g

    @property
    def gain(self) -> float:
        return _lib().disco_hip_surrogate_get_gain(self._h)

    def set_gain(self, g: float):
        self.ctx.check(_lib().disco_hip_surrogate_set_gain(self._h, float(g)))

    def train(self, feats, targets, epochs=20, batch=4096, lr=2e-3, seed=7,
              normalize=True) -> float:
        """Adam on the GPU.  With `normalize`, the network learns targets / RMS
        and the surrogate's gain is set to that RMS."""
        f = np.ascontiguousarray(feats, np.float32)
        t = np.ascontiguousarray(targets, np.float64)
        g = float(np.sqrt((t ** 2).mean())) if normalize else 1.0
        if not g > 0:
            g = 1.0
        t = np.ascontiguousarray(t / g, np.float32)
        self.set_gain(g)
        loss = np.zeros(1, np.float32)
        self.ctx.check(_lib().disco_hip_surrogate_train(self._h, f.ctypes.data_as(_fp),
                                                        t.ctypes.data_as(_fp), f.shape[0],
                                                        epochs, batch, lr, seed,
                                                        loss.ctypes.data_as(_fp)))
        return float(loss[0])

    def apply_torch(self, x, y, z, ux, uy, uz, infos, box, chunk_rows=0):
        """u += sum over loops of scale * MLP (device atoms as torch tensors)."""
        import torch
        arr = (_SLoop * max(1, len(infos)))(*[i.struct() for i in infos])
        L = (ctypes.c_double * 3)(*[float(t) for t in box])
        stream = torch.cuda.current_stream().cuda_stream
        self.ctx.check(_lib().disco_hip_surrogate_apply(
            self._h, x.numel(), x.data_ptr(), y.data_ptr(), z.data_ptr(), ux.data_ptr(),
            uy.data_ptr(), uz.data_ptr(), arr, len(infos), L, chunk_rows, stream))

    def free(self):
        if self._h:
            _lib().disco_hip_surrogate_free(self._h)
            self._h = _vp()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


# ------------------------------------------------------------------ exact far shells
def set_inner(plan: B.Plan, inner: int):
    plan.ctx.check(_lib().disco_hip_plan_set_inner(plan._h, int(inner)))


def set_split(plan: B.Plan, groups: int):
    """Loop groups per launch of a loopon plan evaluated at few atoms."""
    plan.ctx.check(_lib().disco_hip_plan_set_split(plan._h, int(groups)))


def exact_far_shell_torch(ctx, x, y, z, loops, box, img, prec=B.PREC_FP32C):
    """Exact far shell (images with max index in (1, img]) of line loops."""
    import torch
    u = [torch.zeros_like(x) for _ in range(3)]
    plan = B.Plan(ctx, loops, box, img, prec, kind="line")
    set_inner(plan, 1)
    plan.run_torch(x, y, z, *u)
    plan.free()
    return u


def grid_nodes(x, y, z, h, margin=1):
    """Uniform grid of spacing h covering the device atoms with `margin` nodes of
    margin on each side (1 keeps the 4-node stencils of the interpolation
    inside, 2 the 6-node ones): returns the origin, the node counts (nx, ny,
    nz) and the node coordinates (three device tensors of nz * ny * nx, x
    fastest)."""
    import torch
    lo = [float(t.min()) for t in (x, y, z)]
    hi = [float(t.max()) for t in (x, y, z)]
    dims = [int(math.ceil((b - a) / h)) + 2 * margin + 2 for a, b in zip(lo, hi)]
    origin = [a - margin * h for a in lo]
    axes = [origin[k] + h * torch.arange(dims[k], dtype=torch.float64, device=x.device)
            for k in range(3)]
    gz, gy, gx = torch.meshgrid(axes[2], axes[1], axes[0], indexing="ij")
    return origin, dims, (gx.reshape(-1).contiguous(), gy.reshape(-1).contiguous(),
                          gz.reshape(-1).contiguous())


def base_far_shell_torch(ctx, x, y, z, loops, box, img, kind="rule7", u=None,
                         accumulate=False, prec=B.PREC_FP32C, grid_h=None, grid_groups=None):
    """Far shell of the loops' loopon rule (`kind`), summed by the loopon kernels:
    at every atom, or (grid_h set) at the nodes of a grid of that spacing over
    the atoms and interpolated to them -- the far shell is smooth on the scale of
    the box, so a grid of L/32 carries it to ~1e-9 A."""
    import torch
    if u is None:
        u = [torch.zeros_like(x) for _ in range(3)]
    base = base_loopons(loops, kind)
    if grid_h is None:
        plan = B.Plan(ctx, base, box, img, prec, kind="loopon")
        set_inner(plan, 1)
        plan.run_torch(x, y, z, *u, accumulate=accumulate)
        plan.free()
        return u
    origin, dims, nodes = grid_nodes(x, y, z, grid_h)
    # a grid holds few nodes (tens of thousands) for the GPU's 256 CUs: the loops
    # are split into contiguous groups, each summed by its own workgroups of one
    # launch, then added in group order (about 8,192 workgroups in all: 35k waves)
    tiles = (nodes[0].numel() + 255) // 256
    want = grid_groups if grid_groups else -(-8192 // tiles)
    groups = max(1, min(want, len(base), 256))
    plan = B.Plan(ctx, base, box, img, prec, kind="loopon")
    set_inner(plan, 1)
    set_split(plan, groups)
    gu = [torch.zeros_like(nodes[0]) for _ in range(3)]
    plan.run_torch(*nodes, *gu)
    shape = (dims[2], dims[1], dims[0])
    B.grid_interp_torch(ctx, x, y, z, *u, [t.view(shape) for t in gu], origin, grid_h,
                        accumulate=accumulate)
    torch.cuda.synchronize()            # the plan's tables are freed below
    plan.free()
    return u


def dipole_far_shell_torch(ctx, x, y, z, loops, box, img, u=None, accumulate=False):
    """Far shell of the loops' dipole terms, FP64."""
    return base_far_shell_torch(ctx, x, y, z, loops, box, img, "dipole", u, accumulate,
                                B.PREC_FP64)


# ------------------------------------------------------------------ training data
def random_fcc_loops(rng, nloops, a0=3.615, cells=100):
    """Loops drawn as in the 4M-atom workload: radius U(10, 60), centre U(box)
    snapped, n one of the 4 {111}, b one of the 3 in-plane a/2<110>; 60 %
    circles of 48 segments, 40 % hexagons.  Returns (LoopData, info) pairs."""
    loops, geo = G._fcc_loop_set(rng, nloops, a0, cells)
    out = []
    for lp, (kind, R, c, n, b) in zip(loops, geo):
        info = shape_info(lp, R, "circle" if kind == 0 else "polygon")
        out.append((lp, info))
    return out


def make_training_set(ctx, nloops=2000, atoms_per_loop=500, img=4, seed=11, a0=3.615, cells=100,
                      baseline="rule7"):
    """Exact far shells of random (atom, loop) pairs, minus the baseline rule's
    far shell, divided by each loop's scale: (features n x 14, targets n x 3,
    scales n, exact n x 3, baseline n x 3).  Computed on the GPU: the exact
    shells in one batched launch (each loop one sample), in FP64."""
    import torch
    rng = np.random.default_rng(seed)
    L = a0 * cells
    box = (L, L, L)
    pairs = random_fcc_loops(rng, nloops, a0, cells)
    samples, feats, scales, pts = [], [], [], []
    for lp, info in pairs:
        p = rng.uniform(-L / 2, L / 2, size=(atoms_per_loop, 3))
        pts.append(p)
        samples.append(G.Sample("train", p[:, 0].copy(), p[:, 1].copy(), p[:, 2].copy(), [lp],
                                box, img))
        feats.append(features(p, info, box))
        scales.append(np.full(atoms_per_loop, scale(info, box)))
    P = np.concatenate(pts)
    dev = torch.device("cuda")
    x, y, z = (torch.from_numpy(np.ascontiguousarray(P[:, k])).to(dev) for k in range(3))
    u = [torch.zeros_like(x) for _ in range(3)]
    bt = B.Batch(ctx, samples, img, B.PREC_FP64)
    ctx.check(_lib().disco_hip_batch_set_inner(bt._h, 1))
    bt.run_torch(x, y, z, *u)
    torch.cuda.synchronize()
    exact = np.stack([t.cpu().numpy() for t in u], 1)
    bt.free()
    # baseline far shell, loop by loop on its own atoms
    dip = np.zeros_like(exact)
    for k, (lp, _info) in enumerate(pairs):
        s = slice(k * atoms_per_loop, (k + 1) * atoms_per_loop)
        ud = base_far_shell_torch(ctx, x[s].contiguous(), y[s].contiguous(),
                                  z[s].contiguous(), [lp], box, img, baseline, prec=B.PREC_FP64)
        dip[s] = np.stack([t.cpu().numpy() for t in ud], 1)
    S = np.concatenate(scales)
    targets = ((exact - dip) / S[:, None]).astype(np.float32)
    return np.concatenate(feats), targets, S, exact, dip


# ------------------------------------------------------------------ the far-shell pipeline
@dataclass
class FarShellReport:
    used_surrogate: bool
    max_err: float          # surrogate vs exact far shell on the validation atoms (A)
    base_only_err: float    # the same without the network (A)
    baseline: str
    n_valid: int
    seconds_surrogate: float
    seconds_exact: float | None


def far_shell(ctx, sur: Surrogate, x, y, z, u, loops, infos, box, img, tol=1e-5,
              n_valid=4096, seed=3, exact_if_failed=True, baseline="rule7", grid_h=None):
    """u += far shell of the loops at the device atoms (x, y, z): the surrogate
    (baseline rule + MLP) when its error on n_valid random atoms is within tol
    of the exact far shell, else the exact far shell.  Returns a FarShellReport."""
    import torch
    n = x.numel()
    rng = np.random.default_rng(seed)
    idx = torch.from_numpy(np.sort(rng.choice(n, min(n_valid, n), replace=False))).to(x.device)
    xv, yv, zv = x[idx].contiguous(), y[idx].contiguous(), z[idx].contiguous()
    # the reference for the error: exact far shell on the validation atoms
    ex = exact_far_shell_torch(ctx, xv, yv, zv, loops, box, img)
    # the base as the full run computes it (on the grid of the full atom set)
    dv = [torch.zeros_like(xv) for _ in range(3)]
    if grid_h is None:
        base_far_shell_torch(ctx, xv, yv, zv, loops, box, img, baseline, u=dv)
    else:
        full = base_far_shell_torch(ctx, x, y, z, loops, box, img, baseline, grid_h=grid_h)
        dv = [t[idx].contiguous() for t in full]
        del full
    dip_only = max(float((e - d).abs().max()) for e, d in zip(ex, dv))
    sur.apply_torch(xv, yv, zv, *dv, infos, box)
    torch.cuda.synchronize()
    err = max(float((e - d).abs().max()) for e, d in zip(ex, dv))
    if err <= tol:
        t0 = time.perf_counter()
        base_far_shell_torch(ctx, x, y, z, loops, box, img, baseline, u=u, accumulate=True,
                             grid_h=grid_h)
        sur.apply_torch(x, y, z, *u, infos, box)
        torch.cuda.synchronize()
        return FarShellReport(True, err, dip_only, baseline, idx.numel(),
                              time.perf_counter() - t0, None)
    t_ex = None
    if exact_if_failed:
        t0 = time.perf_counter()
        plan = B.Plan(ctx, loops, box, img, B.PREC_FP32C, kind="line")
        set_inner(plan, 1)
        plan.run_torch(x, y, z, *u, accumulate=True)
        torch.cuda.synchronize()
        plan.free()
        t_ex = time.perf_counter() - t0
    return FarShellReport(False, err, dip_only, baseline, idx.numel(), 0.0, t_ex)
