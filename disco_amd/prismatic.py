"""Prismatic slab selection on the device (include/disco_prismatic.h).

Mirrors matterType / prismaticSlab (Code/prismatic.cpp:12-74): for a loop with
b.n != 0, the atoms within k = |b.n| / d planes of the cut and inside the
contour, in ascending index order, with the side of the cut of each -- the
reference's slab bit for bit.  The loop is its constructed state
(``geometry.native_loop_state``).
"""
from __future__ import annotations

import ctypes

from .backend import DiscoError, lib

MATTER = {0: "glide", 1: "vacancy", 2: "interstitial"}

_vp = ctypes.c_void_p
_ll = ctypes.c_longlong
_d = ctypes.c_double


class _SlabInfo(ctypes.Structure):
    _fields_ = [("type", ctypes.c_int), ("planes", ctypes.c_int), ("height", _d),
                ("spacing", _d), ("b", _d * 3), ("count", _ll)]


_bound = False


def _lib():
    global _bound
    L = lib()
    if not _bound:
        L.disco_hip_matter_type.argtypes = [_vp]
        L.disco_hip_prismatic_slab.argtypes = [_vp, _vp, _d, _ll, _vp, _vp, _vp, _d, _d, _d, _vp,
                                               _vp, _ll, ctypes.POINTER(_SlabInfo), _vp]
        _bound = True
    return L


def matter_type(state) -> int:
    """0 glide, 1 vacancy, 2 interstitial (matterType, Code/prismatic.cpp:12-18)."""
    return _lib().disco_hip_matter_type(ctypes.addressof(state))


def prismatic_slab_torch(ctx, state, spacing, x, y, z, box, cap=None):
    """The slab of one loop among device atoms: (atoms int32, side int32) device
    tensors of the slab size, and info (type, planes, height, spacing, b, count)."""
    import torch
    for t in (x, y, z):
        if t.dtype != torch.float64 or not t.is_cuda or not t.is_contiguous():
            raise ValueError("atoms must be contiguous float64 CUDA tensors")
    n = x.numel()
    cap = n if cap is None else int(cap)
    atoms = torch.empty(max(cap, 1), dtype=torch.int32, device=x.device)
    side = torch.empty(max(cap, 1), dtype=torch.int32, device=x.device)
    info = _SlabInfo()
    ctx.check(_lib().disco_hip_prismatic_slab(
        ctx._h, ctypes.addressof(state), float(spacing), n, x.data_ptr(), y.data_ptr(),
        z.data_ptr(), float(box[0]), float(box[1]), float(box[2]), atoms.data_ptr(),
        side.data_ptr(), cap, ctypes.byref(info),
        torch.cuda.current_stream().cuda_stream))
    k = info.count
    return atoms[:k], side[:k], {"type": info.type, "planes": info.planes,
                                 "height": info.height, "spacing": info.spacing,
                                 "b": tuple(info.b), "count": k}
