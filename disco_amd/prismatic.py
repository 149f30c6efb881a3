"""Prismatic slabs on the device (include/disco_prismatic.h).

Mirrors matterType / prismaticSlab (Code/prismatic.cpp:12-74): for a loop with
b.n != 0, the atoms within k = |b.n| / d planes of the cut and inside the
contour, in ascending index order, with the side of the cut of each -- the
reference's slab bit for bit.  The loop is its constructed state
(``geometry.native_loop_state``).

Then the driver's use of the slabs (Code/main.cpp:586-687, 733-742), also on
the device: ``apply_slabs_torch`` filters each slab to the loop's grain and
twin region, deletes vacancy atoms and renumbers the rest in order, and lists
the interstitial copies (source, shift); ``append_copies_torch`` adds the
copies after the displacement sum.
"""
from __future__ import annotations

import ctypes

from .backend import lib

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
        _llp = ctypes.POINTER(_ll)
        L.disco_hip_slab_filter.argtypes = [_vp, _ll, _vp, ctypes.c_int, _vp, ctypes.c_int, _vp,
                                            _vp, _llp, _vp]
        L.disco_hip_slab_mark_removed.argtypes = [_vp, _ll, _ll, _vp, _vp, _vp]
        L.disco_hip_slab_copies.argtypes = [_vp, _ll, _vp, _vp, ctypes.POINTER(_d), _vp, _vp, _vp]
        L.disco_hip_drop_atoms.argtypes = [_vp, _ll, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp,
                                           _ll, _vp, _vp, _vp, _vp, _llp, _llp, _vp]
        L.disco_hip_append_copies.argtypes = [_vp, _ll, _ll, _vp, _vp, _vp, _vp, _vp, _vp, _vp,
                                              _vp, _vp, _vp]
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


def _stream():
    import torch
    return torch.cuda.current_stream().cuda_stream


def _ptr(t):
    return None if t is None else t.data_ptr()


def _check_i32(*ts):
    import torch
    for t in ts:
        if t is not None and (t.dtype != torch.int32 or not t.is_cuda or not t.is_contiguous()):
            raise ValueError("index and label arrays must be contiguous int32 CUDA tensors")


def slab_filter_torch(ctx, natoms, atoms, side, grain=None, g=-1, twin=None, region=0):
    """The slab entries of the loop's grain (grain[a] == g, when grain is given
    and g >= 0) and twin region (twin[a] == region, when twin is given),
    Code/main.cpp:605-615.  Returns (atoms, side) views of the kept prefix."""
    _check_i32(atoms, side, grain, twin)
    if g < 0:
        grain = None
    atoms, side = atoms.contiguous(), side.contiguous()
    count = ctypes.c_longlong(atoms.numel())
    ctx.check(_lib().disco_hip_slab_filter(ctx._h, int(natoms), _ptr(grain), int(g), _ptr(twin),
                                           int(region), _ptr(atoms), _ptr(side),
                                           ctypes.byref(count), _stream()))
    return atoms[:count.value], side[:count.value]


def mark_removed_torch(ctx, atoms, removed):
    """removed[atoms] = 1 (uint8 flags of every atom), Code/main.cpp:626-628."""
    _check_i32(atoms)
    ctx.check(_lib().disco_hip_slab_mark_removed(ctx._h, removed.numel(), atoms.numel(),
                                                 _ptr(atoms), _ptr(removed), _stream()))


def slab_copies_torch(ctx, atoms, side, shifts):
    """The copies of an interstitial slab: (source int32 (k,), shift (k, 3))."""
    import torch
    _check_i32(atoms, side)
    k = atoms.numel()
    ca = torch.empty(max(k, 1), dtype=torch.int32, device=atoms.device)
    cs = torch.empty((max(k, 1), 3), dtype=torch.float64, device=atoms.device)
    sh = (_d * 6)(*[float(v) for v in list(shifts[0]) + list(shifts[1])])
    ctx.check(_lib().disco_hip_slab_copies(ctx._h, k, _ptr(atoms), _ptr(side), sh, _ptr(ca),
                                           _ptr(cs), _stream()))
    return ca[:k], cs[:k]


def drop_atoms_torch(ctx, removed, x, y, z, label=None, copy_atom=None, copy_shift=None):
    """Atoms not flagged, in order (new tensors), and the copies renumbered with
    those of deleted atoms dropped (Code/main.cpp:661-684)."""
    import torch
    _check_i32(label, copy_atom)
    n = x.numel()
    dev = x.device
    xo, yo, zo = (torch.empty(max(n, 1), dtype=torch.float64, device=dev) for _ in range(3))
    lo = None if label is None else torch.empty(max(n, 1), dtype=torch.int32, device=dev)
    nc = 0 if copy_atom is None else copy_atom.numel()
    cao = torch.empty(max(nc, 1), dtype=torch.int32, device=dev)
    cso = torch.empty((max(nc, 1), 3), dtype=torch.float64, device=dev)
    if nc:
        copy_shift = copy_shift.contiguous()
    nk, nck = ctypes.c_longlong(0), ctypes.c_longlong(0)
    ctx.check(_lib().disco_hip_drop_atoms(
        ctx._h, n, _ptr(removed), _ptr(x), _ptr(y), _ptr(z), _ptr(label), _ptr(xo), _ptr(yo),
        _ptr(zo), _ptr(lo), nc, _ptr(copy_atom) if nc else None,
        _ptr(copy_shift) if nc else None, _ptr(cao), _ptr(cso), ctypes.byref(nk),
        ctypes.byref(nck), _stream()))
    k, c = nk.value, nck.value
    return (xo[:k], yo[:k], zo[:k], None if lo is None else lo[:k]), (cao[:c], cso[:c])


def append_copies_torch(ctx, x, y, z, ux, uy, uz, copy_atom, copy_shift, label=None):
    """Code/main.cpp:733-742: new tensors of n + ncopies entries, copy k at its
    source's site (and label) with the source's displacement plus its shift."""
    import torch
    _check_i32(copy_atom, label)
    n, c = x.numel(), copy_atom.numel()
    outs = []
    for t in (x, y, z, ux, uy, uz):
        o = torch.empty(n + c, dtype=torch.float64, device=t.device)
        o[:n].copy_(t)
        outs.append(o)
    lab = None
    if label is not None:
        lab = torch.empty(n + c, dtype=torch.int32, device=label.device)
        lab[:n].copy_(label)
    ctx.check(_lib().disco_hip_append_copies(ctx._h, n, c, _ptr(copy_atom),
                                             _ptr(copy_shift.contiguous()), *(_ptr(o) for o in outs),
                                             _ptr(lab), _stream()))
    return outs, lab


def apply_slabs_torch(ctx, x, y, z, box, loops, grain=None, twin=None):
    """The driver's prismatic step (Code/main.cpp:586-687) on device atoms.

    loops: per loop a dict with ``state`` (disco_loop_state), ``spacing``, and
    optionally ``grain`` (the loop's grain, -1 none), ``region`` (its twin
    region), ``frank`` (FCC Frank loop), ``R`` (its grain's orientation, 3x3 or
    None), ``a0`` and ``matter`` (False: displacement only).  Every slab is
    selected on the given positions; then vacancy atoms are deleted.  Returns a
    dict: x, y, z, grain (compacted; grain None when not given), copy_atom and
    copy_shift (the interstitial copies, renumbered) and per-loop reports
    (type, planes, height, spacing, atoms)."""
    import torch
    from .geometry import native_copy_shifts
    n = x.numel()
    removed = torch.zeros(max(n, 1), dtype=torch.uint8, device=x.device)
    cas, css, reports = [], [], []
    any_loop = False
    for lp in loops:
        st = lp["state"]
        if matter_type(st) == 0 or not lp.get("matter", True):
            reports.append(None)
            continue
        any_loop = True
        atoms, side, info = prismatic_slab_torch(ctx, st, lp["spacing"], x, y, z, box)
        gl = int(lp.get("grain", -1))
        if gl >= 0 or twin is not None:
            atoms, side = slab_filter_torch(ctx, n, atoms, side, grain, gl, twin,
                                            int(lp.get("region", 0)))
        if info["type"] == 1:
            mark_removed_torch(ctx, atoms, removed[:n])
        else:
            sh = native_copy_shifts(info["b"], lp.get("frank", False), list(st.n), lp.get("R"),
                                    lp.get("a0", 1.0))
            ca, cs = slab_copies_torch(ctx, atoms, side, sh)
            cas.append(ca)
            css.append(cs)
        reports.append({**info, "atoms": (-1 if info["type"] == 1 else 1) * atoms.numel()})
    ca = torch.cat(cas) if cas else torch.empty(0, dtype=torch.int32, device=x.device)
    cs = torch.cat(css) if css else torch.empty((0, 3), dtype=torch.float64, device=x.device)
    if not any_loop:
        return {"x": x, "y": y, "z": z, "grain": grain, "copy_atom": ca, "copy_shift": cs,
                "reports": reports}
    (xo, yo, zo, go), (ca, cs) = drop_atoms_torch(ctx, removed[:n], x, y, z, grain, ca, cs)
    return {"x": xo, "y": yo, "z": zo, "grain": go, "copy_atom": ca, "copy_shift": cs,
            "reports": reports}
