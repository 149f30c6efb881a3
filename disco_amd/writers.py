"""The output stage after the sum, on the device (include/disco_output.h).

Mirrors the end of DISCO's driver (Code/main.cpp:797-852): final positions
(x + u wrapped into the periodic box, shifted by +L/2, mapped by U), then the
extended-XYZ file and/or the LAMMPS data file (writeLAMMPS, Code/main.cpp:50-135).
The atom lines are formatted by the device, byte for byte the reference's
iostream text; headers are formatted on the host.
"""
from __future__ import annotations

import ctypes

import numpy as np

from .backend import DiscoError, lib

TEXT_XYZ = 0
TEXT_LAMMPS = 1

_vp = ctypes.c_void_p
_ll = ctypes.c_longlong
_dp = ctypes.POINTER(ctypes.c_double)


class _TextFormat(ctypes.Structure):
    _fields_ = [("format", ctypes.c_int), ("element", ctypes.c_char * 16),
                ("d_grain", _vp), ("d_twin", _vp), ("first_id", _ll),
                ("frame", ctypes.c_double * 9)]


_bound = False


def _lib():
    global _bound
    L = lib()
    if not _bound:
        L.disco_hip_final_positions.argtypes = [_vp, _ll] + [_vp] * 6 + [_dp, _dp] + [_vp] * 4
        L.disco_hip_lammps_box.argtypes = [_dp, _dp, _dp]
        L.disco_hip_format_text.argtypes = [_vp, ctypes.POINTER(_TextFormat), _ll, _vp, _vp, _vp,
                                            _vp, _ll, ctypes.POINTER(_ll), _vp]
        L.disco_hip_write_xyz.argtypes = [_vp, ctypes.c_char_p, _ll, _vp, _vp, _vp, _dp,
                                          ctypes.c_char_p, _vp, _vp, _vp]
        L.disco_hip_write_lammps.argtypes = [_vp, ctypes.c_char_p, _ll, _vp, _vp, _vp, _dp,
                                             ctypes.c_char_p, ctypes.c_double, _vp]
        L.disco_hip_text_profile.argtypes = [_vp, ctypes.c_int]
        L.disco_hip_text_kernel_ms.argtypes = [_vp, _dp]
        _bound = True
    return L


def _d9(a):
    a = np.ascontiguousarray(np.asarray(a, dtype=np.float64).reshape(9))
    return a, a.ctypes.data_as(_dp)


def _dev(ts, dtype=None):
    import torch
    dtype = dtype or torch.float64
    for t in ts:
        if t is not None and (t.dtype != dtype or not t.is_cuda or not t.is_contiguous()):
            raise ValueError(f"device arrays must be contiguous {dtype} CUDA tensors")
    return [None if t is None else t.data_ptr() for t in ts]


def _stream():
    import torch
    return torch.cuda.current_stream().cuda_stream


def final_positions_torch(ctx, x, y, z, ux, uy, uz, box, U=None, out=None):
    """Final positions (Code/main.cpp:797-818) of device atoms: three new device
    tensors, or `out`."""
    import torch
    if out is None:
        out = [torch.empty_like(x) for _ in range(3)]
    ptrs = _dev([x, y, z, ux, uy, uz, *out])
    b = np.ascontiguousarray(np.asarray(box, dtype=np.float64).reshape(3))
    u9 = None if U is None else _d9(U)
    ctx.check(_lib().disco_hip_final_positions(
        ctx._h, x.numel(), *ptrs[:6], b.ctypes.data_as(_dp), None if u9 is None else u9[1],
        *ptrs[6:], _stream()))
    return out


def lammps_box(Hp):
    """(lx, ly, lz, xy, xz, yz) and the frame rows e1, e2, e3 of the deformed box
    Hp (columns = box vectors; Code/main.cpp:64-86)."""
    h, hp = _d9(Hp)
    b, fr = np.empty(6), np.empty(9)
    r = _lib().disco_hip_lammps_box(hp, b.ctypes.data_as(_dp), fr.ctypes.data_as(_dp))
    if r:
        raise DiscoError(r, "invalid box")
    return b, fr.reshape(3, 3)


def format_text_torch(ctx, fmt, fx, fy, fz, element="Fe", grain=None, twin=None, first_id=1,
                      frame=None, host=True):
    """The atom lines of the XYZ (fmt "xyz") or LAMMPS (fmt "lammps") writer for
    device positions: bytes on the host (host=True) or a device uint8 tensor."""
    import torch
    f = _TextFormat()
    f.format = TEXT_XYZ if fmt == "xyz" else TEXT_LAMMPS
    el = element.encode()
    if len(el) > 15:
        raise DiscoError(1, "element name longer than 15 characters")
    f.element = el
    f.d_grain, f.d_twin = _dev([grain, twin], torch.int32)
    f.first_id = int(first_id)
    if frame is not None:
        f.frame[:] = [float(v) for v in np.asarray(frame, dtype=np.float64).reshape(9)]
    ptrs = _dev([fx, fy, fz])
    nbytes = _ll(0)
    L = _lib()
    ctx.check(L.disco_hip_format_text(ctx._h, ctypes.byref(f), fx.numel(), *ptrs, None, 0,
                                      ctypes.byref(nbytes), _stream()))
    buf = torch.empty(max(nbytes.value, 1), dtype=torch.uint8, device=fx.device)
    ctx.check(L.disco_hip_format_text(ctx._h, ctypes.byref(f), fx.numel(), *ptrs,
                                      buf.data_ptr(), buf.numel(), ctypes.byref(nbytes),
                                      _stream()))
    buf = buf[:nbytes.value]
    return buf.cpu().numpy().tobytes() if host else buf


def text_profile(ctx, on=True):
    """Record the durations of the text kernels of every format call."""
    ctx.check(_lib().disco_hip_text_profile(ctx._h, 1 if on else 0))


def text_kernel_ms(ctx):
    """(count, scan, write) kernel ms of the last format call (profiling on)."""
    ms = np.zeros(3)
    ctx.check(_lib().disco_hip_text_kernel_ms(ctx._h, ms.ctypes.data_as(_dp)))
    return tuple(float(v) for v in ms)


def write_xyz(ctx, path, fx, fy, fz, Hp, element, grain=None, twin=None):
    """The extended-XYZ file (Code/main.cpp:820-845) of device final positions."""
    import torch
    g, t = _dev([grain, twin], torch.int32)
    _, hp = _d9(Hp)
    ctx.check(_lib().disco_hip_write_xyz(ctx._h, str(path).encode(), fx.numel(), *_dev([fx, fy, fz]),
                                         hp, element.encode(), g, t, _stream()))


def write_lammps(ctx, path, fx, fy, fz, Hp, element, mass):
    """The LAMMPS data file (writeLAMMPS, Code/main.cpp:50-135) of device final positions."""
    _, hp = _d9(Hp)
    ctx.check(_lib().disco_hip_write_lammps(ctx._h, str(path).encode(), fx.numel(),
                                            *_dev([fx, fy, fz]), hp, element.encode(),
                                            float(mass), _stream()))


def write_outputs(ctx, out_name, out_format, x, y, z, ux, uy, uz, box, element, mass, U=None,
                  grain=None, twin=None):
    """The driver's output step (Code/main.cpp:797-852): final positions, then
    `out_name` (XYZ) for out_format "xyz" / "both" and out_name[:-3] + "lmp"
    (LAMMPS) for "lammps" / "both".  Returns the device final positions."""
    if out_format not in ("xyz", "lammps", "both", "none"):
        raise ValueError(f"unknown output format {out_format!r}")
    f = final_positions_torch(ctx, x, y, z, ux, uy, uz, box, U)
    Uh = np.eye(3) if U is None else np.asarray(U, dtype=np.float64).reshape(3, 3)
    Hp = Uh * np.asarray(box, dtype=np.float64)[None, :]     # H'[i][j] = U[i][j] L_j
    if out_format in ("xyz", "both"):
        write_xyz(ctx, out_name, *f, Hp, element, grain, twin)
    if out_format in ("lammps", "both"):
        write_lammps(ctx, out_name[:-3] + "lmp", *f, Hp, element, mass)
    return f
