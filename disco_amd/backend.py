"""Python mirror of the reference's backend interface, on the HIP C ABI.

Mirrors ``Code/backend.hpp``:

* ``LoopData``       -- Code/backend.hpp:20-26 (loopons SoA or polyline + apex)
* ``BackendTiming``  -- Code/backend.hpp:36-39
* ``GpuOptions``     -- Code/backend.hpp:53-56 (``fp32`` selects DISCO_PREC_FP32C;
                        ``chunk_pairs`` is accepted and ignored: each loop's pass
                        is one launch over all atoms, and no watchdog needs chunking)
* ``gpu_available``, ``print_gpu_info``, ``gpu_line_fp64``, ``kroupa_gpu``
                     -- Code/backend.hpp:57-63 / Code/backend_cuda.cu:525-722

``kroupa_gpu`` keeps the reference's contract: the caller owns the host arrays,
``ux, uy, uz`` end holding (their input +) the sum over loops and images.  All
compute goes through ``libdisco_hip.so``; there is no CPU path here -- a
missing library or device raises.
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass, field

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libdisco_hip.so")

PREC_FP64 = 0
PREC_FP32C = 1

ERR_NAMES = {0: "DISCO_OK", 1: "DISCO_ERR_ARG", 2: "DISCO_ERR_NO_DEVICE", 3: "DISCO_ERR_HIP",
             4: "DISCO_ERR_OOM", 5: "DISCO_ERR_TOO_LARGE", 6: "DISCO_ERR_STATE",
             7: "DISCO_ERR_IO"}


# ------------------------------------------------------------------ types
@dataclass
class LoopData:
    """One loop for the backend (Code/backend.hpp:20-26)."""
    phys: tuple = (0.0,) * 10            # bx by bz mx my mz nx ny nz dS
    line: bool = False
    x: np.ndarray = field(default_factory=lambda: np.zeros(0))
    y: np.ndarray = field(default_factory=lambda: np.zeros(0))
    z: np.ndarray = field(default_factory=lambda: np.zeros(0))
    vx: np.ndarray = field(default_factory=lambda: np.zeros(0))
    vy: np.ndarray = field(default_factory=lambda: np.zeros(0))
    vz: np.ndarray = field(default_factory=lambda: np.zeros(0))
    cx: float = 0.0
    cy: float = 0.0
    cz: float = 0.0

    def burgers(self):
        """b = (bx mx, by my, bz mz), Code/backend_cpu.cpp:44-45."""
        p = self.phys
        return (p[0] * p[3], p[1] * p[4], p[2] * p[5])


@dataclass
class BackendTiming:
    upload: float = 0.0
    kernels: float = 0.0
    download: float = 0.0
    kernel_ms: float = 0.0
    loop_sec: list = field(default_factory=list)


@dataclass
class GpuOptions:
    fp32: bool = False
    chunk_pairs: float = -1.0


class DiscoError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{ERR_NAMES.get(code, code)}: {msg}")
        self.code = code


# ------------------------------------------------------------------ C ABI
class _LineLoop(ctypes.Structure):
    _fields_ = [("nv", ctypes.c_int), ("vx", ctypes.POINTER(ctypes.c_double)),
                ("vy", ctypes.POINTER(ctypes.c_double)), ("vz", ctypes.POINTER(ctypes.c_double)),
                ("c", ctypes.c_double * 3), ("b", ctypes.c_double * 3)]


class _LooponLoop(ctypes.Structure):
    _fields_ = [("np", ctypes.c_int), ("x", ctypes.POINTER(ctypes.c_double)),
                ("y", ctypes.POINTER(ctypes.c_double)), ("z", ctypes.POINTER(ctypes.c_double)),
                ("phys", ctypes.c_double * 10)]


class _Timing(ctypes.Structure):
    _fields_ = [("upload", ctypes.c_double), ("kernels", ctypes.c_double),
                ("download", ctypes.c_double), ("kernel_ms", ctypes.c_double),
                ("loop_sec", ctypes.POINTER(ctypes.c_double))]


class _Sample(ctypes.Structure):
    _fields_ = [("atom_offset", ctypes.c_longlong), ("n", ctypes.c_int), ("nloops", ctypes.c_int),
                ("loops", ctypes.POINTER(_LineLoop)), ("L", ctypes.c_double * 3)]


class _Tuning(ctypes.Structure):
    _fields_ = [("block", ctypes.c_int), ("tile", ctypes.c_int), ("groups", ctypes.c_int)]


class _DeviceInfo(ctypes.Structure):
    _fields_ = [("name", ctypes.c_char * 256), ("compute_units", ctypes.c_int),
                ("global_mem", ctypes.c_size_t), ("lds_per_block", ctypes.c_size_t),
                ("arch_major", ctypes.c_int), ("arch_minor", ctypes.c_int),
                ("arch", ctypes.c_char * 64)]


_lib = None
_dp = ctypes.POINTER(ctypes.c_double)
_vp = ctypes.c_void_p
_i = ctypes.c_int
_d = ctypes.c_double


def lib():
    """libdisco_hip.so; raises if it was not built (there is no fallback)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"HIP backend library missing: {LIB_PATH} "
                               "(build it with __graft_entry__.build())")
        # One HIP runtime per process: torch links its bundled libamdhip64 by
        # the unversioned name, this library by SONAME libamdhip64.so.7.  Loading
        # torch first makes the SONAME resolve to torch's copy, so device
        # pointers and stream handles from torch are valid here.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = ctypes.CDLL(LIB_PATH)
        L.disco_hip_device_count.argtypes = [ctypes.POINTER(_i)]
        L.disco_hip_device_info.argtypes = [_i, ctypes.POINTER(_DeviceInfo)]
        L.disco_hip_line_fp64.argtypes = []
        L.disco_hip_strerror.argtypes = [_i]
        L.disco_hip_strerror.restype = ctypes.c_char_p
        L.disco_hip_open.argtypes = [_i, ctypes.POINTER(_vp)]
        L.disco_hip_close.argtypes = [_vp]
        L.disco_hip_last_error.argtypes = [_vp]
        L.disco_hip_last_error.restype = ctypes.c_char_p
        L.disco_hip_line_sum.argtypes = [_vp, _i, _dp, _dp, _dp, _dp, _dp, _dp,
                                         ctypes.POINTER(_LineLoop), _i, _d, _d, _d, _i, _i,
                                         ctypes.POINTER(_Timing)]
        L.disco_hip_loopon_sum.argtypes = [_vp, _i, _dp, _dp, _dp, _dp, _dp, _dp,
                                           ctypes.POINTER(_LooponLoop), _i, _d, _d, _d, _i, _i,
                                           ctypes.POINTER(_Timing)]
        L.disco_hip_plan_line.argtypes = [_vp, ctypes.POINTER(_LineLoop), _i, _d, _d, _d, _i, _i,
                                          ctypes.POINTER(_vp)]
        L.disco_hip_plan_loopon.argtypes = [_vp, ctypes.POINTER(_LooponLoop), _i, _d, _d, _d, _i,
                                            _i, ctypes.POINTER(_vp)]
        L.disco_hip_plan_run.argtypes = [_vp, _i, _vp, _vp, _vp, _vp, _vp, _vp, _i, _vp]
        L.disco_hip_plan_run_noalloc.argtypes = [_vp, _i, _vp, _vp, _vp, _vp, _vp, _vp, _i, _vp]
        L.disco_hip_plan_set_cubature.argtypes = [_vp, _d, _d, ctypes.POINTER(_i)]
        L.disco_hip_line_cubature.argtypes = [ctypes.POINTER(_LineLoop), _i, _dp]
        L.disco_hip_plan_set_band.argtypes = [_vp, _d, ctypes.POINTER(_i)]
        L.disco_hip_line_band.argtypes = [ctypes.POINTER(_LineLoop), _i, _dp]
        L.disco_hip_plan_set_image_skip.argtypes = [_vp, ctypes.POINTER(ctypes.c_uint), _i]
        L.disco_hip_plan_set_image_skip_octants.argtypes = [_vp, ctypes.POINTER(ctypes.c_uint), _i,
                                                            ctypes.POINTER(ctypes.c_double)]
        L.disco_hip_plan_profile.argtypes = [_vp, _i]
        L.disco_hip_plan_kernel_ms.argtypes = [_vp, _dp, _i]
        L.disco_hip_plan_evals_per_atom.argtypes = [_vp]
        L.disco_hip_plan_evals_per_atom.restype = _d
        L.disco_hip_plan_free.argtypes = [_vp]
        L.disco_hip_batch_plan.argtypes = [_vp, ctypes.POINTER(_Sample), _i, _i, _i,
                                           ctypes.POINTER(_vp)]
        L.disco_hip_batch_set_tuning.argtypes = [_vp, ctypes.POINTER(_Tuning)]
        L.disco_hip_batch_get_tuning.argtypes = [_vp, ctypes.POINTER(_Tuning)]
        L.disco_hip_batch_autotune.argtypes = [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _i, _vp,
                                               ctypes.POINTER(_Tuning), _dp]
        L.disco_hip_batch_run.argtypes = [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _i, _vp]
        L.disco_hip_batch_evals.argtypes = [_vp]
        L.disco_hip_batch_evals.restype = _d
        L.disco_hip_batch_free.argtypes = [_vp]
        L.disco_hip_lattice_atoms.argtypes = [_i, _i, _i, _i]
        L.disco_hip_lattice_atoms.restype = ctypes.c_longlong
        L.disco_hip_lattice.argtypes = [_vp, _i, _i, _i, _i, _d, _vp, _vp, _vp, _vp]
        L.disco_hip_grid_interp.argtypes = [_vp, _i] + [_vp] * 9 + [_i, _i, _i, _dp, _d, _i,
                                                                     _i, _vp]
        _lib = L
    return _lib


def _check(code, ctx=None):
    if code != 0:
        L = lib()
        detail = L.disco_hip_last_error(ctx).decode() if ctx else ""
        raise DiscoError(code, (L.disco_hip_strerror(code).decode() +
                                (f" ({detail})" if detail else "")))


# ------------------------------------------------------------------ lattice
LATTICE = {"BCC": 0, "FCC": 1}


def lattice_atoms(kind: str, nx: int, ny: int, nz: int) -> int:
    n = lib().disco_hip_lattice_atoms(LATTICE[kind], nx, ny, nz)
    if n < 0:
        raise DiscoError(1, "invalid lattice arguments")
    return int(n)


def lattice_torch(ctx, kind: str, nx: int, ny: int, nz: int, a0: float, out=None):
    """Perfect lattice generated on the device (generateLattice, Code/lattice.hpp:35-51):
    returns device tensors x, y, z (FP64), or fills `out` (three tensors of
    lattice_atoms() elements each)."""
    import torch
    n = lattice_atoms(kind, nx, ny, nz)
    if out is None:
        out = [torch.empty(n, dtype=torch.float64, device=f"cuda:{ctx.device}") for _ in range(3)]
    for t in out:
        if t.numel() < n or t.dtype != torch.float64 or not t.is_contiguous():
            raise DiscoError(1, "output tensors must be contiguous FP64 with room for every site")
    stream = torch.cuda.current_stream().cuda_stream
    ctx.check(lib().disco_hip_lattice(ctx._h, LATTICE[kind], nx, ny, nz, float(a0),
                                      out[0].data_ptr(), out[1].data_ptr(), out[2].data_ptr(),
                                      stream))
    return out


# ------------------------------------------------------------------ grid field
def grid_interp_torch(ctx, x, y, z, ux, uy, uz, g, origin, h, accumulate=True, order=4):
    """u (+)= the field g (three device tensors of shape (nz, ny, nx), x fastest;
    node (i, j, k) at origin + (i, j, k) h) interpolated at the device atoms by
    tensor-product Lagrange on `order` nodes per axis (4: cubic, 6: quintic)."""
    import torch
    gz_, gy_, gx_ = g[0].shape
    gc = [t.contiguous() for t in g]
    o = (ctypes.c_double * 3)(*[float(v) for v in origin])
    stream = torch.cuda.current_stream().cuda_stream
    ctx.check(lib().disco_hip_grid_interp(
        ctx._h, x.numel(), x.data_ptr(), y.data_ptr(), z.data_ptr(), ux.data_ptr(),
        uy.data_ptr(), uz.data_ptr(), gc[0].data_ptr(), gc[1].data_ptr(), gc[2].data_ptr(),
        gx_, gy_, gz_, o, float(h), int(order), 1 if accumulate else 0, stream))


# ------------------------------------------------------------------ devices
def gpu_available() -> bool:
    """A HIP device can be used (Code/backend_cuda.cu:537-541)."""
    n = _i(0)
    return lib().disco_hip_device_count(ctypes.byref(n)) == 0 and n.value > 0


def gpu_info(device: int = 0) -> dict:
    info = _DeviceInfo()
    _check(lib().disco_hip_device_info(device, ctypes.byref(info)))
    return {"name": info.name.decode(), "compute_units": info.compute_units,
            "global_mem": info.global_mem, "lds_per_block": info.lds_per_block,
            "arch": info.arch.decode()}


def print_gpu_info(device: int = 0) -> None:
    """The "--- GPU: <name> ---" banner (Code/backend_cuda.cu:543-553), in CUs."""
    g = gpu_info(device)
    print(f"\n--- GPU: {g['name']} ({g['arch']}) ---")
    print(f"  CUs              : {g['compute_units']}")
    print(f"  Global memory    : {g['global_mem'] >> 20} MB")
    print(f"  LDS/block        : {g['lds_per_block'] // 1024} KB\n")


def gpu_line_fp64() -> bool:
    return bool(lib().disco_hip_line_fp64())


# ------------------------------------------------------------------ context
class Context:
    """A device context of the C ABI (one stream, cached buffers)."""

    def __init__(self, device: int = 0):
        self._h = _vp()
        self.device = device
        _check(lib().disco_hip_open(device, ctypes.byref(self._h)))

    @property
    def handle(self):
        return self._h

    def close(self):
        if self._h:
            lib().disco_hip_close(self._h)
            self._h = _vp()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def check(self, code):
        _check(code, self._h)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def line_structs(loops):
    """ctypes array of disco_line_loop; keeps the numpy buffers alive."""
    keep = []
    arr = (_LineLoop * max(1, len(loops)))()
    for i, lp in enumerate(loops):
        vx, vy, vz = _f64(lp.vx), _f64(lp.vy), _f64(lp.vz)
        keep += [vx, vy, vz]
        arr[i].nv = int(vx.shape[0])
        arr[i].vx = vx.ctypes.data_as(_dp)
        arr[i].vy = vy.ctypes.data_as(_dp)
        arr[i].vz = vz.ctypes.data_as(_dp)
        arr[i].c[:] = [lp.cx, lp.cy, lp.cz]
        arr[i].b[:] = list(lp.burgers())
    return arr, keep


def loopon_structs(loops):
    keep = []
    arr = (_LooponLoop * max(1, len(loops)))()
    for i, lp in enumerate(loops):
        x, y, z = _f64(lp.x), _f64(lp.y), _f64(lp.z)
        keep += [x, y, z]
        arr[i].np = int(x.shape[0])
        arr[i].x = x.ctypes.data_as(_dp)
        arr[i].y = y.ctypes.data_as(_dp)
        arr[i].z = z.ctypes.data_as(_dp)
        arr[i].phys[:] = [float(t) for t in lp.phys]
    return arr, keep


def _sum_host(ctx, kind, x, y, z, u, loops, box, img, prec, timing):
    x, y, z = _f64(x), _f64(y), _f64(z)
    n = x.shape[0]
    ux, uy, uz = u
    for a in (ux, uy, uz):
        if a.dtype != np.float64 or not a.flags.c_contiguous or a.shape[0] != n:
            raise ValueError("u arrays must be C-contiguous float64 of length n")
    t = _Timing()
    sec = np.zeros(max(1, len(loops)))
    # per-loop times only when asked for: they run each loop in its own launch
    t.loop_sec = sec.ctypes.data_as(_dp) if timing is not None else None
    if kind == "line":
        arr, keep = line_structs(loops)
        fn = lib().disco_hip_line_sum
    else:
        arr, keep = loopon_structs(loops)
        fn = lib().disco_hip_loopon_sum
    ctx.check(fn(ctx.handle, n, x.ctypes.data_as(_dp), y.ctypes.data_as(_dp),
                 z.ctypes.data_as(_dp), ux.ctypes.data_as(_dp), uy.ctypes.data_as(_dp),
                 uz.ctypes.data_as(_dp), arr, len(loops), float(box[0]), float(box[1]),
                 float(box[2]), int(img), int(prec), ctypes.byref(t)))
    del keep
    if timing is not None:
        timing.upload, timing.kernels, timing.download = t.upload, t.kernels, t.download
        timing.kernel_ms = t.kernel_ms
        timing.loop_sec = list(sec[:len(loops)])
    return ux, uy, uz


def line_cubature(loops) -> np.ndarray:
    """Per loop: (eligible, centre weight, ring weight, ring radius, cut radius)
    of the 9-point cubature the cubature form uses (host only)."""
    arr, keep = line_structs(loops)
    out = np.zeros((max(1, len(loops)), 5))
    _check(lib().disco_hip_line_cubature(arr, len(loops), out.ctypes.data_as(_dp)))
    del keep
    return out[:len(loops)]


def line_band(loops) -> np.ndarray:
    """Per loop: (eligible, centre weight, per-point weight of ring 1, of ring 2,
    radius of ring 1, of ring 2) of the two-ring cubature of the band form
    (host only)."""
    arr, keep = line_structs(loops)
    out = np.zeros((max(1, len(loops)), 6))
    _check(lib().disco_hip_line_band(arr, len(loops), out.ctypes.data_as(_dp)))
    del keep
    return out[:len(loops)]


def line_sum(ctx, x, y, z, loops, box, img=1, prec=PREC_FP64, u=None, timing=None):
    """Host-buffer line sum: u (+)= sum over loops x images (one launch; one per
    loop, each timed, when `timing` is given)."""
    n = len(x)
    u = (np.zeros(n), np.zeros(n), np.zeros(n)) if u is None else u
    return _sum_host(ctx, "line", x, y, z, u, loops, box, img, prec, timing)


def loopon_sum(ctx, x, y, z, loops, box, img=1, prec=PREC_FP64, u=None, timing=None):
    """Host-buffer loopon sum: u (+)= sum over loops x images x loopons."""
    n = len(x)
    u = (np.zeros(n), np.zeros(n), np.zeros(n)) if u is None else u
    return _sum_host(ctx, "loopon", x, y, z, u, loops, box, img, prec, timing)


class Plan:
    """Device-resident geometry of one sample (disco_hip_plan_*)."""

    def __init__(self, ctx, loops, box, img=1, prec=PREC_FP32C, kind=None):
        self.ctx = ctx
        self._h = _vp()
        kind = kind or ("line" if (loops and loops[0].line) else "loopon")
        self.kind = kind
        if kind == "line":
            arr, keep = line_structs(loops)
            ctx.check(lib().disco_hip_plan_line(ctx.handle, arr, len(loops), float(box[0]),
                                                float(box[1]), float(box[2]), int(img),
                                                int(prec), ctypes.byref(self._h)))
        else:
            arr, keep = loopon_structs(loops)
            ctx.check(lib().disco_hip_plan_loopon(ctx.handle, arr, len(loops), float(box[0]),
                                                  float(box[1]), float(box[2]), int(img),
                                                  int(prec), ctypes.byref(self._h)))
        del keep

    def set_cubature(self, k_exact: float, k_rule: float | None = None) -> int:
        """Cubature form of the far pass (FP32C line plans): exact within k_exact
        loop radii of a loop image, the 9-point cubature beyond k_rule (default
        k_exact: a hard switch), a smooth blend between; k_exact <= 0 switches it
        off.  Returns the number of loops that use it."""
        k_rule = k_exact if k_rule is None else k_rule
        n = _i(0)
        self.ctx.check(lib().disco_hip_plan_set_cubature(self._h, float(k_exact), float(k_rule),
                                                         ctypes.byref(n)))
        return n.value

    def set_band(self, k_band: float) -> int:
        """Band form of a cubature plan: the two-ring cubature in place of the
        exact field for lanes beyond k_band loop radii of a loop image (loops it
        represents); k_band <= 0 switches it off.  Returns the number of loops
        that use it."""
        n = _i(0)
        self.ctx.check(lib().disco_hip_plan_set_band(self._h, float(k_band), ctypes.byref(n)))
        return n.value

    def set_image_skip(self, masks=None, split=None):
        """Leave images of the +-1 shell out of the sums (the caller sums them
        another way): masks[l] (bit 9 (i + 1) + 3 (j + 1) + (k + 1)) for every
        atom, or masks[l, o] for the atoms of octant o = [x >= split[0]] +
        2 [y >= split[1]] + 4 [z >= split[2]] when split is given; None clears."""
        if masks is None:
            self.ctx.check(lib().disco_hip_plan_set_image_skip(self._h, None, 0))
            return
        m = np.ascontiguousarray(masks, dtype=np.uint32)
        ptr = m.ctypes.data_as(ctypes.POINTER(ctypes.c_uint))
        if split is None:
            if m.ndim != 1:
                raise ValueError("masks per octant need the octant planes (split)")
            self.ctx.check(lib().disco_hip_plan_set_image_skip(self._h, ptr, m.shape[0]))
            return
        if m.ndim != 2 or m.shape[1] != 8:
            raise ValueError("octant masks must have shape (loops, 8)")
        sp = (ctypes.c_double * 3)(*[float(v) for v in split])
        self.ctx.check(lib().disco_hip_plan_set_image_skip_octants(self._h, ptr, m.shape[0], sp))

    def profile(self, on: bool = True):
        """Record HIP events around each kernel of later runs."""
        self.ctx.check(lib().disco_hip_plan_profile(self._h, 1 if on else 0))

    def kernel_ms(self):
        """(main pass ms, near pass ms) of the last profiled run (waits for it)."""
        ms = (ctypes.c_double * 2)()
        self.ctx.check(lib().disco_hip_plan_kernel_ms(self._h, ms, 2))
        return float(ms[0]), float(ms[1])

    def evals_per_atom(self) -> float:
        return lib().disco_hip_plan_evals_per_atom(self._h)

    def run_ptr(self, n, px, py, pz, pux, puy, puz, accumulate=False, stream=None):
        """Launch on device pointers (ints), asynchronously on `stream` (int or None)."""
        self.ctx.check(lib().disco_hip_plan_run(self._h, int(n), px, py, pz, pux, puy, puz,
                                                1 if accumulate else 0, stream))

    def run_torch(self, x, y, z, ux, uy, uz, accumulate=False, stream=None):
        """Launch on torch CUDA float64 tensors, on torch's current stream by default."""
        import torch
        for t in (x, y, z, ux, uy, uz):
            if t.dtype != torch.float64 or not t.is_cuda or not t.is_contiguous():
                raise ValueError("device arrays must be contiguous float64 CUDA tensors")
        if stream is None:
            stream = torch.cuda.current_stream().cuda_stream
        self.run_ptr(x.numel(), x.data_ptr(), y.data_ptr(), z.data_ptr(), ux.data_ptr(),
                     uy.data_ptr(), uz.data_ptr(), accumulate, stream)

    def free(self):
        if self._h:
            lib().disco_hip_plan_free(self._h)
            self._h = _vp()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Batch:
    """Many samples summed by one persistent launch per pass, replayed as a
    HIP graph (disco_hip_batch_*).  `samples` is a list of geometry.Sample
    (or any object with .n, .loops, .box) laid out back to back in the atom
    arrays, in list order."""

    def __init__(self, ctx, samples, img=1, prec=PREC_FP32C):
        self.ctx = ctx
        self._h = _vp()
        arr = (_Sample * max(1, len(samples)))()
        self._keep = []
        off = 0
        self.offsets = []
        for i, s in enumerate(samples):
            la, keep = line_structs(s.loops)
            self._keep += [la, keep]
            arr[i].atom_offset = off
            arr[i].n = int(s.n)
            arr[i].nloops = len(s.loops)
            arr[i].loops = ctypes.cast(la, ctypes.POINTER(_LineLoop))
            arr[i].L[:] = [float(t) for t in s.box]
            self.offsets.append(off)
            off += int(s.n)
        self.n = off
        ctx.check(lib().disco_hip_batch_plan(ctx.handle, arr, len(samples), int(img), int(prec),
                                             ctypes.byref(self._h)))
        self._keep = None          # the plan copied everything it needs

    def evals(self) -> float:
        return lib().disco_hip_batch_evals(self._h)

    @property
    def tuning(self):
        t = _Tuning()
        self.ctx.check(lib().disco_hip_batch_get_tuning(self._h, ctypes.byref(t)))
        return {"block": t.block, "tile": t.tile, "groups": t.groups}

    def set_tuning(self, block, tile, groups):
        t = _Tuning(block, tile, groups)
        self.ctx.check(lib().disco_hip_batch_set_tuning(self._h, ctypes.byref(t)))

    @staticmethod
    def _ptrs(ts):
        import torch
        for t in ts:
            if t.dtype != torch.float64 or not t.is_cuda or not t.is_contiguous():
                raise ValueError("device arrays must be contiguous float64 CUDA tensors")
        return [t.data_ptr() for t in ts]

    def autotune(self, x, y, z, ux, uy, uz, stride=8, stream=None):
        import torch
        if stream is None:
            stream = torch.cuda.current_stream().cuda_stream
        t = _Tuning()
        ms = ctypes.c_double(0.0)
        self.ctx.check(lib().disco_hip_batch_autotune(self._h, *self._ptrs((x, y, z, ux, uy, uz)),
                                                      int(stride), stream, ctypes.byref(t),
                                                      ctypes.byref(ms)))
        return {"block": t.block, "tile": t.tile, "groups": t.groups}, ms.value

    def run_torch(self, x, y, z, ux, uy, uz, accumulate=False, stream=None):
        import torch
        if stream is None:
            stream = torch.cuda.current_stream().cuda_stream
        self.ctx.check(lib().disco_hip_batch_run(self._h, *self._ptrs((x, y, z, ux, uy, uz)),
                                                 1 if accumulate else 0, stream))

    def free(self):
        if self._h:
            lib().disco_hip_batch_free(self._h)
            self._h = _vp()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def kroupa_gpu(x, y, z, ux, uy, uz, loops, dx, dy, dz, opt: GpuOptions | None = None,
               timing: BackendTiming | None = None, device: int = 0, img: int = 1):
    """kroupaGPU (Code/backend_cuda.cu:568-722): every loop, 27 images (img = 1),
    line and loopon loops alike; ux, uy, uz accumulate the result."""
    opt = opt or GpuOptions()
    prec = PREC_FP32C if opt.fp32 else PREC_FP64
    line = [l for l in loops if l.line]
    lps = [l for l in loops if not l.line]
    tl, tp = BackendTiming(), BackendTiming()
    with Context(device) as ctx:
        if line:
            line_sum(ctx, x, y, z, line, (dx, dy, dz), img, prec, (ux, uy, uz), tl)
        if lps:
            loopon_sum(ctx, x, y, z, lps, (dx, dy, dz), img, prec, (ux, uy, uz), tp)
    if timing is not None:
        timing.upload = tl.upload + tp.upload
        timing.kernels = tl.kernels + tp.kernels
        timing.download = tl.download + tp.download
        timing.kernel_ms = tl.kernel_ms + tp.kernel_ms
        it_l, it_p = iter(tl.loop_sec), iter(tp.loop_sec)
        timing.loop_sec = [next(it_l) if l.line else next(it_p) for l in loops]
    return ux, uy, uz
