"""CPU oracle for the displacement hot path -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline``
leg may import this package, and only as the checker.  The product
(``disco_amd``) never imports it and has no CPU fallback.

Two libraries are wrapped:

* ``liboracle.so`` -- the repo's FP64 restatement (``disco_oracle.c``), built
  here and shipped with the tree;
* ``_ref/libdisco_ref.so`` -- the reference's own ``Code/backend_cpu.cpp`` and
  ``Code/loop.cpp`` compiled where they lie (``oracle/Makefile``, target
  ``ref``).  Present only where the reference tree was available at build time.

Parity pins of each oracle function (checked in ``tests/test_oracle_*.py``):

============================  =================================================
oracle function               pinned by
============================  =================================================
barnett_loop (FP64)           golden vectors of barnettLoop<double> (reference
                              build); Dunavant quadrature of the Kroupa kernel;
                              jump -b across the cut; apex / start invariance;
                              on-axis solid angle of a regular pyramid; disc
                              limit; linearity in b; collinear subdivision
kroupa_pair (FP64)            golden vectors of kroupaPair; oddness in rho;
                              rho < 1e-10 -> 0; loopon sum -> line integral
line_sum (+-img images)       golden kroupaCPU output at +-1 (4,096-atom C1
                              subsample); +-N = sum of single-image fields
loopon_sum (+-img images)     golden kroupaCPU output at +-1 (C3 loop, 4,096
                              atoms); +-N = sum of single-image fields
line_sum_ref32                golden barnettLoop<float> vectors (the float
                              instantiation the reference's --fp32 uses)
prismatic_slab, matter_type   the compiled reference's prismaticSlab on every
                              loop shape (rounded corners, axes, polygons);
                              golden slabs; brute force in exact arithmetic
classify_burgers,             the compiled reference's classifyBurgers,
frank_copy_shift, copy_shifts frankCopyShift and rotations on random and
                              enumerated inputs; the reference unit test's
                              Frank-copy checks (in-plane, |s| = a0/sqrt6, the
                              third stacking position)
prismatic_apply,              invariants of the driver's bookkeeping: kept
append_copies                 atoms = the complement of the filtered vacancy
                              slabs, in order; copies = the filtered
                              interstitial entries in loop then slab order,
                              renumbered, those of deleted atoms dropped;
                              copies at their source's site with u + shift
final_positions               exact rational arithmetic of the reference's
                              expression shapes under GCC contraction; wrap
                              closed forms; identity U leaves wrapped + L/2
lammps_box                    exact rational Gram-Schmidt; orthogonal boxes
                              give (dx, dy, dz) and the unit frame
fixed8, write_xyz,            Python's correctly rounded '%.8f' (an
write_lammps                  independent implementation) on random values,
                              exact ties, -0.0, tiny negatives, nan/inf; the
                              file layout line by line
============================  =================================================
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(_HERE, "liboracle.so")
OUT_SO = os.path.join(_HERE, "liboracle_out.so")
REF_SO = os.path.join(_HERE, "_ref", "libdisco_ref.so")

_dp = ctypes.POINTER(ctypes.c_double)
_fp = ctypes.POINTER(ctypes.c_float)
_ip = ctypes.POINTER(ctypes.c_int)
_d = ctypes.c_double
_i = ctypes.c_int
_f = ctypes.c_float

_lib = None
_ref = None
_out = None


def _ptr(a, kind=_dp):
    return a.ctypes.data_as(kind)


def lib():
    """The FP64 restatement (raises if it was not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_SO):
            raise RuntimeError(f"oracle library missing: {ORACLE_SO} (run `make -C oracle`)")
        L = ctypes.CDLL(ORACLE_SO)
        L.oracle_threads.restype = _i
        L.oracle_k1.restype = _d
        L.oracle_barnett_loop.argtypes = [_d, _d, _d, _dp, _dp, _dp, _i, _d, _d, _d, _d, _d, _d, _dp]
        L.oracle_barnett_loop_f.argtypes = [_f, _f, _f, _fp, _fp, _fp, _i, _f, _f, _f, _f, _f, _f, _fp]
        L.oracle_kroupa_pair.argtypes = [_d, _d, _d, _dp, _d, _d, _dp]
        L.oracle_line_sum.argtypes = [_i, _dp, _dp, _dp, _dp, _dp, _dp, _i, _ip, _dp, _dp, _dp,
                                      _dp, _dp, _d, _d, _d, _i]
        L.oracle_loopon_sum.argtypes = [_i, _dp, _dp, _dp, _dp, _dp, _dp, _i, _ip, _dp, _dp, _dp,
                                        _dp, _d, _d, _d, _i]
        L.oracle_line_sum_ref32.argtypes = [_i, _dp, _dp, _dp, _dp, _dp, _dp, _i, _ip, _dp, _dp,
                                            _dp, _dp, _dp, _d, _d, _d, _i, _fp]
        L.oracle_matter_type.argtypes = [ctypes.c_void_p]
        L.oracle_matter_type.restype = _i
        L.oracle_prismatic_slab.argtypes = [ctypes.c_void_p, _d, ctypes.c_longlong, _dp, _dp,
                                            _dp, _d, _d, _d, _ip, _ip, ctypes.c_longlong, _dp]
        L.oracle_prismatic_slab.restype = ctypes.c_longlong
        L.oracle_classify_burgers.argtypes = [_i, _dp, _dp]
        L.oracle_classify_burgers.restype = _i
        L.oracle_frank_copy_shift.argtypes = [_dp, _dp, _d, _dp]
        L.oracle_copy_shifts.argtypes = [_dp, _i, _dp, ctypes.c_void_p, _d, _dp]
        _llp = ctypes.POINTER(ctypes.c_longlong)
        L.oracle_prismatic_apply.argtypes = [ctypes.c_longlong, _dp, _dp, _dp, ctypes.c_void_p,
                                             ctypes.c_void_p, _i, _ip, _llp, _ip, _ip, _ip, _ip,
                                             _dp, _ip, _dp, _llp]
        L.oracle_prismatic_apply.restype = ctypes.c_longlong
        L.oracle_append_copies.argtypes = [ctypes.c_longlong, ctypes.c_longlong, _ip, _dp, _dp,
                                           _dp, _dp, _dp, _dp, _dp, ctypes.c_void_p]
        _lib = L
    return _lib


def ref_available() -> bool:
    return os.path.exists(REF_SO)


def ref():
    """The reference's own sources, compiled (raises if not built here)."""
    global _ref
    if _ref is None:
        if not ref_available():
            raise RuntimeError(f"reference build missing: {REF_SO} (run `make -C oracle ref`)")
        R = ctypes.CDLL(REF_SO)
        R.ref_barnett_loop.argtypes = [_d, _d, _d, _dp, _dp, _dp, _i, _d, _d, _d, _d, _d, _d, _dp]
        R.ref_barnett_loop_f.argtypes = [_f, _f, _f, _fp, _fp, _fp, _i, _f, _f, _f, _f, _f, _f, _fp]
        R.ref_kroupa_pair.argtypes = [_d, _d, _d, _dp, _d, _d, _dp]
        R.ref_kroupa_k1.restype = _d
        R.ref_kroupa_cpu_line.argtypes = [_i, _dp, _dp, _dp, _dp, _dp, _dp, _i, _ip, _dp, _dp, _dp,
                                          _dp, _dp, _d, _d, _d, _dp]
        R.ref_kroupa_cpu_loopon.argtypes = [_i, _dp, _dp, _dp, _dp, _dp, _dp, _i, _ip, _dp, _dp,
                                            _dp, _dp, _d, _d, _d, _dp]
        R.ref_loop_polyline.argtypes = [_i, _dp, _dp, _dp, _dp, _d, _d, _dp, _i]
        R.ref_loop_polyline.restype = _i
        R.ref_loop_phys.argtypes = [_i, _dp, _d, _dp, _dp, _dp, _d, _dp, _dp]
        R.ref_loop_loopons.argtypes = [_i, _dp, _d, _dp, _dp, _dp, _d, _dp, _i]
        R.ref_loop_loopons.restype = _i
        R.ref_regular_polygon.argtypes = [_i, _d, _dp, _dp, _dp]
        R.ref_snap_mid_plane.argtypes = [_dp, _dp, _d, _i, _i, _i, _i, _dp, _dp]
        R.ref_lattice.argtypes = [_i, _i, _i, _i, _d, _dp, _dp, _dp, _i]
        R.ref_lattice.restype = _i
        R.ref_prismatic_slab.argtypes = [_i, _dp, _dp, _dp, _dp, _d, _d, _i, _dp, _dp, _dp, _d,
                                         _d, _d, _ip, _ip, _i, _dp]
        R.ref_prismatic_slab.restype = _i
        R.ref_classify_burgers.argtypes = [_i, _dp, _dp]
        R.ref_classify_burgers.restype = _i
        R.ref_frank_copy_shift.argtypes = [_dp, _dp, _d, _dp]
        R.ref_rotate.argtypes = [_dp, _i, _dp, _dp]
        _ref = R
    return _ref


def threads() -> int:
    return lib().oracle_threads()


def k1() -> float:
    return lib().oracle_k1()


# ---------------------------------------------------------------- pair terms
def _as3(v):
    return np.ascontiguousarray(v, dtype=np.float64)


def barnett_loop(p, V, c, b, use_ref=False):
    """Field at p of the closed polyline V (nv x 3), apex c, Burgers b (FP64)."""
    V = np.ascontiguousarray(np.asarray(V, dtype=np.float64).T)
    u = np.zeros(3)
    fn = ref().ref_barnett_loop if use_ref else lib().oracle_barnett_loop
    fn(float(p[0]), float(p[1]), float(p[2]), _ptr(V[0]), _ptr(V[1]), _ptr(V[2]), V.shape[1],
       float(c[0]), float(c[1]), float(c[2]), float(b[0]), float(b[1]), float(b[2]), _ptr(u))
    return u


def barnett_loop_f(p, V, c, b, use_ref=False):
    """The float instantiation (R = float) of the same evaluation."""
    V = np.ascontiguousarray(np.asarray(V, dtype=np.float32).T)
    u = np.zeros(3, dtype=np.float32)
    fn = ref().ref_barnett_loop_f if use_ref else lib().oracle_barnett_loop_f
    fn(*[float(np.float32(t)) for t in p], _ptr(V[0], _fp), _ptr(V[1], _fp), _ptr(V[2], _fp),
       V.shape[1], *[float(np.float32(t)) for t in c], *[float(np.float32(t)) for t in b],
       _ptr(u, _fp))
    return u


def kroupa_pair(rho, phys, use_ref=False):
    """One loopon image at separation rho; phys = (b3, m3, n3, dS)."""
    ph = _as3(phys)
    kk = k1()
    b_n = float(ph[3] * ph[6] + ph[4] * ph[7] + ph[5] * ph[8])
    u = np.zeros(3)
    fn = ref().ref_kroupa_pair if use_ref else lib().oracle_kroupa_pair
    fn(float(rho[0]), float(rho[1]), float(rho[2]), _ptr(ph), kk, b_n, _ptr(u))
    return u


# ---------------------------------------------------------------- packed sums
def pack_line(loops):
    """(voff, vx, vy, vz, c, b, phys) arrays from a list of line LoopData."""
    voff = np.zeros(len(loops) + 1, dtype=np.int32)
    for i, lp in enumerate(loops):
        voff[i + 1] = voff[i] + len(lp.vx)
    vx = np.concatenate([np.asarray(lp.vx, np.float64) for lp in loops]) if loops else np.zeros(0)
    vy = np.concatenate([np.asarray(lp.vy, np.float64) for lp in loops]) if loops else np.zeros(0)
    vz = np.concatenate([np.asarray(lp.vz, np.float64) for lp in loops]) if loops else np.zeros(0)
    c = np.array([[lp.cx, lp.cy, lp.cz] for lp in loops], np.float64).reshape(-1)
    b = np.array([lp.burgers() for lp in loops], np.float64).reshape(-1)
    phys = np.array([lp.phys for lp in loops], np.float64).reshape(-1)
    return voff, vx, vy, vz, c, b, phys


def pack_loopon(loops):
    poff = np.zeros(len(loops) + 1, dtype=np.int32)
    for i, lp in enumerate(loops):
        poff[i + 1] = poff[i] + len(lp.x)
    px = np.concatenate([np.asarray(lp.x, np.float64) for lp in loops])
    py = np.concatenate([np.asarray(lp.y, np.float64) for lp in loops])
    pz = np.concatenate([np.asarray(lp.z, np.float64) for lp in loops])
    phys = np.array([lp.phys for lp in loops], np.float64).reshape(-1)
    return poff, px, py, pz, phys


def _atoms(x, y, z):
    return (np.ascontiguousarray(x, np.float64), np.ascontiguousarray(y, np.float64),
            np.ascontiguousarray(z, np.float64))


def line_sum(x, y, z, loops, box, img=1, u=None):
    """Sum over loops x images (+-img) of the line field, FP64 (returns ux, uy, uz)."""
    x, y, z = _atoms(x, y, z)
    n = x.shape[0]
    ux, uy, uz = (np.zeros(n), np.zeros(n), np.zeros(n)) if u is None else u
    voff, vx, vy, vz, c, b, _ = pack_line(loops)
    lib().oracle_line_sum(n, _ptr(x), _ptr(y), _ptr(z), _ptr(ux), _ptr(uy), _ptr(uz),
                          len(loops), _ptr(voff, _ip), _ptr(vx), _ptr(vy), _ptr(vz), _ptr(c),
                          _ptr(b), float(box[0]), float(box[1]), float(box[2]), int(img))
    return ux, uy, uz


def line_sum_ref32(x, y, z, loops, box, img=1):
    """The reference's --fp32 line scheme restated on the CPU (yardstick)."""
    x, y, z = _atoms(x, y, z)
    n = x.shape[0]
    ux, uy, uz = np.zeros(n), np.zeros(n), np.zeros(n)
    voff, vx, vy, vz, c, b, _ = pack_line(loops)
    scratch = np.zeros(3 * max(1, int(np.max(np.diff(voff))) if len(loops) else 1), np.float32)
    lib().oracle_line_sum_ref32(n, _ptr(x), _ptr(y), _ptr(z), _ptr(ux), _ptr(uy), _ptr(uz),
                                len(loops), _ptr(voff, _ip), _ptr(vx), _ptr(vy), _ptr(vz),
                                _ptr(c), _ptr(b), float(box[0]), float(box[1]), float(box[2]),
                                int(img), _ptr(scratch, _fp))
    return ux, uy, uz


def loopon_sum(x, y, z, loops, box, img=1, u=None):
    """Sum over loops x images (+-img) x loopons of the Kroupa term, FP64."""
    x, y, z = _atoms(x, y, z)
    n = x.shape[0]
    ux, uy, uz = (np.zeros(n), np.zeros(n), np.zeros(n)) if u is None else u
    poff, px, py, pz, phys = pack_loopon(loops)
    lib().oracle_loopon_sum(n, _ptr(x), _ptr(y), _ptr(z), _ptr(ux), _ptr(uy), _ptr(uz),
                            len(loops), _ptr(poff, _ip), _ptr(px), _ptr(py), _ptr(pz),
                            _ptr(phys), float(box[0]), float(box[1]), float(box[2]), int(img))
    return ux, uy, uz


# ---------------------------------------------------------------- reference calls
def ref_kroupa_cpu_line(x, y, z, loops, box):
    """The reference's kroupaCPU (line loops, +-1 images)."""
    x, y, z = _atoms(x, y, z)
    n = x.shape[0]
    ux, uy, uz = np.zeros(n), np.zeros(n), np.zeros(n)
    voff, vx, vy, vz, c, _, phys = pack_line(loops)
    sec = np.zeros(len(loops))
    ref().ref_kroupa_cpu_line(n, _ptr(x), _ptr(y), _ptr(z), _ptr(ux), _ptr(uy), _ptr(uz),
                              len(loops), _ptr(voff, _ip), _ptr(vx), _ptr(vy), _ptr(vz), _ptr(c),
                              _ptr(phys), float(box[0]), float(box[1]), float(box[2]), _ptr(sec))
    return ux, uy, uz, sec


def ref_kroupa_cpu_loopon(x, y, z, loops, box):
    """The reference's kroupaCPU (loopon loops, +-1 images)."""
    x, y, z = _atoms(x, y, z)
    n = x.shape[0]
    ux, uy, uz = np.zeros(n), np.zeros(n), np.zeros(n)
    poff, px, py, pz, phys = pack_loopon(loops)
    sec = np.zeros(len(loops))
    ref().ref_kroupa_cpu_loopon(n, _ptr(x), _ptr(y), _ptr(z), _ptr(ux), _ptr(uy), _ptr(uz),
                                len(loops), _ptr(poff, _ip), _ptr(px), _ptr(py), _ptr(pz),
                                _ptr(phys), float(box[0]), float(box[1]), float(box[2]), _ptr(sec))
    return ux, uy, uz, sec


SHAPE_CODE = {"circle": 0, "rect": 1, "parallelogram": 2, "polygon": 3, "polygon_vertices": 4}


def _params(params):
    """Shape parameters, zero-padded (optional trailing directions read as empty)."""
    p = np.zeros(max(16, len(params) + 8))
    p[:len(params)] = params
    return p


def ref_polyline(shape, params, c, m, n, a0, eps, cap=4096):
    out = np.zeros(3 * cap)
    p = _params(params)
    nv = ref().ref_loop_polyline(SHAPE_CODE[shape], _ptr(p), _ptr(_as3(c)), _ptr(_as3(m)),
                                 _ptr(_as3(n)), a0, eps, _ptr(out), cap)
    return out[:3 * min(nv, cap)].reshape(-1, 3)


def ref_loop_phys(shape, params, da, c, m, n, a0):
    phys = np.zeros(10)
    area = np.zeros(1)
    ref().ref_loop_phys(SHAPE_CODE[shape], _ptr(_params(params)), da, _ptr(_as3(c)), _ptr(_as3(m)),
                        _ptr(_as3(n)), a0, _ptr(phys), _ptr(area))
    return phys, float(area[0])


def ref_loopons(shape, params, da, c, m, n, a0, cap=200000):
    out = np.zeros(3 * cap)
    npn = ref().ref_loop_loopons(SHAPE_CODE[shape], _ptr(_params(params)), da, _ptr(_as3(c)),
                                 _ptr(_as3(m)), _ptr(_as3(n)), a0, _ptr(out), cap)
    return out[:3 * min(npn, cap)].reshape(-1, 3)


def ref_regular_polygon(sides, R, m, n):
    out = np.zeros(3 * sides)
    ref().ref_regular_polygon(sides, R, _ptr(_as3(m)), _ptr(_as3(n)), _ptr(out))
    return out.reshape(-1, 3)


def ref_snap(center, n, a0, nx, ny, nz, lattice):
    out = np.zeros(3)
    sp = np.zeros(1)
    ref().ref_snap_mid_plane(_ptr(_as3(center)), _ptr(_as3(n)), a0, nx, ny, nz,
                             1 if lattice == "FCC" else 0, _ptr(out), _ptr(sp))
    return out, float(sp[0])


def ref_lattice(lattice, nx, ny, nz, a0):
    cap = (4 if lattice == "FCC" else 2) * nx * ny * nz
    x, y, z = np.zeros(cap), np.zeros(cap), np.zeros(cap)
    ref().ref_lattice(1 if lattice == "FCC" else 0, nx, ny, nz, a0, _ptr(x), _ptr(y), _ptr(z), cap)
    return x, y, z


def prismatic_slab(state, spacing, x, y, z, box):
    """prismaticSlab restated (Code/prismatic.cpp:30-74) for a constructed loop
    state (a disco_loop_state ctypes structure): (atoms, side, info), info =
    (type, planes, height, spacing, b0, b1, b2); raises ValueError where the
    reference aborts (|b.n| not a whole number of plane spacings)."""
    x, y, z = _atoms(x, y, z)
    cap = x.shape[0]
    atoms, side = np.zeros(cap, dtype=np.int32), np.zeros(cap, dtype=np.int32)
    info = np.zeros(7)
    ns = lib().oracle_prismatic_slab(ctypes.addressof(state), float(spacing), cap, _ptr(x),
                                     _ptr(y), _ptr(z), float(box[0]), float(box[1]),
                                     float(box[2]), _ptr(atoms, _ip), _ptr(side, _ip), cap,
                                     _ptr(info))
    if ns < 0:
        raise ValueError("|b.n| is not a whole number of atomic-plane spacings")
    return atoms[:ns], side[:ns], info


def matter_type(state) -> int:
    """matterType (Code/prismatic.cpp:12-18): 0 glide, 1 vacancy, 2 interstitial."""
    return lib().oracle_matter_type(ctypes.addressof(state))


def ref_prismatic_slab(shape, params, c, m, n, a0, spacing, x, y, z, box):
    """The reference's prismaticSlab: (atoms, side, info) with info = (type, planes,
    height, spacing, b0, b1, b2), type 0 glide, 1 vacancy, 2 interstitial."""
    x, y, z = _atoms(x, y, z)
    cap = x.shape[0]
    atoms, side = np.zeros(cap, dtype=np.int32), np.zeros(cap, dtype=np.int32)
    info = np.zeros(7)
    ns = ref().ref_prismatic_slab(SHAPE_CODE[shape], _ptr(_params(params)), _ptr(_as3(c)),
                                  _ptr(_as3(m)), _ptr(_as3(n)), a0, spacing, cap, _ptr(x),
                                  _ptr(y), _ptr(z), float(box[0]), float(box[1]), float(box[2]),
                                  _ptr(atoms, _ip), _ptr(side, _ip), cap, _ptr(info))
    return atoms[:ns], side[:ns], info


def classify_burgers(lattice, m, n) -> int:
    """classifyBurgers restated (Code/lattice.hpp:189-212): 0 perfect, 1 Frank,
    2 Shockley, 3 other; lattice "BCC" / "FCC"."""
    return lib().oracle_classify_burgers(1 if lattice == "FCC" else 0, _ptr(_as3(m)),
                                         _ptr(_as3(n)))


def frank_copy_shift(v, nhat, a0):
    """frankCopyShift restated (Code/lattice.hpp:239-256)."""
    out = np.zeros(3)
    lib().oracle_frank_copy_shift(_ptr(_as3(v)), _ptr(_as3(nhat)), float(a0), _ptr(out))
    return out


def copy_shifts(b, frank, n, R, a0):
    """The two copy shifts of an interstitial loop (Code/main.cpp:633-650):
    shape (2, 3), row 0 for side +1; R a 3x3 orientation (rows: crystal
    directions of the box axes) or None."""
    out = np.zeros(6)
    Rm = None if R is None else np.ascontiguousarray(np.asarray(R, dtype=np.float64).reshape(9))
    lib().oracle_copy_shifts(_ptr(_as3(b)), int(bool(frank)), _ptr(_as3(n)),
                             None if Rm is None else Rm.ctypes.data, float(a0), _ptr(out))
    return out.reshape(2, 3)


def prismatic_apply(x, y, z, grain, twin, slabs):
    """The driver's use of the slabs (Code/main.cpp:593-687), restated.

    slabs: per loop, dict(type=0|1|2, atoms, side, grain=int (< 0 none),
    region=int, shifts=(2, 3) array).  Returns (x, y, z, grain) compacted (grain
    None when not given) and (copy_atom, copy_shift (k, 3))."""
    x, y, z = (np.array(a, dtype=np.float64) for a in (x, y, z))
    n = x.shape[0]
    if grain is None and any(s.get("grain", -1) >= 0 for s in slabs):
        raise ValueError("a loop names a grain but the atoms have no grain labels")
    g = None if grain is None else np.array(grain, dtype=np.int32)
    tw = None if twin is None else np.ascontiguousarray(twin, dtype=np.int32)
    nl = len(slabs)
    typ = np.array([s["type"] for s in slabs] or [0], dtype=np.int32)
    off = np.zeros(nl + 1, dtype=np.int64)
    for l, s in enumerate(slabs):
        off[l + 1] = off[l] + len(s["atoms"])
    atoms = np.ascontiguousarray(np.concatenate([np.asarray(s["atoms"], dtype=np.int32)
                                                 for s in slabs] + [np.zeros(1, np.int32)]))
    side = np.ascontiguousarray(np.concatenate([np.asarray(s["side"], dtype=np.int32)
                                                for s in slabs] + [np.zeros(1, np.int32)]))
    lg = np.array([s.get("grain", -1) for s in slabs] or [0], dtype=np.int32)
    lr = np.array([s.get("region", 0) for s in slabs] or [0], dtype=np.int32)
    sh = np.ascontiguousarray(np.concatenate([np.asarray(s.get("shifts", np.zeros((2, 3))),
                                                         dtype=np.float64).reshape(6)
                                              for s in slabs] + [np.zeros(6)]))
    cap = max(int(off[-1]), 1)
    ca, cs = np.zeros(cap, dtype=np.int32), np.zeros(3 * cap)
    nc = ctypes.c_longlong(0)
    k = lib().oracle_prismatic_apply(
        n, _ptr(x), _ptr(y), _ptr(z), None if g is None else g.ctypes.data,
        None if tw is None else tw.ctypes.data, nl, _ptr(typ, _ip),
        off.ctypes.data_as(ctypes.POINTER(ctypes.c_longlong)), _ptr(atoms, _ip), _ptr(side, _ip),
        _ptr(lg, _ip), _ptr(lr, _ip), _ptr(sh), _ptr(ca, _ip), _ptr(cs), ctypes.byref(nc))
    if k < 0:
        raise MemoryError("oracle_prismatic_apply")
    c = nc.value
    return ((x[:k], y[:k], z[:k], None if g is None else g[:k]),
            (ca[:c].copy(), cs[:3 * c].reshape(c, 3).copy()))


def append_copies(x, y, z, ux, uy, uz, grain, copy_atom, copy_shift):
    """The copies appended after the sum (Code/main.cpp:733-742), restated:
    new arrays of n + ncopies entries."""
    n, c = len(x), len(copy_atom)
    out = [np.concatenate([np.asarray(a, dtype=np.float64), np.zeros(c)])
           for a in (x, y, z, ux, uy, uz)]
    g = None if grain is None else np.concatenate([np.asarray(grain, dtype=np.int32),
                                                   np.zeros(c, dtype=np.int32)])
    ca = np.ascontiguousarray(copy_atom, dtype=np.int32)
    cs = np.ascontiguousarray(np.asarray(copy_shift, dtype=np.float64).reshape(-1))
    lib().oracle_append_copies(n, c, _ptr(ca, _ip), _ptr(cs), *(_ptr(a) for a in out),
                               None if g is None else g.ctypes.data)
    return out, g


def ref_classify_burgers(lattice, m, n) -> int:
    return ref().ref_classify_burgers(1 if lattice == "FCC" else 0, _ptr(_as3(m)), _ptr(_as3(n)))


def ref_frank_copy_shift(v, nhat, a0):
    out = np.zeros(3)
    ref().ref_frank_copy_shift(_ptr(_as3(v)), _ptr(_as3(nhat)), float(a0), _ptr(out))
    return out


def ref_rotate(R, to_lab, v):
    out = np.zeros(3)
    Rm = np.ascontiguousarray(np.asarray(R, dtype=np.float64).reshape(9))
    ref().ref_rotate(_ptr(Rm), int(bool(to_lab)), _ptr(_as3(v)), _ptr(out))
    return out


# ---------------------------------------------------------------- output stage
def out_lib():
    """The output-stage restatement (raises if it was not built)."""
    global _out
    if _out is None:
        if not os.path.exists(OUT_SO):
            raise RuntimeError(f"oracle library missing: {OUT_SO} (run `make -C oracle`)")
        L = ctypes.CDLL(OUT_SO)
        _ll = ctypes.c_longlong
        L.oracle_final_positions.argtypes = [_ll, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp,
                                             _dp, _dp]
        L.oracle_lammps_box.argtypes = [_dp, _dp, _dp]
        L.oracle_write_xyz.argtypes = [ctypes.c_char_p, _ll, _dp, _dp, _dp, _dp,
                                       ctypes.c_char_p, _ip, _ip]
        L.oracle_write_xyz.restype = _i
        L.oracle_write_lammps.argtypes = [ctypes.c_char_p, _ll, _dp, _dp, _dp, _dp,
                                          ctypes.c_char_p, _d]
        L.oracle_write_lammps.restype = _i
        L.oracle_fixed8.argtypes = [_d, ctypes.c_char_p, _i]
        L.oracle_fixed8.restype = _i
        _out = L
    return _out


def _c(a, dtype=np.float64):
    return np.ascontiguousarray(a, dtype=dtype)


def final_positions(x, y, z, ux, uy, uz, box, U=None):
    """Code/main.cpp:797-818: x + u wrapped into the box centred at the origin,
    shifted by +L/2, times U (row-major; identity by default)."""
    x, y, z, ux, uy, uz = (_c(a) for a in (x, y, z, ux, uy, uz))
    U = _c(np.eye(3) if U is None else U).reshape(9)
    box = _c(box)
    f = [np.empty_like(x) for _ in range(3)]
    out_lib().oracle_final_positions(x.size, *(_ptr(a) for a in (x, y, z, ux, uy, uz)),
                                     _ptr(box), _ptr(U), *(_ptr(a) for a in f))
    return f


def lammps_box(Hp):
    """Code/main.cpp:64-86: (lx, ly, lz, xy, xz, yz) and the frame rows e1, e2, e3."""
    Hp = _c(Hp).reshape(9)
    b, fr = np.empty(6), np.empty(9)
    out_lib().oracle_lammps_box(_ptr(Hp), _ptr(b), _ptr(fr))
    return b, fr.reshape(3, 3)


def write_xyz(path, fx, fy, fz, Hp, element, grain=None, twin=None):
    fx, fy, fz = _c(fx), _c(fy), _c(fz)
    g = None if grain is None else _c(grain, np.int32)
    t = None if twin is None else _c(twin, np.int32)
    r = out_lib().oracle_write_xyz(str(path).encode(), fx.size, _ptr(fx), _ptr(fy), _ptr(fz),
                                   _ptr(_c(Hp).reshape(9)), element.encode(),
                                   None if g is None else _ptr(g, _ip),
                                   None if t is None else _ptr(t, _ip))
    if r:
        raise OSError(f"oracle could not write {path}")


def write_lammps(path, fx, fy, fz, Hp, element, mass):
    fx, fy, fz = _c(fx), _c(fy), _c(fz)
    r = out_lib().oracle_write_lammps(str(path).encode(), fx.size, _ptr(fx), _ptr(fy), _ptr(fz),
                                      _ptr(_c(Hp).reshape(9)), element.encode(), float(mass))
    if r:
        raise OSError(f"oracle could not write {path}")


def fixed8(v) -> str:
    """One value as the writers print it (iostream fixed, 8 decimals)."""
    buf = ctypes.create_string_buffer(400)
    n = out_lib().oracle_fixed8(float(v), buf, 400)
    if n < 0:
        raise ValueError("value too long")
    return buf.value.decode()
