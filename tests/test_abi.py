"""The C-ABI library: it loads, exports every entry point include/*.h declares,
its struct layouts match the Python bindings, and argument errors are reported
without a device (no compute call is made here)."""
import ctypes
import os
import re
import shutil
import subprocess

import pytest

from disco_amd import backend as B

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INCLUDE = os.path.join(ROOT, "include")


def _declared():
    names = set()
    for h in sorted(os.listdir(INCLUDE)):
        if h.endswith(".h"):
            text = open(os.path.join(INCLUDE, h)).read()
            names |= set(re.findall(r"^\s*(?:int|double|long long|const char \*)\s*"
                                    r"(disco_(?:hip|geo)_\w+)\s*\(", text, flags=re.M))
    return sorted(names)


def test_header_declares_the_entry_points():
    names = _declared()
    for must in ("disco_hip_line_sum", "disco_hip_loopon_sum", "disco_hip_plan_line",
                 "disco_hip_plan_run", "disco_hip_device_count", "disco_hip_open",
                 "disco_geo_polyline", "disco_geo_polylines", "disco_geo_snap_mid_plane"):
        assert must in names


def test_library_exports_every_declared_symbol():
    lib = B.lib()
    missing = [n for n in _declared() if not hasattr(lib, n)]
    assert not missing, missing


def test_struct_layout_matches_the_bindings(tmp_path):
    cc = shutil.which("gcc") or shutil.which("cc")
    if not cc:
        pytest.skip("no C compiler")
    src = tmp_path / "layout.c"
    src.write_text('#include <stdio.h>\n#include <stddef.h>\n#include "disco_hip.h"\n'
                   'int main(void){printf("%zu %zu %zu %zu %zu %zu %zu\\n",'
                   'sizeof(disco_line_loop), offsetof(disco_line_loop, c),'
                   'sizeof(disco_loopon_loop), offsetof(disco_loopon_loop, phys),'
                   'sizeof(disco_timing), sizeof(disco_device_info),'
                   'offsetof(disco_device_info, arch));return 0;}\n')
    exe = tmp_path / "layout"
    subprocess.run([cc, "-I", INCLUDE, str(src), "-o", str(exe)], check=True)
    got = [int(v) for v in subprocess.run([str(exe)], capture_output=True, text=True).stdout.split()]
    want = [ctypes.sizeof(B._LineLoop), B._LineLoop.c.offset, ctypes.sizeof(B._LooponLoop),
            B._LooponLoop.phys.offset, ctypes.sizeof(B._Timing), ctypes.sizeof(B._DeviceInfo),
            B._DeviceInfo.arch.offset]
    assert got == want


def test_shape_layout_matches_the_geometry_binding(tmp_path):
    from disco_amd import geometry as G
    cc = shutil.which("gcc") or shutil.which("cc")
    if not cc:
        pytest.skip("no C compiler")
    src = tmp_path / "shape.c"
    src.write_text('#include <stdio.h>\n#include <stddef.h>\n#include "disco_geometry.h"\n'
                   'int main(void){printf("%zu %zu %zu %zu %zu\\n",'
                   'sizeof(disco_loop_shape), offsetof(disco_loop_shape, a0),'
                   'offsetof(disco_loop_shape, dir2), offsetof(disco_loop_shape, nvert),'
                   'offsetof(disco_loop_shape, vert));return 0;}\n')
    exe = tmp_path / "shape"
    subprocess.run([cc, "-I", INCLUDE, str(src), "-o", str(exe)], check=True)
    got = [int(v) for v in subprocess.run([str(exe)], capture_output=True, text=True).stdout.split()]
    S = G._Shape
    assert got == [ctypes.sizeof(S), S.a0.offset, S.dir2.offset, S.nvert.offset, S.vert.offset]


def test_error_codes_without_a_device():
    lib = B.lib()
    assert lib.disco_hip_strerror(0) == b"success"
    assert lib.disco_hip_strerror(2) == b"no usable HIP device"
    assert lib.disco_hip_line_fp64() == 1
    # a null context is an argument error, before any device work
    rc = lib.disco_hip_line_sum(None, 0, None, None, None, None, None, None, None, 0,
                                1.0, 1.0, 1.0, 1, 0, None)
    assert rc == 1
    assert lib.disco_hip_plan_run(None, 0, None, None, None, None, None, None, 0, None) == 1
    # lattice sizes are host arithmetic; a null context is an argument error
    assert B.lattice_atoms("FCC", 100, 100, 100) == 4_000_000
    assert B.lattice_atoms("BCC", 40, 41, 42) == 2 * 40 * 41 * 42
    assert lib.disco_hip_lattice_atoms(7, 1, 1, 1) == -1
    assert lib.disco_hip_lattice_atoms(0, -1, 1, 1) == -1
    assert lib.disco_hip_lattice(None, 0, 1, 1, 1, 1.0, None, None, None, None) == 1
    n = ctypes.c_int(-1)
    assert lib.disco_hip_device_count(ctypes.byref(n)) == 0 and n.value >= 0
    if n.value == 0:
        h = ctypes.c_void_p()
        assert lib.disco_hip_open(0, ctypes.byref(h)) == 2
        assert not B.gpu_available()
        with pytest.raises(B.DiscoError):
            B.Context(0)


def test_driver_binary_reports_usage():
    exe = os.path.join(ROOT, "disco_amd", "kroupa_gpu_driver")
    if not os.path.exists(exe):
        pytest.skip("driver not built (reference header absent at build time)")
    r = subprocess.run([exe], capture_output=True, text=True)
    assert r.returncode == 1 and "usage" in r.stderr


@pytest.mark.skipif(not os.path.exists("/root/reference/Code/backend.hpp"),
                    reason="reference tree not present")
def test_backend_hip_compiles_against_the_reference_interface(tmp_path):
    """backend_hip.cpp defines exactly the GPU entry points Code/backend.hpp
    declares under WITH_CUDA (a mismatch fails to compile or link)."""
    obj = tmp_path / "backend_hip.o"
    subprocess.run(["g++", "-std=c++14", "-DWITH_CUDA", "-I/root/reference/Code",
                    "-I", os.path.join(ROOT, "include"), "-c",
                    os.path.join(ROOT, "disco_amd", "csrc", "backend_hip.cpp"), "-o", str(obj)],
                   check=True)
    syms = subprocess.run(["nm", "-C", "--defined-only", str(obj)], capture_output=True,
                          text=True).stdout
    for sig in ("gpuAvailable()", "printGpuInfo()", "gpuLineFp64()",
                "kroupaGPU(AtomArrays const&, std::vector<LoopData, std::allocator<LoopData> > const&, double, double, double, GpuOptions const&, BackendTiming&)"):
        assert sig in syms, sig
