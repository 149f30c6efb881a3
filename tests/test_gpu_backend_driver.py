"""The drop-in C++ backend (backend_hip.cpp: kroupaGPU & co. with the
reference's Code/backend.hpp signatures) driven by a main.cpp-shaped caller."""
import os
import struct
import subprocess

import numpy as np
import pytest

import oracle
from disco_amd import geometry as G

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "disco_amd", "kroupa_gpu_driver")


def _write_input(path, x, y, z, box, fp32, loops):
    with open(path, "wb") as f:
        f.write(struct.pack("<i", len(x)))
        for a in (x, y, z):
            f.write(np.ascontiguousarray(a, "<f8").tobytes())
        f.write(np.asarray(box, "<f8").tobytes())
        f.write(struct.pack("<ii", int(fp32), len(loops)))
        for lp in loops:
            f.write(struct.pack("<i", int(lp.line)))
            f.write(np.asarray(lp.phys, "<f8").tobytes())
            arrs = (lp.vx, lp.vy, lp.vz) if lp.line else (lp.x, lp.y, lp.z)
            f.write(struct.pack("<i", len(arrs[0])))
            for a in arrs:
                f.write(np.ascontiguousarray(a, "<f8").tobytes())
            if lp.line:
                f.write(np.asarray([lp.cx, lp.cy, lp.cz], "<f8").tobytes())


@pytest.mark.parametrize("fp32,tol", [(0, 1e-10), (1, 1e-5)])
def test_kroupa_gpu_driver_mixed_methods(tmp_path, fp32, tol):
    assert os.path.exists(DRIVER), "kroupa_gpu_driver was not built"
    c1 = G.config_c1()
    c3 = G.config_c3(nloops=1)
    idx = np.arange(0, c1.n, 61)[:2048]
    x, y, z = c1.x[idx], c1.y[idx], c1.z[idx]
    loops = [c1.loops[0], c3.loops[0], c1.loops[1]]          # line, loopons, line
    inp, out = tmp_path / "in.bin", tmp_path / "out.bin"
    _write_input(inp, x, y, z, c1.box, fp32, loops)
    r = subprocess.run([DRIVER, str(inp), str(out)], capture_output=True, text=True, timeout=120)
    print(r.stdout[-2000:], r.stderr[-2000:])
    assert r.returncode == 0
    assert "--- GPU:" in r.stdout and "Running GPU kernels" in r.stdout
    data = np.fromfile(out, "<f8")
    n = len(x)
    u = data[:3 * n].reshape(3, n)
    loop_sec = data[3 * n:3 * n + 3]
    ref = np.stack(oracle.line_sum(x, y, z, [loops[0], loops[2]], c1.box, 1))
    ref += np.stack(oracle.loopon_sum(x, y, z, [loops[1]], c1.box, 1))
    assert np.abs(u - ref).max() < tol
    assert loop_sec.shape == (3,) and np.all(loop_sec > 0)      # measured per loop
    upload, kernels, download = data[3 * n + 3:3 * n + 6]
    assert loop_sec.sum() <= kernels * 1.05 + 1e-3
