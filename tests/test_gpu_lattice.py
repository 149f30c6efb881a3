"""On-device lattice generation (generateLattice, Code/lattice.hpp:35-51) through
the C ABI: bit-identical to the host restatement, which the CPU tests pin to
the compiled reference (tests/test_geometry.py)."""
import numpy as np
import pytest

from disco_amd import backend as B
from disco_amd import geometry as G

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("kind,n,a0", [("BCC", (3, 4, 5), 2.8553), ("FCC", (5, 3, 4), 3.615),
                                       ("BCC", (40, 40, 40), 2.8553),
                                       ("BCC", (60, 60, 60), 2.8553),
                                       ("FCC", (100, 100, 100), 3.615),
                                       ("BCC", (60, 1, 37), 2.8665), ("FCC", (1, 1, 1), 4.05)])
def test_lattice_bit_identical(hip_ctx, kind, n, a0):
    import torch
    x, y, z = B.lattice_torch(hip_ctx, kind, *n, a0)
    torch.cuda.synchronize()
    rx, ry, rz = G.lattice(kind, *n, a0)
    assert x.numel() == rx.size == B.lattice_atoms(kind, *n)
    for got, want in ((x, rx), (y, ry), (z, rz)):
        assert np.array_equal(got.cpu().numpy(), want)


def test_lattice_edge_cases(hip_ctx):
    import torch
    assert B.lattice_atoms("FCC", 0, 5, 5) == 0
    x, y, z = B.lattice_torch(hip_ctx, "FCC", 0, 5, 5, 3.6)       # empty: nothing written
    assert x.numel() == 0
    sentinel = [torch.full((100,), 7.0, dtype=torch.float64, device="cuda") for _ in range(3)]
    B.lattice_torch(hip_ctx, "BCC", 2, 3, 4, 2.9, out=sentinel)   # 48 sites into 100
    torch.cuda.synchronize()
    assert torch.all(sentinel[0][48:] == 7.0)
    assert np.array_equal(sentinel[2][:48].cpu().numpy(), G.lattice("BCC", 2, 3, 4, 2.9)[2])
    with pytest.raises(B.DiscoError):
        B.lattice_torch(hip_ctx, "BCC", 70000, 1, 1, 2.9)
    with pytest.raises(B.DiscoError):
        B.lattice_torch(hip_ctx, "FCC", 2, 2, 2, float("nan"))
    small = [torch.zeros(10, dtype=torch.float64, device="cuda") for _ in range(3)]
    with pytest.raises(B.DiscoError):
        B.lattice_torch(hip_ctx, "FCC", 2, 2, 2, 3.6, out=small)   # 32 sites do not fit
