"""Overlapped batched datagen (datagen.OverlappedBatches): three batches through
the pipeline give, bit for bit, what each batch gives summed on its own, and the
per-batch arrays handed to the callback are the batch's lattices."""
import numpy as np
import pytest

from disco_amd import backend as B
from disco_amd import datagen as D
from disco_amd import geometry as G

pytestmark = pytest.mark.gpu


def _specs(k):
    return G.config_c4_specs(seed=100 + k, nsamples=10)


def test_overlapped_batches_equal_separate_runs(hip_ctx):
    import torch
    got = {}

    def keep(k, specs, bt, xyz, u):
        got[k] = ([t.cpu().numpy() for t in xyz], [t.cpu().numpy() for t in u], bt.offsets)

    pipe = D.OverlappedBatches(hip_ctx, _specs, 3)
    times = pipe.run(on_batch=keep)
    assert sorted(got) == [0, 1, 2] and len(times.wall_ms) == 3
    assert all(d > 0 for d in times.device_ms)
    for k in range(3):
        specs = _specs(k)
        n = sum(s.n for s in specs)
        x, y, z = (torch.empty(n, dtype=torch.float64, device="cuda") for _ in range(3))
        off = 0
        for s in specs:
            N = s.meta["cells"]
            B.lattice_torch(hip_ctx, "BCC", N, N, N, s.meta["a0"],
                            out=[t[off:off + s.n] for t in (x, y, z)])
            off += s.n
        bt = B.Batch(hip_ctx, specs, 1, B.PREC_FP32C)
        bt.set_tuning(pipe.tuning["block"], pipe.tuning["tile"], pipe.tuning["groups"])
        u = [torch.zeros_like(x) for _ in range(3)]
        bt.run_torch(x, y, z, *u)
        torch.cuda.synchronize()
        xyz, uu, offs = got[k]
        assert offs == bt.offsets
        for a, b in zip(xyz, (x, y, z)):
            assert np.array_equal(a, b.cpu().numpy())
        for a, b in zip(uu, u):
            assert np.array_equal(a, b.cpu().numpy()), k
        bt.free()
