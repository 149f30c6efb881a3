"""Batched samples (one persistent launch per pass, HIP-graph replay) against
one-sample plans and the oracle."""
import numpy as np
import pytest

import oracle
from loop_kinds import BOX, atoms_near, kind_loops
from disco_amd import backend as B
from disco_amd import geometry as G

pytestmark = pytest.mark.gpu


def _small_batch(nsamples=10, seed=7):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(nsamples):
        s = G.config_c4_sample(rng)
        keep = np.arange(0, s.n, 7)            # smaller atom sets, uneven sizes
        out.append(s.subsample(keep))
    return out


def _concat(samples):
    import torch
    cat = lambda k: torch.from_numpy(np.concatenate([getattr(s, k) for s in samples])).cuda()
    return cat("x"), cat("y"), cat("z")


@pytest.mark.parametrize("prec", [B.PREC_FP64, B.PREC_FP32C])
def test_batch_equals_per_sample_plans(hip_ctx, prec):
    import torch
    samples = _small_batch()
    x, y, z = _concat(samples)
    u = [torch.zeros_like(x) for _ in range(3)]
    bt = B.Batch(hip_ctx, samples, 1, prec)
    bt.run_torch(x, y, z, *u)
    torch.cuda.synchronize()
    got = np.stack([t.cpu().numpy() for t in u])
    for s, off in zip(samples, bt.offsets):
        ref = np.stack(B.line_sum(hip_ctx, s.x, s.y, s.z, s.loops, s.box, 1, prec))
        # same per-atom device functions, same order: bit-identical
        assert np.array_equal(got[:, off:off + s.n], ref)
    assert bt.evals() == sum(s.n * s.segments() * 27 for s in samples)
    bt.free()


def test_batch_vs_oracle_and_graph_replays(hip_ctx):
    import torch
    samples = _small_batch(6, seed=3)
    x, y, z = _concat(samples)
    bt = B.Batch(hip_ctx, samples, 1, B.PREC_FP32C)
    u = [torch.zeros_like(x) for _ in range(3)]
    bt.run_torch(x, y, z, *u)
    torch.cuda.synchronize()
    first = np.stack([t.cpu().numpy() for t in u])
    for rep in range(3):                     # replays with eager work in between
        for t in u:
            t.fill_(123.0)
        junk = torch.randn(1000, device="cuda").sum().item()
        bt.run_torch(x, y, z, *u)
        torch.cuda.synchronize()
        again = np.stack([t.cpu().numpy() for t in u])
        assert np.array_equal(again, first), rep
        assert np.isfinite(junk)
    for s, off in zip(samples, bt.offsets):
        idx = np.arange(0, s.n, max(1, s.n // 64))
        ref = np.stack(oracle.line_sum(s.x[idx], s.y[idx], s.z[idx], s.loops, s.box, 1))
        assert np.abs(first[:, off + idx] - ref).max() < 1e-5
    bt.free()


def test_autotune_and_tunings_agree(hip_ctx):
    import torch
    samples = _small_batch(8, seed=11)
    x, y, z = _concat(samples)
    bt = B.Batch(hip_ctx, samples, 1, B.PREC_FP32C)
    u = [torch.zeros_like(x) for _ in range(3)]
    bt.run_torch(x, y, z, *u)
    torch.cuda.synchronize()
    ref = np.stack([t.cpu().numpy() for t in u])
    best, ms = bt.autotune(x, y, z, *u, stride=2)
    assert best["block"] in (128, 256) and best["tile"] % best["block"] == 0 and ms > 0
    for shape in ((64, 64, 37), (128, 512, 300), (256, 256, 2048), (best["block"], best["tile"], best["groups"])):
        bt.set_tuning(*shape)
        for t in u:
            t.zero_()
        bt.run_torch(x, y, z, *u)
        torch.cuda.synchronize()
        got = np.stack([t.cpu().numpy() for t in u])
        assert np.array_equal(got, ref), shape
    with pytest.raises(B.DiscoError):
        bt.set_tuning(96, 96, 10)
    bt.free()


def test_batch_accumulate_and_empty_samples(hip_ctx):
    import torch
    samples = _small_batch(4, seed=5)
    samples[1] = G.Sample("empty-loops", samples[1].x, samples[1].y, samples[1].z, [],
                          samples[1].box, 1)
    x, y, z = _concat(samples)
    bt = B.Batch(hip_ctx, samples, 1, B.PREC_FP64)
    u = [torch.full_like(x, 0.5) for _ in range(3)]
    bt.run_torch(x, y, z, *u, accumulate=True)
    torch.cuda.synchronize()
    got = np.stack([t.cpu().numpy() for t in u])
    off = bt.offsets[1]
    assert np.all(got[:, off:off + samples[1].n] == 0.5)       # no loops: u unchanged
    s = samples[2]
    ref = np.stack(B.line_sum(hip_ctx, s.x, s.y, s.z, s.loops, s.box, 1, B.PREC_FP64))
    o2 = bt.offsets[2]
    assert np.abs(got[:, o2:o2 + s.n] - 0.5 - ref).max() < 1e-13
    bt.free()


@pytest.mark.parametrize("prec", [B.PREC_FP64, B.PREC_FP32C])
def test_batch_loop_kinds_and_loopless_sample_overwrite(hip_ctx, prec):
    """Samples holding different loop forms (one far launch per form present)
    and a sample without loops: with accumulate = 0 every atom's u is written
    (zero for the loop-less sample), equal to one-sample plans."""
    import torch
    rng = np.random.default_rng(11)
    samples = []
    for forms in (("glide",), ("prism", "bent"), (), ("mixed", "glide")):
        loops = sum((kind_loops(k, rng, 2) for k in forms), [])
        x, y, z = atoms_near(loops, rng, nrand=700, per_loop=100) if loops else \
            tuple(rng.uniform(0, 150, 500) for _ in range(3))
        samples.append(G.Sample("kinds", x, y, z, loops, BOX, 1))
    x, y, z = _concat(samples)
    bt = B.Batch(hip_ctx, samples, 1, prec)
    u = [torch.full_like(x, 7.0) for _ in range(3)]
    for rep in range(2):                       # capture, then a graph replay
        bt.run_torch(x, y, z, *u, accumulate=False)
        torch.cuda.synchronize()
    got = np.stack([t.cpu().numpy() for t in u])
    for k, s in enumerate(samples):
        o = bt.offsets[k]
        part = got[:, o:o + s.n]
        if not s.loops:
            assert np.all(part == 0.0)
            continue
        ref = np.stack(B.line_sum(hip_ctx, s.x, s.y, s.z, s.loops, s.box, 1, prec))
        assert np.abs(part - ref).max() < 1e-12
        orc = np.stack(oracle.line_sum(s.x, s.y, s.z, s.loops, s.box, 1))
        assert np.abs(part - orc).max() < (1e-10 if prec == B.PREC_FP64 else 1e-5)
    bt.free()


def test_c4_full_batch_one_launch(hip_ctx):
    """The whole batched workload: 256 BCC samples (42.5M atoms, lattices written
    on the device) summed by one persistent launch per pass (FP32C, +-1).  Every
    16th sample is compared with its own one-sample plan (bit-identical: the same
    per-atom device functions in the same order) and with the FP64 sum (within
    1e-5 A on every atom of the sample)."""
    import torch
    specs = G.config_c4_specs()
    ntot = sum(s.n for s in specs)
    x, y, z = (torch.empty(ntot, dtype=torch.float64, device="cuda") for _ in range(3))
    off = 0
    for s in specs:
        N = s.meta["cells"]
        B.lattice_torch(hip_ctx, "BCC", N, N, N, s.meta["a0"],
                        out=[t[off:off + s.n] for t in (x, y, z)])
        off += s.n
    u = [torch.zeros_like(x) for _ in range(3)]
    bt = B.Batch(hip_ctx, specs, 1, B.PREC_FP32C)
    assert bt.n == ntot and len(bt.offsets) == 256
    bt.run_torch(x, y, z, *u)
    torch.cuda.synchronize()
    assert all(torch.isfinite(t).all().item() for t in u)
    worst = 0.0
    for k in range(0, len(specs), 16):
        s, o = specs[k], bt.offsets[k]
        xs, ys, zs = (t[o:o + s.n] for t in (x, y, z))
        mine = [t[o:o + s.n] for t in u]
        for prec in (B.PREC_FP32C, B.PREC_FP64):
            v = [torch.zeros_like(xs) for _ in range(3)]
            plan = B.Plan(hip_ctx, s.loops, s.box, 1, prec)
            plan.run_torch(xs, ys, zs, *v)
            torch.cuda.synchronize()
            plan.free()
            if prec == B.PREC_FP32C:
                assert all(torch.equal(a, b) for a, b in zip(mine, v)), k
            else:
                e = max(float((a - b).abs().max()) for a, b in zip(mine, v))
                worst = max(worst, e)
                assert e < 1e-5, (k, e)
    print(f"C4 full batch: {ntot} atoms; every 16th sample bit-equal to its own plan, "
          f"max|du| vs FP64 {worst:.3e} A")
    bt.free()
