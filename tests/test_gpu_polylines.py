"""GPU parity on closed polylines that are not simple flat polygons (off their
plane, winding twice, lobes of opposite sense, apex off the plane: the shapes of
polylines.py), which Code/barnett.hpp:44-107 sums like any other: through the C
ABI against the FP64 oracle, in plans and in batches, and against what the cut
fixes (moving the apex of a cone adds b inside it).  Tolerances: FP64 1e-10 A,
FP32C 1e-5 A."""
import numpy as np
import pytest

import oracle
import polylines as PL
from disco_amd import backend as B
from disco_amd import geometry as G
from loop_kinds import kind_loops

pytestmark = pytest.mark.gpu
TOL = {B.PREC_FP64: 1e-10, B.PREC_FP32C: 1e-5}
PRECS = [B.PREC_FP64, B.PREC_FP32C]


@pytest.mark.parametrize("prec", PRECS)
@pytest.mark.parametrize("name", PL.SHAPES)
def test_polyline_vs_oracle(hip_ctx, name, prec):
    rng = np.random.default_rng(PL.SHAPES.index(name) + 60)
    loops = PL.loops(rng, (name, name))
    x, y, z = PL.atoms(loops, rng)
    ref = np.stack(oracle.line_sum(x, y, z, loops, PL.BOX, 1))
    got = np.stack(B.line_sum(hip_ctx, x, y, z, loops, PL.BOX, 1, prec))
    err = np.abs(got - ref).max()
    print(f"{name} prec {prec}: max |du| {err:.2e}, |u|max {np.abs(ref).max():.2f}")
    assert err < TOL[prec]


@pytest.mark.parametrize("prec", PRECS)
def test_cone_apex_move_adds_b_inside(hip_ctx, prec):
    """The same cone fanned from its tip and from its base's centre: u differs by
    +-b inside the solid cone (a solid angle beyond 2 pi there) and by 0 outside."""
    rng = np.random.default_rng(61)
    V, apex, n, _, _, R = PL.shape("cone", rng)
    b = PL.burgers(rng)
    P_in, P_out = PL.inside_cone(V, apex, n, R, rng, 400)
    P = np.concatenate([P_in, P_out])
    x, y, z = (P[:, k].copy() for k in range(3))
    tip = np.stack(B.line_sum(hip_ctx, x, y, z, [PL.loop_data(V, apex, b, n)], PL.BOX, 0, prec))
    base = np.stack(B.line_sum(hip_ctx, x, y, z, [PL.loop_data(V, V.mean(0), b, n)], PL.BOX, 0,
                               prec))
    d = (tip - base).T
    k = len(P_in)
    sgn = np.sign(d[0] @ b)
    assert np.abs(d[:k] - sgn * b).max() < TOL[prec] * 10
    assert np.abs(d[k:]).max() < TOL[prec] * 10


@pytest.mark.parametrize("prec", PRECS)
def test_polylines_mixed_with_flat_loops_plan_and_batch(hip_ctx, prec):
    """One plan holding every shape beside flat loops of each frame form, and a
    batch of samples holding them: equal to the oracle, and the batch equal to
    one-sample plans."""
    import torch
    rng = np.random.default_rng(62)
    loops = PL.loops(rng) + sum((kind_loops(k, rng, 1) for k in ("glide", "prism", "mixed")), [])
    rng.shuffle(loops)
    x, y, z = PL.atoms(loops, rng, nrand=1000, per_loop=80)
    ref = np.stack(oracle.line_sum(x, y, z, loops, PL.BOX, 1))
    got = np.stack(B.line_sum(hip_ctx, x, y, z, loops, PL.BOX, 1, prec))
    assert np.abs(got - ref).max() < TOL[prec]
    samples = []
    for part in (loops[:4], loops[4:7], loops[7:]):
        sx, sy, sz = PL.atoms(part, rng, nrand=400, per_loop=60)
        samples.append(G.Sample("polylines", sx, sy, sz, part, PL.BOX, 1))
    xs, ys, zs = (torch.tensor(np.concatenate([getattr(s, a) for s in samples]), device="cuda")
                  for a in "xyz")
    bt = B.Batch(hip_ctx, samples, 1, prec)
    u = [torch.zeros_like(xs) for _ in range(3)]
    bt.run_torch(xs, ys, zs, *u, accumulate=False)
    torch.cuda.synchronize()
    out = np.stack([t.cpu().numpy() for t in u])
    for k, s in enumerate(samples):
        o = bt.offsets[k]
        part = out[:, o:o + s.n]
        one = np.stack(B.line_sum(hip_ctx, s.x, s.y, s.z, s.loops, s.box, 1, prec))
        assert np.abs(part - one).max() < 1e-12
        orc = np.stack(oracle.line_sum(s.x, s.y, s.z, s.loops, s.box, 1))
        assert np.abs(part - orc).max() < TOL[prec]
    bt.free()
