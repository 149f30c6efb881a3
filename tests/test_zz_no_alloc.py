"""disco_hip_plan_run_noalloc through the C ABI: the same results as
disco_hip_plan_run, with a split loopon plan's partial sums in the library's
static device scratch instead of a buffer the plan allocates."""
import ctypes

import numpy as np
import pytest

from disco_amd import backend as B
from disco_amd import geometry as G
from disco_amd import surrogate as S

pytestmark = pytest.mark.gpu

SCRATCH_BYTES = 64 << 20        # DISCO_NOALLOC_SCRATCH_BYTES


def _noalloc(plan, x, y, z, u, accumulate=False):
    import torch
    stream = torch.cuda.current_stream().cuda_stream
    return B.lib().disco_hip_plan_run_noalloc(plan._h, x.numel(), x.data_ptr(), y.data_ptr(),
                                              z.data_ptr(), u[0].data_ptr(), u[1].data_ptr(),
                                              u[2].data_ptr(), 1 if accumulate else 0, stream)


def _grid_plan(ctx, groups):
    """A far-shell loopon plan of 30 loops' 7-point rules (images +-3 without the
    +-1 shell), split into `groups` loop groups, as the grid far shell uses."""
    s = G.config_c2(seed=5, nloops=30)
    plan = B.Plan(ctx, S.base_loopons(s.loops, "rule7"), s.box, 3, B.PREC_FP32C, kind="loopon")
    S.set_inner(plan, 1)
    S.set_split(plan, groups)
    return plan, s.box


def _points(n, box, seed):
    import torch
    rng = np.random.default_rng(seed)
    p = rng.uniform(-box[0] / 2, box[0] / 2, size=(3, n))
    return [torch.from_numpy(np.ascontiguousarray(r)).cuda() for r in p]


def test_noalloc_equals_plan_run_on_two_calls(hip_ctx):
    import torch
    plan, box = _grid_plan(hip_ctx, 8)
    for n, seed in ((6000, 1), (1234, 2)):          # the second call is smaller
        x, y, z = _points(n, box, seed)
        ref = [torch.zeros_like(x) for _ in range(3)]
        plan.run_torch(x, y, z, *ref)
        got = [torch.full_like(x, 3.0) for _ in range(3)]
        hip_ctx.check(_noalloc(plan, x, y, z, got))
        torch.cuda.synchronize()
        assert all(torch.equal(a, b) for a, b in zip(got, ref)), n
        # accumulate = 1 adds the sum to u (one FP64 add per component)
        acc = [torch.full_like(x, 0.25) for _ in range(3)]
        hip_ctx.check(_noalloc(plan, x, y, z, acc, accumulate=True))
        torch.cuda.synchronize()
        assert all(torch.equal(a, b + 0.25) for a, b in zip(acc, ref)), n
    plan.free()


def test_noalloc_line_plan_and_limits(hip_ctx):
    import torch
    s = G.config_c1()
    idx = np.arange(0, s.n, 31)
    x, y, z = (torch.from_numpy(np.ascontiguousarray(a[idx])).cuda() for a in (s.x, s.y, s.z))
    plan = B.Plan(hip_ctx, s.loops, s.box, 1, B.PREC_FP32C)
    ref = [torch.zeros_like(x) for _ in range(3)]
    plan.run_torch(x, y, z, *ref)
    got = [torch.zeros_like(x) for _ in range(3)]
    hip_ctx.check(_noalloc(plan, x, y, z, got))
    torch.cuda.synchronize()
    assert all(torch.equal(a, b) for a, b in zip(got, ref))
    plan.free()
    # a split plan whose partial sums exceed the static scratch is refused
    # before any launch: u is untouched
    groups = 8
    gp, box = _grid_plan(hip_ctx, groups)
    n = SCRATCH_BYTES // (24 * groups) + 1
    x, y, z = _points(n, box, 3)
    u = [torch.full_like(x, 5.0) for _ in range(3)]
    rc = _noalloc(gp, x, y, z, u)
    torch.cuda.synchronize()
    assert rc == 5                                  # DISCO_ERR_TOO_LARGE
    assert all(bool((t == 5.0).all()) for t in u)
    n_fit = SCRATCH_BYTES // (24 * groups)           # the largest size that fits
    ref = [torch.zeros_like(x[:n_fit]) for _ in range(3)]
    gp.run_torch(x[:n_fit], y[:n_fit], z[:n_fit], *ref)
    got = [torch.zeros_like(x[:n_fit]) for _ in range(3)]
    hip_ctx.check(_noalloc(gp, x[:n_fit], y[:n_fit], z[:n_fit], got))
    torch.cuda.synchronize()
    assert all(torch.equal(a, b) for a, b in zip(got, ref))
    assert B.lib().disco_hip_plan_run_noalloc(None, 0, None, None, None, None, None, None, 0,
                                              None) == 1
    gp.free()
