"""The learned far shell on the 200-loop, images +-4 sample: the rule's far
shell (7-point rule on an L/32 grid) plus the trained network, validated and
applied by surrogate.far_shell, with the exact +-1 shell, against the FP64
plain sum over all 729 images; and the exact fallback when the baseline is too
coarse for the network to bring within the tolerance."""
import numpy as np
import pytest

from disco_amd import backend as B
from disco_amd import geometry as G
from disco_amd import surrogate as S
from test_gpu_cubature import _dev, _err, _near_cut_atoms, _plain64

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def c5_case(hip_ctx):
    s = G.config_c2(seed=20245, nloops=200)
    rng = np.random.default_rng(8)
    L = s.box[0]
    P = np.concatenate([rng.uniform(-L / 2, L / 2, (4096, 3)),
                        _near_cut_atoms(s.loops, rng, 20, 6.0)])
    P = np.clip(P, -L / 2, L / 2 - 1e-9)
    x, y, z = _dev(P[:, 0], P[:, 1], P[:, 2])
    infos = [S.shape_info(lp, g[1], "circle" if g[0] == 0 else "polygon")
             for lp, g in zip(s.loops, s.meta["geo"])]
    ref = _plain64(hip_ctx, x, y, z, s.loops, s.box, 4)
    return s, x, y, z, infos, ref


def _near_shell(ctx, s, x, y, z):
    import torch
    u = [torch.zeros_like(x) for _ in range(3)]
    p = B.Plan(ctx, s.loops, s.box, 1, B.PREC_FP32C)
    p.run_torch(x, y, z, *u)
    torch.cuda.synchronize()
    p.free()
    return u


def test_accepted_surrogate_path(hip_ctx, c5_case):
    """Trained on 1e6 exact far shells (2,000 loops x 500 atoms): the surrogate
    passes its 1e-5 A validation and is applied; the sample's sum is within
    1e-5 A of the FP64 plain sum at every test atom, and the network lowers the
    held-out rms error of the rule alone."""
    s, x, y, z, infos, ref = c5_case
    F, T, Sc, _, _ = S.make_training_set(hip_ctx, nloops=2000, atoms_per_loop=500, img=4,
                                         seed=11, baseline="rule7")
    ntr = int(0.9 * F.shape[0])
    sur = S.Surrogate(hip_ctx, seed=1)
    sur.train(F[:ntr], T[:ntr], epochs=30, batch=4096, lr=2e-3)
    pred = sur.forward(F[ntr:])
    rms_net = float(np.sqrt((((pred - T[ntr:]) * Sc[ntr:, None]) ** 2).mean()))
    rms_base = float(np.sqrt(((T[ntr:] * Sc[ntr:, None]) ** 2).mean()))
    u = _near_shell(hip_ctx, s, x, y, z)
    rep = S.far_shell(hip_ctx, sur, x, y, z, u, s.loops, infos, s.box, 4, tol=1e-5,
                      baseline="rule7", grid_h=s.box[0] / 32)
    e = _err(u, ref)
    print(f"surrogate path: validation {rep.max_err:.3e} A (rule only {rep.base_only_err:.3e}); "
          f"vs FP64 plain at {x.numel()} atoms {e:.3e} A; held-out rms {rms_net:.3e} "
          f"(rule only {rms_base:.3e})")
    assert rep.used_surrogate
    assert e < 1e-5
    assert rms_net < rms_base
    sur.free()


def test_coarse_baseline_falls_back_to_exact(hip_ctx, c5_case):
    """The dipole baseline with a network trained for it stays above 1e-5 A on
    validation: far_shell sums the exact far shell instead, bit for bit the exact
    far-shell plan's output."""
    import torch
    s, x, y, z, infos, ref = c5_case
    F, T, Sc, _, _ = S.make_training_set(hip_ctx, nloops=300, atoms_per_loop=200, img=4,
                                         seed=3, baseline="dipole")
    sur = S.Surrogate(hip_ctx, seed=2)
    sur.train(F, T, epochs=10, batch=2048, lr=2e-3)
    u = [torch.zeros_like(x) for _ in range(3)]
    rep = S.far_shell(hip_ctx, sur, x, y, z, u, s.loops, infos, s.box, 4, tol=1e-5,
                      baseline="dipole")
    assert not rep.used_surrogate and rep.max_err > 1e-5
    ex = S.exact_far_shell_torch(hip_ctx, x, y, z, s.loops, s.box, 4)
    assert all(torch.equal(a, b) for a, b in zip(u, ex))
    near = _near_shell(hip_ctx, s, x, y, z)
    tot = [a + b for a, b in zip(near, u)]
    assert _err(tot, ref) < 1e-5
    sur.free()
