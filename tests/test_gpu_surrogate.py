"""The learned far shell: MFMA forward and backward kernels against the numpy
FP64 reference (same bf16 rounding points), training, the feature and apply
paths, and the validation / exact-fallback pipeline."""
import numpy as np
import pytest

import mlp_ref as R
from disco_amd import backend as B
from disco_amd import geometry as G
from disco_amd import surrogate as S

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def sur(hip_ctx):
    s = S.Surrogate(hip_ctx, seed=5)
    yield s
    s.free()


def _rows(n, seed=0):
    rng = np.random.default_rng(seed)
    X = rng.normal(size=(n, 14)).astype(np.float32)
    X[:, 10:14] = np.eye(4)[rng.integers(0, 4, n)]
    T = (rng.normal(size=(n, 3)) * 0.3).astype(np.float32)
    return X, T


def test_forward_matches_reference(sur):
    p = R.init_params(3)
    sur.set_params(p)
    assert np.array_equal(sur.params(), p)
    X, _ = _rows(333)                          # not a multiple of 32
    got = sur.forward(X)
    want, _ = R.forward(p, X.astype(np.float64))
    err = np.abs(got - want).max() / np.abs(want).max()
    print("forward rel err", err)
    assert err < 5e-3


def test_gradient_matches_reference(sur):
    p = R.init_params(4)
    sur.set_params(p)
    X, T = _rows(257, seed=1)
    loss, g = sur.loss_grad(X, T)
    rl, rg = R.loss_grad(p, X.astype(np.float64), T.astype(np.float64))
    assert abs(loss - rl) <= 1e-4 * rl
    off = 0
    for name, shp in R.SIZES:
        n = int(np.prod(shp))
        a, b = g[off:off + n], rg[off:off + n]
        rel = np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30)
        print(name, "grad rel err", rel)
        assert rel < 2e-2, name
        off += n


def test_training_reduces_loss(hip_ctx):
    s = S.Surrogate(hip_ctx, seed=9)
    rng = np.random.default_rng(2)
    X = rng.uniform(-1, 1, size=(20000, 14)).astype(np.float32)
    T = np.stack([np.sin(2 * X[:, 0]) * X[:, 3], X[:, 1] * X[:, 2], np.cos(X[:, 9])], 1).astype(np.float32)
    mse = lambda: float(((s.forward(X[:4096]) - T[:4096]) ** 2).mean())
    l0 = mse()
    s.train(X, T, epochs=15, batch=1024, lr=3e-3)          # targets normalised, gain set
    l1 = mse()
    print("mse", l0, "->", l1, "gain", s.gain)
    assert abs(s.gain - np.sqrt((T.astype(np.float64) ** 2).mean())) < 1e-6 * s.gain
    assert l1 < 0.05 * l0
    s.free()


def test_features_and_apply(hip_ctx, sur):
    import torch
    p = R.init_params(6)
    sur.set_params(p)
    c2 = G.config_c2(nloops=3)
    infos = [S.shape_info(lp, g[1], "circle" if g[0] == 0 else "polygon")
             for lp, g in zip(c2.loops, c2.meta["geo"])]
    idx = np.arange(0, c2.n, 99991)
    P = np.stack([c2.x[idx], c2.y[idx], c2.z[idx]], 1)
    want = np.zeros((idx.size, 3))
    for info in infos:
        F = S.features(P, info, c2.box)
        assert F.shape == (idx.size, 14)
        assert np.allclose(F[:, 0], (P[:, 0] - info.c[0]) / c2.box[0], atol=1e-6)
        want += S.scale(info, c2.box) * sur.forward(F)
    x, y, z = (torch.from_numpy(np.ascontiguousarray(P[:, k])).cuda() for k in range(3))
    u = [torch.zeros_like(x) for _ in range(3)]
    sur.apply_torch(x, y, z, *u, infos, c2.box, chunk_rows=64)   # several chunks
    got = np.stack([t.cpu().numpy() for t in u], 1)
    assert np.abs(got - want).max() <= 1e-6 * np.abs(want).max()


def test_dipole_form_is_the_far_field_of_the_loop(hip_ctx):
    """One loopon of the loop's area reproduces its far shell to O((R/r)^2)."""
    import torch
    c2 = G.config_c2(nloops=2)
    lp = c2.loops[0]
    rng = np.random.default_rng(1)
    P = rng.uniform(-c2.box[0] / 2, c2.box[0] / 2, size=(64, 3))
    x, y, z = (torch.from_numpy(np.ascontiguousarray(P[:, k])).cuda() for k in range(3))
    ex = S.exact_far_shell_torch(hip_ctx, x, y, z, [lp], c2.box, 2, B.PREC_FP64)
    dp = S.dipole_far_shell_torch(hip_ctx, x, y, z, [lp], c2.box, 2)
    e = max(float((a - b).abs().max()) for a, b in zip(ex, dp))
    m = max(float(a.abs().max()) for a in ex)
    R_ = c2.meta["geo"][0][1]
    print("dipole rel err", e / m, "R", R_)
    assert e / m < 3 * (R_ / c2.box[0]) ** 2 + 1e-3


def test_rule7_far_shell_is_within_the_bar(hip_ctx):
    """The moment-matched 7-loopon rule reproduces each loop's far shell."""
    import torch
    c2 = G.config_c2(nloops=6)
    rng = np.random.default_rng(3)
    P = rng.uniform(-c2.box[0] / 2, c2.box[0] / 2, size=(256, 3))
    x, y, z = (torch.from_numpy(np.ascontiguousarray(P[:, k])).cuda() for k in range(3))
    ex = S.exact_far_shell_torch(hip_ctx, x, y, z, c2.loops, c2.box, 3, B.PREC_FP64)
    r7 = S.base_far_shell_torch(hip_ctx, x, y, z, c2.loops, c2.box, 3, "rule7", prec=B.PREC_FP64)
    dp = S.base_far_shell_torch(hip_ctx, x, y, z, c2.loops, c2.box, 3, "dipole", prec=B.PREC_FP64)
    e7 = max(float((a - b).abs().max()) for a, b in zip(ex, r7))
    ed = max(float((a - b).abs().max()) for a, b in zip(ex, dp))
    print("far shell err: rule7", e7, "dipole", ed)
    assert e7 < 1e-6 and e7 < ed / 100


def test_trained_network_gains_over_its_baseline(hip_ctx):
    """A network trained on exact far shells lowers the error of the dipole
    baseline on held-out (atom, loop) pairs."""
    F, T, Sc, ex, base = S.make_training_set(hip_ctx, nloops=240, atoms_per_loop=250, img=3,
                                             seed=4, baseline="dipole")
    n = F.shape[0]
    ntr = int(0.85 * n)
    s = S.Surrogate(hip_ctx, seed=2)
    s.train(F[:ntr], T[:ntr], epochs=40, batch=2048, lr=2e-3)
    pred = s.forward(F[ntr:])
    err_net = np.sqrt((((pred - T[ntr:]) * Sc[ntr:, None]) ** 2).mean())
    err_base = np.sqrt(((T[ntr:] * Sc[ntr:, None]) ** 2).mean())
    print("held-out rms: baseline only", err_base, "with network", err_net)
    assert err_net < 0.7 * err_base
    s.free()


def test_far_shell_falls_back_to_exact_when_untrained(hip_ctx, sur):
    import torch
    sur.set_params(R.init_params(8))                 # an untrained network
    c2 = G.config_c2(nloops=2)
    infos = [S.shape_info(lp, g[1], "circle" if g[0] == 0 else "polygon")
             for lp, g in zip(c2.loops, c2.meta["geo"])]
    idx = np.arange(0, c2.n, 7919)
    x, y, z = (torch.from_numpy(a[idx].copy()).cuda() for a in (c2.x, c2.y, c2.z))
    u = [torch.zeros_like(x) for _ in range(3)]
    rep = S.far_shell(hip_ctx, sur, x, y, z, u, c2.loops, infos, c2.box, 2, tol=1e-9,
                      n_valid=64)
    assert not rep.used_surrogate and rep.max_err > 1e-9
    ex = S.exact_far_shell_torch(hip_ctx, x, y, z, c2.loops, c2.box, 2)
    assert max(float((a - b).abs().max()) for a, b in zip(u, ex)) == 0.0


def test_grid_interp_reproduces_cubics(hip_ctx):
    """Tensor-product 4-node cubic Lagrange is exact for degree <= 3 per axis;
    accumulate adds, overwrite writes."""
    import torch
    rng = np.random.default_rng(7)
    h, origin, dims = 1.7, (-3.0, 2.0, 0.5), (9, 7, 11)
    f = lambda X, Y, Z: ((1 + 2 * X - 0.3 * X ** 2 + 0.05 * X ** 3) * (3 - Y + 0.2 * Y * Y)
                         * (0.01 * Z ** 3 - 2 * Z + 1))
    ax = [origin[k] + h * np.arange(dims[k]) for k in range(3)]
    Zg, Yg, Xg = np.meshgrid(ax[2], ax[1], ax[0], indexing="ij")
    g = [torch.from_numpy(np.ascontiguousarray(c * f(Xg, Yg, Zg))).cuda() for c in (1.0, -2.0, 0.5)]
    P = np.stack([rng.uniform(origin[k] + h, origin[k] + h * (dims[k] - 2), 500) for k in range(3)], 1)
    x, y, z = (torch.from_numpy(np.ascontiguousarray(P[:, k])).cuda() for k in range(3))
    u = [torch.full_like(x, 5.0) for _ in range(3)]
    B.grid_interp_torch(hip_ctx, x, y, z, *u, g, origin, h, accumulate=False)
    want = f(P[:, 0], P[:, 1], P[:, 2])
    for c, t in zip((1.0, -2.0, 0.5), u):
        assert np.abs(t.cpu().numpy() - c * want).max() <= 1e-10 * np.abs(want).max()
    B.grid_interp_torch(hip_ctx, x, y, z, *u, g, origin, h, accumulate=True)
    assert np.abs(u[0].cpu().numpy() - 2 * want).max() <= 1e-10 * np.abs(want).max()


def test_grid_interp_quintic_reproduces_degree_5(hip_ctx):
    """The 6-node stencil is exact for degree <= 5 per axis; its grid needs 6 nodes
    per axis, and orders other than 4 and 6 are refused."""
    import torch
    rng = np.random.default_rng(9)
    h, origin, dims = 1.3, (1.0, -2.0, 0.25), (11, 9, 13)
    f = lambda X, Y, Z: ((1 - X + 0.2 * X ** 3 - 0.01 * X ** 5) * (2 + Y ** 4 * 0.01 - Y)
                         * (0.5 + 0.001 * Z ** 5 - 0.02 * Z ** 2))
    ax = [origin[k] + h * np.arange(dims[k]) for k in range(3)]
    Zg, Yg, Xg = np.meshgrid(ax[2], ax[1], ax[0], indexing="ij")
    g = [torch.from_numpy(np.ascontiguousarray(c * f(Xg, Yg, Zg))).cuda() for c in (1.0, 3.0, -1.0)]
    P = np.stack([rng.uniform(origin[k] + 2 * h, origin[k] + h * (dims[k] - 3), 700)
                  for k in range(3)], 1)
    x, y, z = (torch.from_numpy(np.ascontiguousarray(P[:, k])).cuda() for k in range(3))
    u = [torch.zeros_like(x) for _ in range(3)]
    B.grid_interp_torch(hip_ctx, x, y, z, *u, g, origin, h, accumulate=False, order=6)
    want = f(P[:, 0], P[:, 1], P[:, 2])
    for c, t in zip((1.0, 3.0, -1.0), u):
        assert np.abs(t.cpu().numpy() - c * want).max() <= 1e-10 * np.abs(want).max()
    # the cubic stencil is not exact for this field
    v = [torch.zeros_like(x) for _ in range(3)]
    B.grid_interp_torch(hip_ctx, x, y, z, *v, g, origin, h, accumulate=False, order=4)
    assert np.abs(v[0].cpu().numpy() - want).max() > 1e-6 * np.abs(want).max()
    small = [t[:5, :5, :5].contiguous() for t in g]
    with pytest.raises(B.DiscoError):
        B.grid_interp_torch(hip_ctx, x, y, z, *u, small, origin, h, order=6)
    with pytest.raises(B.DiscoError):
        B.grid_interp_torch(hip_ctx, x, y, z, *u, g, origin, h, order=5)


@pytest.mark.parametrize("div,tol64", [(32, 2e-7), (24, 6e-7)])
def test_grid_far_shell_matches_the_direct_sum(hip_ctx, div, tol64):
    """The rule's far shell (images 2..4) tabulated on an L/32 (L/24) grid over
    the atoms and interpolated equals the direct loopon sum at every atom (to
    2e-7 (6e-7) A: the interpolation error falls as h^4)."""
    import torch
    c2 = G.config_c2(nloops=6)
    rng = np.random.default_rng(4)
    idx = np.sort(rng.choice(c2.n, 3000, replace=False))
    x, y, z = (torch.from_numpy(np.ascontiguousarray(a[idx])).cuda() for a in (c2.x, c2.y, c2.z))
    # the interpolation error of an L/32 grid: ~6e-6 of the far shell's ~1e-2 A
    for prec, tol in ((B.PREC_FP64, tol64), (B.PREC_FP32C, tol64 + 8e-7)):
        direct = S.base_far_shell_torch(hip_ctx, x, y, z, c2.loops, c2.box, 4, "rule7",
                                        prec=prec)
        grid = S.base_far_shell_torch(hip_ctx, x, y, z, c2.loops, c2.box, 4, "rule7",
                                      prec=prec, grid_h=c2.box[0] / div)
        d = max(float((a - b).abs().max()) for a, b in zip(direct, grid))
        m = max(float(a.abs().max()) for a in direct)
        print(f"grid vs direct far shell, prec {prec}: {d:.3e} A (field up to {m:.3e} A)")
        assert d < tol
