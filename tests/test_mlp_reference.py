"""The numpy MLP used to check the GPU kernels, pinned by finite differences
(FP64, no rounding) -- so the gradient check does not rest on a retyped formula."""
import numpy as np

import mlp_ref as R


def test_backward_matches_central_differences():
    rng = np.random.default_rng(0)
    p = R.init_params(1).astype(np.float64)
    X = rng.normal(size=(7, 14))
    T = rng.normal(size=(7, 3)) * 0.1
    _, g = R.loss_grad(p, X, T, rnd=False)
    idx = rng.choice(p.size, 60, replace=False)
    for name, shp in R.SIZES:           # make sure every block is sampled
        off = sum(int(np.prod(s)) for n, s in R.SIZES[:[n for n, _ in R.SIZES].index(name)])
        k = off + (int(np.prod(shp)) // 2 if name not in ("W4", "b4") else 1)
        idx = np.append(idx, k)
    h = 1e-6
    for k in idx:
        e = np.zeros_like(p)
        e[k] = h
        lp, _ = R.loss_grad(p + e, X, T, rnd=False)
        lm, _ = R.loss_grad(p - e, X, T, rnd=False)
        fd = (lp - lm) / (2 * h)
        assert abs(fd - g[k]) <= 1e-6 * max(1.0, abs(fd)), (k, fd, g[k])


def test_bf16_rounding_is_round_to_nearest_even():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -9, -2.5, 3.3], np.float32)
    got = R.bf16(x)
    assert got[0] == 1.0
    assert got[1] == 1.0                     # tie -> even
    assert got[2] == 1.0 + 2 ** -7           # above the tie -> up
    assert got[3] == -2.5
    assert abs(got[4] - 3.3) < 3.3 * 2 ** -8
