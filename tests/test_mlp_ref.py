"""Pins the CPU reference of the proposal network before it judges the GPU kernels: its
analytic gradient against central finite differences (float64), its loss against a direct
per-sample formula, and the BF16 rounding helper against exact cases."""
import numpy as np

import mlp_ref as R


def test_bf16_rounding():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 2 ** -7, 1.0 + 3 * 2 ** -8, -2.5, 0.0, 3.0e38], np.float32)
    got = R.bf16(x)
    # 1 + 2^-8 is a tie between 1 and 1 + 2^-7: even mantissa (1.0) wins; 1 + 3*2^-8 ties to 1 + 2^-6
    assert got.tolist() == [1.0, 1.0, 1.0 + 2 ** -7, 1.0 + 2 ** -6, -2.5, 0.0, R.bf16(np.float32(3.0e38))]
    y = np.random.default_rng(0).normal(size=1000).astype(np.float32)
    assert np.all(np.abs(R.bf16(y) - y) <= np.abs(y) * 2 ** -8)


def test_loss_matches_direct_formula():
    rng = np.random.default_rng(1)
    p = R.init_params(1)
    x, y = R.random_samples(rng, 50)
    loss, _, z = R.forward_backward(p, x, y, bf16_ops=False)
    want = 0.0
    for i in range(50):
        for j in range(8):
            if (y[i] >> (8 + j)) & 1:
                t = (y[i] >> j) & 1
                q = 1.0 / (1.0 + np.exp(-z[i, j]))
                want -= t * np.log(q) + (1 - t) * np.log(1 - q)
        c = z[i, 8:]
        b = (y[i] >> 16) & 7
        want -= c[b] - np.log(np.exp(c).sum())
    assert abs(loss - want) <= 1e-9 * abs(want)


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(2)
    p = R.init_params(2).astype(np.float64)
    x, y = R.random_samples(rng, 40)
    _, g, _ = R.forward_backward(p, x, y, bf16_ops=False)
    # every bias, and random weights of each matrix (padded rows of W1 excluded)
    idx = list(range(R.OFF_B1, R.OFF_W2)) + list(range(R.OFF_B2, R.OFF_W3)) + list(range(R.OFF_B3, R.NPARAM))
    idx += [int(i) for i in rng.choice(R.F * R.H, 60, replace=False)]
    idx += [R.OFF_W2 + int(i) for i in rng.choice(R.H * R.H, 60, replace=False)]
    idx += [R.OFF_W3 + int(i) for i in rng.choice(R.H * R.OUT, 60, replace=False)]
    eps = 1e-6
    for i in idx:
        q = p.copy()
        q[i] += eps
        lp, _, _ = R.forward_backward(q, x, y, bf16_ops=False)
        q[i] -= 2 * eps
        lm, _, _ = R.forward_backward(q, x, y, bf16_ops=False)
        fd = (lp - lm) / (2 * eps)
        assert abs(fd - g[i]) <= 1e-5 * max(1.0, abs(fd)), (i, fd, g[i])
    # padded input rows carry no gradient
    assert np.all(g[R.F * R.H:R.IN * R.H] == 0)


def test_bf16_reference_stays_close_to_float64():
    rng = np.random.default_rng(3)
    p = R.init_params(3)
    x, y = R.random_samples(rng, 500)
    l32, g32, z32 = R.forward_backward(p, x, y, bf16_ops=True)
    l64, g64, z64 = R.forward_backward(p, x, y, bf16_ops=False)
    assert abs(l32 - l64) <= 0.02 * abs(l64)
    assert np.linalg.norm(g32 - g64) <= 0.05 * np.linalg.norm(g64)
