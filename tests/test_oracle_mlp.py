"""Pins the oracle of the band-tail closure network (oracle/mlp.py): its
analytic gradient against central finite differences of its own loss (FP64,
unrounded weights, where the loss is smooth), the exact GELU against erf
identities, and the BF16 rounding against the round-to-nearest-even rule.
"""
import numpy as np

from oracle import mlp


def _net(seed=3, dims=(6, 16, 16, 2)):
    _, n = mlp.layout(dims)
    rng = np.random.default_rng(seed)
    p = rng.standard_normal(n) * 0.3
    x = rng.standard_normal((40, dims[0]))
    t = rng.standard_normal((40, dims[-1]))
    return list(dims), p, x, t


def test_gradient_matches_finite_differences():
    dims, p, x, t = _net()
    loss, g = mlp.gradient(p, dims, x, t, round_bf16=False)
    rng = np.random.default_rng(1)
    idx = rng.choice(len(p), 60, replace=False)
    h = 1e-6
    for i in idx:
        pp = p.copy(); pp[i] += h
        pm = p.copy(); pm[i] -= h
        lp, _ = mlp.gradient(pp, dims, x, t, round_bf16=False)
        lm, _ = mlp.gradient(pm, dims, x, t, round_bf16=False)
        fd = (lp - lm) / (2 * h)
        assert abs(fd - g[i]) <= 1e-7 * max(1.0, abs(g[i])), (i, fd, g[i])
    # the loss is the mean squared error of forward()
    y = mlp.forward(p, dims, x, round_bf16=False)
    assert abs(loss - np.mean((y - t) ** 2)) <= 1e-14


def test_gelu_and_bf16_rounding():
    z = np.linspace(-6, 6, 101)
    g = mlp._gelu(z)
    # gelu(z) - gelu(-z) = z and gelu(0) = 0
    np.testing.assert_allclose(g - mlp._gelu(-z), z, atol=1e-13)
    assert mlp._gelu(np.array([0.0]))[0] == 0.0
    # bf16: 1 + 2^-8 is a tie between 1 and 1 + 2^-7 and rounds to even (1); 1 + 3*2^-8 rounds up
    assert mlp.bf16(np.float32(1.0 + 2.0 ** -8)) == 1.0
    assert mlp.bf16(np.float32(1.0 + 3 * 2.0 ** -8)) == 1.0 + 2.0 ** -6
    v = np.random.default_rng(0).standard_normal(1000).astype(np.float32)
    r = mlp.bf16(v)
    assert np.all(np.abs(r - v) <= np.abs(v) * 2.0 ** -8)
