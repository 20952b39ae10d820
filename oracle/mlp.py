"""Test infrastructure: the oracle of the band-tail closure network (no
counterpart in the reference).  A numpy FP64 restatement of the network the
library trains (respackdev_amd/csrc/rp_mlp.*): fully connected layers with the
exact GELU (x/2 (1 + erf(x/sqrt 2))) on the hidden layers, a linear output,
weights rounded to BF16 (round to nearest even) as the GPU uses them, and the
mean-squared error over all outputs.  Parameters are one vector: per layer the
weights W[out][in], then the biases.

Pinned by central finite differences of the loss (tests/test_oracle_mlp.py);
only tests/ use it, as the checker of the GPU gradients.
"""
import numpy as np
from scipy.special import erf


def layout(dims):
    """(W offset, b offset) per layer and the parameter count."""
    off, out = 0, []
    for l in range(len(dims) - 1):
        wo = off
        off += dims[l + 1] * dims[l]
        out.append((wo, off))
        off += dims[l + 1]
    return out, off


def bf16(x):
    """FP32 -> BF16 -> FP32, round to nearest even."""
    u = np.asarray(x, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def _gelu(z):
    return 0.5 * z * (1.0 + erf(z / np.sqrt(2.0)))


def _gelu_grad(z):
    return 0.5 * (1.0 + erf(z / np.sqrt(2.0))) + z * np.exp(-0.5 * z * z) / np.sqrt(2.0 * np.pi)


def _weights(p, dims, round_bf16):
    lay, _ = layout(dims)
    ws = []
    for l, (wo, bo) in enumerate(lay):
        w = np.asarray(p[wo:bo], np.float64).reshape(dims[l + 1], dims[l])
        b = np.asarray(p[bo:bo + dims[l + 1]], np.float64)
        ws.append((bf16(w) if round_bf16 else w, b))
    return ws


def forward(p, dims, x, round_bf16=True):
    ws = _weights(p, dims, round_bf16)
    a = np.asarray(x, np.float64)
    L = len(ws)
    for l, (w, b) in enumerate(ws):
        z = a @ w.T + b
        a = _gelu(z) if l < L - 1 else z
    return a


def gradient(p, dims, x, t, round_bf16=True):
    """(loss, grad) of mean((y - t)^2) over all samples and outputs."""
    ws = _weights(p, dims, round_bf16)
    lay, n = layout(dims)
    L = len(ws)
    acts, zs = [np.asarray(x, np.float64)], []
    for l, (w, b) in enumerate(ws):
        z = acts[-1] @ w.T + b
        zs.append(z)
        acts.append(_gelu(z) if l < L - 1 else z)
    y = acts[-1]
    m = y.size
    d = y - np.asarray(t, np.float64)
    loss = float(np.sum(d * d) / m)
    dz = 2.0 * d / m
    g = np.zeros(n)
    for l in range(L - 1, -1, -1):
        wo, bo = lay[l]
        g[wo:bo] = (dz.T @ acts[l]).reshape(-1)
        g[bo:bo + dims[l + 1]] = dz.sum(axis=0)
        if l > 0:
            dz = (dz @ ws[l][0]) * _gelu_grad(zs[l - 1])
    return loss, g
