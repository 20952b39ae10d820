"""The closure network's HIP kernels (rp_mlp_*) against the oracle
(oracle/mlp.py, FP64 with the same BF16-rounded weights): forward outputs and
the gradient of the mean-squared error for every layer, within the FP32
accumulation of the GPU (1e-4 relative per tensor); and a training run that
lowers the loss of a smooth regression problem tenfold.
"""
import numpy as np
import pytest

from oracle import mlp as omlp
from respackdev_amd.closure import Mlp

pytestmark = pytest.mark.gpu


def test_forward_and_gradient_match_oracle():
    net = Mlp(0, nin=8, width=128, depth=3, nout=2, seed=11)
    try:
        rng = np.random.default_rng(4)
        p = (rng.standard_normal(net.nparams) * 0.08).astype(np.float32)
        net.params = p
        x = rng.standard_normal((1000, 8)).astype(np.float32)
        t = rng.standard_normal((1000, 2)).astype(np.float32)
        y = net.forward(x)
        yo = omlp.forward(p, net.dims, x)
        assert np.linalg.norm(y - yo) <= 1e-5 * np.linalg.norm(yo)
        loss, g = net.gradient(x, t)
        lo, go = omlp.gradient(p, net.dims, x, t)
        assert abs(loss - lo) <= 1e-5 * lo
        lay, _ = omlp.layout(net.dims)
        for l, (wo, bo) in enumerate(lay):
            for a, b, what in ((wo, bo, "W"), (bo, bo + net.dims[l + 1], "b")):
                err = np.linalg.norm(g[a:b] - go[a:b]) / np.linalg.norm(go[a:b])
                print(f"layer {l} {what}: rel {err:.2e}", flush=True)
                assert err <= 1e-4
    finally:
        net.close()


def test_training_lowers_loss():
    net = Mlp(0, nin=8, width=128, depth=3, nout=2, seed=5)
    try:
        rng = np.random.default_rng(9)
        x = rng.uniform(-1, 1, (20000, 8)).astype(np.float32)
        t = np.stack([np.sin(2 * x[:, 0]) * x[:, 1], x[:, 2] ** 2 - x[:, 3]], 1).astype(np.float32)
        l0, _ = net.gradient(x[:4096], t[:4096])
        losses = net.train(x, t, steps=600, batch=1024, lr=2e-3)
        l1, _ = net.gradient(x[:4096], t[:4096])
        print(f"loss {l0:.4f} -> {l1:.5f} (last minibatch {losses[-1]:.5f})", flush=True)
        assert np.all(np.isfinite(losses))
        assert l1 < 0.1 * l0
    finally:
        net.close()
