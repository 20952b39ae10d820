"""Band-tail closure for chi (no counterpart in the reference).

chi_GG'(q, w) is a sum over occupied x empty band pairs; truncating the empty
bands at n_trunc leaves out a tail that decays with the band energy.  A small
MLP (3 hidden layers of 128, GELU; BF16 weights, FP32 accumulation; trained and
evaluated by the library's own HIP kernels through rp_mlp_*) predicts that tail
for the diagonal and near-diagonal entries (G' = G or G' = G +- b_i) from the
truncated chi at the same entry, |q+G|, |q+G'| and w.  The full-band method
stays the fallback and the judge: the closure is trained on chi data the method
itself produces at n_trunc and at n_full empty bands.
"""
from __future__ import annotations

import numpy as np

from . import _lib
from ._lib import P, check, ptr

NIN, NOUT = 8, 2
WIDTH, DEPTH = 128, 3


class Mlp:
    """The closure network on one GPU (its own handle and HIP stream)."""

    def __init__(self, device: int = 0, nin: int = NIN, width: int = WIDTH, depth: int = DEPTH,
                 nout: int = NOUT, seed: int = 1):
        self._h = P()
        check(_lib.lib().rp_init(device, _lib.C.byref(self._h)))
        check(_lib.lib().rp_mlp_open(self._h, nin, width, depth, nout, seed))
        self.dims = [nin] + [width] * depth + [nout]
        n = np.zeros(1, np.int64)
        check(_lib.lib().rp_mlp_nparams(self._h, ptr(n)))
        self.nparams = int(n[0])

    def close(self):
        if self._h:
            _lib.lib().rp_finalize(self._h)
            self._h = None

    @property
    def params(self) -> np.ndarray:
        p = np.zeros(self.nparams, np.float32)
        check(_lib.lib().rp_mlp_get_params(self._h, ptr(p)))
        return p

    @params.setter
    def params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        assert p.size == self.nparams
        check(_lib.lib().rp_mlp_set_params(self._h, ptr(p)))

    def forward(self, x) -> np.ndarray:
        x = np.ascontiguousarray(x, np.float32)
        y = np.zeros((len(x), self.dims[-1]), np.float32)
        check(_lib.lib().rp_mlp_forward(self._h, len(x), ptr(x), ptr(y)))
        return y

    def gradient(self, x, t):
        x = np.ascontiguousarray(x, np.float32)
        t = np.ascontiguousarray(t, np.float32)
        g = np.zeros(self.nparams, np.float32)
        loss = np.zeros(1)
        check(_lib.lib().rp_mlp_gradient(self._h, len(x), ptr(x), ptr(t), ptr(g), ptr(loss)))
        return float(loss[0]), g

    def train(self, x, t, steps: int, batch: int = 4096, lr: float = 1e-3, seed: int = 7) -> np.ndarray:
        x = np.ascontiguousarray(x, np.float32)
        t = np.ascontiguousarray(t, np.float32)
        losses = np.zeros(steps)
        check(_lib.lib().rp_mlp_train(self._h, len(x), ptr(x), ptr(t), steps, batch, lr, seed,
                                       ptr(losses)))
        return losses
