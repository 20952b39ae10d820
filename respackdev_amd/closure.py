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


AU_EV = 27.21151


def near_diagonal_pairs(lg: np.ndarray, q) -> np.ndarray:
    """(ig, jg) with G' = G or G' = G +- b_i, both in the chi sphere; at q = 0 the
    G = 0 row and column are left out (their heads come from the momentum
    matrix elements, chiqw.F:1780-1848, not from a band sum of this form)."""
    idx = {tuple(int(v) for v in g): i for i, g in enumerate(lg)}
    q0 = all(abs(v) < 1e-12 for v in q)
    out = []
    for i, g in enumerate(lg):
        if q0 and not g.any():
            continue
        out.append((i, i))
        for d in range(3):
            for s in (1, -1):
                h = [int(v) for v in g]
                h[d] += s
                j = idx.get(tuple(h))
                if j is not None and not (q0 and not any(h)):
                    out.append((i, j))
    return np.array(out, np.int64)


def features(chi, lg, q, b, em, volume, e_cut):
    """Inputs per (direction, frequency, entry): V chi_trunc (re, im), |q+G|,
    |q+G'|, w (re, im; Ha), [G == G'], truncation energy above E_F (Ha)."""
    pairs = near_diagonal_pairs(lg, q)
    qg = np.linalg.norm((np.asarray(q)[None, :] + lg) @ b, axis=1)
    ig, jg = pairs[:, 0], pairs[:, 1]
    vals = chi[:, :, jg, ig] * volume             # chi is [dir][ie][jg][ig]
    nd, nen, npr = vals.shape
    x = np.empty((nd, nen, npr, NIN), np.float64)
    x[..., 0] = vals.real
    x[..., 1] = vals.imag
    x[..., 2] = qg[ig]
    x[..., 3] = qg[jg]
    x[..., 4] = np.asarray(em).real[None, :, None]
    x[..., 5] = np.asarray(em).imag[None, :, None]
    x[..., 6] = (ig == jg)
    x[..., 7] = e_cut
    return x.reshape(-1, NIN), pairs


def targets(chi_full, chi_trunc, pairs, volume):
    ig, jg = pairs[:, 0], pairs[:, 1]
    d = (chi_full[:, :, jg, ig] - chi_trunc[:, :, jg, ig]) * volume
    return np.stack([d.real, d.imag], -1).reshape(-1, NOUT)


class Scaler:
    """Standardisation of inputs and targets by training-set statistics."""

    def __init__(self, x, t):
        self.mx, self.sx = x.mean(0), x.std(0)
        self.mt, self.st = t.mean(0), t.std(0)
        self.sx[self.sx == 0] = 1.0
        self.st[self.st == 0] = 1.0

    def x(self, x):
        return ((x - self.mx) / self.sx).astype(np.float32)

    def t(self, t):
        return ((t - self.mt) / self.st).astype(np.float32)

    def t_inv(self, tn):
        return tn.astype(np.float64) * self.st + self.mt


def family_system(i: int, nk: int, nocc: int, nvir: int, ecut_psi: float, seed0: int):
    """Member i of the training family: lattice 6.5-8.0 bohr and band gap 1-4 eV
    drawn from seed0 + i (the precision configuration's generator)."""
    from . import gen
    rng = np.random.default_rng(seed0 + i)
    a = float(rng.uniform(6.5, 8.0))
    gap = float(rng.uniform(1.0, 4.0))
    s = gen.make_system(a=a, nk=nk, nocc=nocc, nvir=nvir, ecut_psi=ecut_psi, seed=seed0 + 1000 + i,
                        gap_ev=gap)
    return s, a, gap


def truncation_energy(s, nvir: int) -> float:
    """Highest kept empty-band energy above E_F (Ha)."""
    return float(np.max(s.eig[:, s.nocc + nvir - 1]) - s.ef)


def training_samples(s, ecut_eps: float, nen: int, n_trunc: int, n_full: int, device: int = 0,
                     n_q: int | None = None, seed: int = 0):
    """(inputs, targets) of one system: every irreducible q (or n_q of them drawn
    with `seed`), every direction and frequency, the near-diagonal entries of chi
    at n_trunc against n_full empty bands (FP64 chiqw on the GPU)."""
    from .chiqw import Chiqw, from_system
    from . import gen
    w = from_system(s)
    e_cut = truncation_energy(s, n_trunc)
    ks = list(w.ski)
    if n_q is not None and n_q < len(ks):
        pick = np.random.default_rng(seed).choice(len(ks), n_q, replace=False)
        ks = [ks[i] for i in sorted(pick)]
    xs, ts = [], []
    c = Chiqw(device)
    try:
        info = c.open(w, ecut_for_eps=ecut_eps, num_freq_grid=nen)
        em, vol = np.asarray(info["em"]), float(info["volume"])
        for k in ks:
            q = tuple(gen._fold(v) for v in k)
            c.set_window(0.0, n_full)
            full = c.run_q(q, "fp64", want_chi=True, want_eps=False)
            c.set_window(0.0, n_trunc)
            tr = c.run_q(q, "fp64", want_chi=True, want_eps=False)
            x, pairs = features(tr["chi"], tr["lg"], q, s.bvec, em, vol, e_cut)
            xs.append(x)
            ts.append(targets(full["chi"], tr["chi"], pairs, vol))
    finally:
        c.close()
    return np.concatenate(xs), np.concatenate(ts)


def eps_variants(s, ecut_eps: float, nen: int, n_trunc: int, n_full: int, plain, net=None,
                 scaler=None, device: int = 0):
    """eps^-1 per irreducible q for: the full band set, plain truncation at each
    count in `plain`, and (with a trained net) n_trunc plus the closure.
    Returns (sqi, em, {variant: [eps per q]})."""
    from .chiqw import Chiqw, from_system
    from . import gen
    w = from_system(s)
    e_cut = truncation_energy(s, n_trunc)
    sqi = [tuple(gen._fold(v) for v in k) for k in w.ski]
    out = {"full": [], "closure": []}
    for nv in plain:
        out[f"plain{nv}"] = []
    c = Chiqw(device)
    try:
        info = c.open(w, ecut_for_eps=ecut_eps, num_freq_grid=nen)
        em, vol = np.asarray(info["em"]), float(info["volume"])
        for q in sqi:
            c.set_window(0.0, n_full)
            out["full"].append(c.run_q(q, "fp64")["epsinv"])
            for nv in plain:
                c.set_window(0.0, nv)
                out[f"plain{nv}"].append(c.run_q(q, "fp64")["epsinv"])
            if net is None:
                continue
            c.set_window(0.0, n_trunc)
            tr = c.run_q(q, "fp64", want_chi=True, want_eps=False)
            x, pairs = features(tr["chi"], tr["lg"], q, s.bvec, em, vol, e_cut)
            d = scaler.t_inv(net.forward(scaler.x(x)))
            nd, ne = tr["chi"].shape[:2]
            dz = (d[:, 0] + 1j * d[:, 1]).reshape(nd, ne, len(pairs)) / vol
            chi = tr["chi"].copy()
            chi[:, :, pairs[:, 1], pairs[:, 0]] += dz
            out["closure"].append(c.dyson_from(chi))
    finally:
        c.close()
    if net is None:
        del out["closure"]
    return sqi, em, out


def w_matrix(s, sqi, em, eps_list, ecut_eps: float, nwf: int, wan_seed: int, device: int = 0):
    """W_ij(R, w) of calc_w3d (GPU) for model Wannier functions of system s and
    eps^-1 per irreducible q (at Gamma the average of the three directions, as
    calc_intW.F:595-612 does)."""
    from .chiqw import from_system
    from . import gen
    from .w3d import W3d
    w = from_system(s)
    wan, _ = gen.make_wannier(s, nwf, wan_seed)
    ntg = max(c.shape[1] for c in wan)
    W = np.zeros((len(wan), nwf, ntg), np.complex128)
    for k, cf in enumerate(wan):
        W[k, :, :cf.shape[1]] = cf
    g = W3d(device)
    try:
        g.open(w, ecut_eps, em)
        g.set_wannier(W)
        g.set_q(np.array(sqi))
        for iq, e in enumerate(eps_list):
            g.set_epsinv(iq, e.mean(axis=0))
        _, wm, _, _ = g.vw()
    finally:
        g.close()
    return wm


def run(n_train: int = 16, n_test: int = 4, nk: int = 3, nocc: int = 8, n_trunc: int = 50,
        n_full: int = 400, plain=(50, 100, 200), ecut_psi: float = 24.0, ecut_eps: float = 8.0,
        nen: int = 16, steps: int = 3000, batch: int = 4096, lr: float = 1e-3, nwf: int = 5,
        seed0: int = 5000, device: int = 0, log=print, q_per_system: int | None = None,
        test_nen: int | None = None) -> dict:
    """Trains the closure on n_train family systems and reports, on n_test
    held-out ones, the error of the static W_ij (w = 0, every i, j and R; meV,
    against the full band set) of plain truncation at each count in `plain`
    and of n_trunc + closure.  q_per_system: training q drawn per system (all
    irreducible q when None); test_nen: frequency points of the held-out runs
    (w = 0 is the first point of every grid, so 1 gives the static W alone)."""
    import time
    t0 = time.perf_counter()
    xs, ts = [], []
    for i in range(n_train):
        s, a, gap = family_system(i, nk, nocc, n_full, ecut_psi, seed0)
        x, t = training_samples(s, ecut_eps, nen, n_trunc, n_full, device, q_per_system, seed0 + i)
        xs.append(x)
        ts.append(t)
        log(f"train system {i}: a={a:.2f} gap={gap:.2f} eV, {len(x)} samples")
    x = np.concatenate(xs)
    t = np.concatenate(ts)
    sc = Scaler(x, t)
    t_data = time.perf_counter() - t0
    net = Mlp(device, nin=NIN, nout=NOUT, width=WIDTH, depth=DEPTH, seed=seed0)
    try:
        t1 = time.perf_counter()
        losses = net.train(sc.x(x), sc.t(t), steps=steps, batch=batch, lr=lr, seed=seed0)
        t_train = time.perf_counter() - t1
        log(f"trained {steps} steps on {len(x)} samples in {t_train:.1f} s: "
            f"loss {losses[:50].mean():.4f} -> {losses[-50:].mean():.5f}")
        errs, wscale = {}, []
        for j in range(n_test):
            s, a, gap = family_system(n_train + j, nk, nocc, n_full, ecut_psi, seed0)
            sqi, em, eps = eps_variants(s, ecut_eps, test_nen or nen, n_trunc, n_full, plain, net,
                                        sc, device)
            # static W (w = 0, the calc_ifreq = 1 default of calc_int), every i, j and R
            wfull = w_matrix(s, sqi, em, eps["full"], ecut_eps, nwf, seed0 + j, device)[0]
            wscale.append(float(np.abs(wfull).max() * AU_EV))
            for key, lst in eps.items():
                if key == "full":
                    continue
                wv = w_matrix(s, sqi, em, lst, ecut_eps, nwf, seed0 + j, device)[0]
                errs.setdefault(key, []).append(float(np.abs(wv - wfull).max() * AU_EV * 1e3))
            log(f"test system {j}: a={a:.2f} gap={gap:.2f} eV, max |W| {wscale[-1]:.3f} eV: " +
                ", ".join(f"{k} {v[-1]:.2f} meV" for k, v in errs.items()))
    finally:
        net.close()
    return {"samples": int(len(x)), "steps": steps, "train_systems": n_train,
            "test_systems": n_test, "loss_first": float(losses[:50].mean()),
            "loss_last": float(losses[-50:].mean()), "data_s": t_data, "train_s": t_train,
            "w_static_max_ev": float(np.mean(wscale)),
            "w_err_mev": {k: float(np.mean(v)) for k, v in errs.items()},
            "w_err_mev_max": {k: float(np.max(v)) for k, v in errs.items()}}
