"""Constrained RPA on the GPU (rp_chiqw_set_crpa, through the C ABI): a
projector U(n,w,k) on a k-dependent band window (Ns(k), Nb(k), as read from
dir-wan/dat.ns-nb and dat.umat, chiqw.F:870-905) gives the oracle's chi and
eps^-1 for P(n,k) = sum_w |U|^2 (FP64 within 1e-6 relative Frobenius, q != 0
and q = 0); the split-BF16 path stays within 1e-5 of the FP64 GPU path with the
projector on; reopening the system clears it.
"""
import numpy as np
import pytest

from respackdev_amd import gen
from respackdev_amd.chiqw import Chiqw, from_system

pytestmark = pytest.mark.gpu


def relf(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def _system(name):
    wl = gen.WORKLOADS[name]
    s = gen.make_system(**{k: wl[k] for k in ("a", "nk", "nocc", "nvir", "ecut_psi", "seed")})
    return wl, from_system(s)


def _projector(ntk, nbands, nocc, nwf, seed):
    """A window that starts at or one below the top occupied band and holds
    3 or 4 bands, varying with k, and a random U scaled so that every band's
    weight lies in (0, 1]."""
    rng = np.random.default_rng(seed)
    ns = (nocc - 1 - rng.integers(0, 2, ntk)).astype(np.int32)
    nb = np.minimum(rng.integers(3, 5, ntk), nbands - ns).astype(np.int32)
    mb = int(nb.max())
    u = rng.standard_normal((ntk, nwf, mb)) + 1j * rng.standard_normal((ntk, nwf, mb))
    u /= np.sqrt((np.abs(u) ** 2).sum(axis=1, keepdims=True)) * rng.uniform(1.0, 1.6, (ntk, 1, mb))
    prob = np.zeros((nbands, ntk), np.complex128)
    for k in range(ntk):
        for ib in range(nb[k]):
            prob[ns[k] + ib, k] = np.sum(np.abs(u[k, :, ib]) ** 2)
    return ns, nb, u, prob


@pytest.mark.parametrize("name,q", [("w3d_tiny", None), ("w3d_tiny", (0.0, 0.0, 0.0)),
                                    ("c1_plain_fp64", None)])
def test_crpa_matches_oracle(name, q):
    from oracle import pipeline
    wl, w = _system(name)
    q = wl["q"] if q is None else q
    g = Chiqw(0)
    try:
        info = g.open(w, ecut_for_eps=wl["ecut_eps"], num_freq_grid=wl["nen"], flg_crpa=1)
        ntk, nocc = info["ntk"], info["n_occ"]
        nbands = info["n_docc"] + info["n_vir"]
        ns, nb, u, prob = _projector(ntk, nbands, nocc, 2, seed=31)
        g.set_crpa(ns, nb, u)
        r = g.run_q(q, "fp64", want_chi=True)
        o = pipeline.run_q(w, q, ecut_for_eps=wl["ecut_eps"], num_freq_grid=wl["nen"], prob=prob)
        plain = pipeline.run_q(w, q, ecut_for_eps=wl["ecut_eps"], num_freq_grid=wl["nen"])
        ec, ee = relf(r["chi"], o["chi"]), relf(r["epsinv"], o["epsinv"])
        shift = relf(o["chi"], plain["chi"])
        print(f"{name} q={q}: chi {ec:.3e} epsinv {ee:.3e} (cRPA moves chi by {shift:.3f})",
              flush=True)
        assert shift > 0.01  # the projector is not a no-op on this case
        assert ec <= 1e-6 and ee <= 1e-6
        if name == "c1_plain_fp64":
            m = g.run_q(q, "bf16x3", want_chi=True)
            em, eme = relf(m["chi"], r["chi"]), relf(m["epsinv"], r["epsinv"])
            print(f"bf16x3 vs fp64 with cRPA: chi {em:.3e} epsinv {eme:.3e}", flush=True)
            assert em <= 1e-5 and eme <= 1e-5
        g.open(w, ecut_for_eps=wl["ecut_eps"], num_freq_grid=wl["nen"])
        r2 = g.run_q(q, "fp64", want_chi=True)
        assert relf(r2["chi"], plain["chi"]) <= 1e-6
    finally:
        g.close()
