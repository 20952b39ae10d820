"""Pins the constrained-RPA factor of the oracle (oracle/oracle.c, chi
accumulation, after chiqw.F:1397-1410): every (occupied j at k, empty i at
k+q) term is weighted by 1 - P(j,k) P(i,k+q), where P(n,k) = sum_w |U(n,w,k)|^2
is the band's weight in the Wannier subspace (chiqw.F:897-905).

Checks that fix the factor without retyping it:
- a constant P = p on every band scales chi by 1 - p^2 (sign and square);
- P on the occupied bands alone, or on the empty bands alone, changes nothing
  (the factor is a product of an occupied and an empty weight, not a sum);
- P = 1 on all occupied bands and on the lowest empty band removes exactly the
  chi of the pair list restricted to that empty band, computed by the band
  window instead of the factor;
- k indexing: removing the occupied bands at k0 and removing the empty bands
  at k0+q (and not at k0) take out the same contribution, on a mesh where k->k+q
  is not an involution.
"""
import numpy as np
import pytest

from respackdev_amd import gen
from respackdev_amd.chiqw import from_system

_cache = {}


def _case():
    if "w" not in _cache:
        wl = gen.WORKLOADS["w3d_tiny"]
        s = gen.make_system(**{k: wl[k] for k in ("a", "nk", "nocc", "nvir", "ecut_psi", "seed")})
        _cache["w"] = (wl, from_system(s))
    return _cache["w"]


def _run(q=None, **kw):
    from oracle import pipeline
    wl, w = _case()
    return pipeline.run_q(w, wl["q"] if q is None else q, ecut_for_eps=wl["ecut_eps"],
                          num_freq_grid=wl["nen"], **kw)


def _dims(o):
    s = o["setup"]
    return s["coef"].shape[0], s["ntk"], s["n_occ"]


def _rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.mark.parametrize("q", [None, (0.0, 0.0, 0.0)])
def test_constant_projector_scales_chi(q):
    o = _run(q)
    nb, ntk, _ = _dims(o)
    p = 0.6
    c = _run(q, prob=np.full((nb, ntk), p, np.complex128))["chi"]
    assert _rel(c, (1.0 - p * p) * o["chi"]) <= 1e-13


def test_one_sided_projector_changes_nothing():
    o = _run()
    nb, ntk, nocc = _dims(o)
    P = np.zeros((nb, ntk), np.complex128)
    P[:nocc] = 1.0
    assert _rel(_run(prob=P)["chi"], o["chi"]) <= 1e-15
    P = np.zeros((nb, ntk), np.complex128)
    P[nocc:] = 1.0
    assert _rel(_run(prob=P)["chi"], o["chi"]) <= 1e-15


def test_projector_equals_band_window():
    o = _run()
    nb, ntk, nocc = _dims(o)
    P = np.zeros((nb, ntk), np.complex128)
    P[:nocc + 1] = 1.0
    removed = o["chi"] - _run(prob=P)["chi"]
    lowest_empty = _run(n_vir_window=1)["chi"]
    assert np.linalg.norm(removed) > 0.05 * np.linalg.norm(o["chi"])
    assert _rel(removed, lowest_empty) <= 1e-13


def test_empty_band_weight_is_taken_at_k_plus_q():
    o = _run()
    nb, ntk, nocc = _dims(o)
    ikq = o["ikq"]
    k0 = 4
    assert ikq[ikq[k0]] != k0  # k -> k+q is not an involution here
    P = np.zeros((nb, ntk), np.complex128)
    P[:nocc, k0] = 1.0
    P[nocc:, :] = 1.0
    by_occ = o["chi"] - _run(prob=P)["chi"]
    P = np.zeros((nb, ntk), np.complex128)
    P[:nocc, :] = 1.0
    P[nocc:, ikq[k0]] = 1.0
    by_vir = o["chi"] - _run(prob=P)["chi"]
    P = np.zeros((nb, ntk), np.complex128)
    P[:nocc, :] = 1.0
    P[nocc:, k0] = 1.0
    at_k = o["chi"] - _run(prob=P)["chi"]
    assert np.linalg.norm(by_occ) > 0.01 * np.linalg.norm(o["chi"])
    assert _rel(by_vir, by_occ) <= 1e-12
    assert _rel(at_k, by_occ) > 0.1
