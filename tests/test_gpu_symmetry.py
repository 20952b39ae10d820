"""Symmetry-general input on the GPU: the irreducible k list of a 4_1 screw
axis (four rotations with fractional translations, time reversal) unfolded by
librespack_gpu gives the same FP64 chi and eps^-1 as the oracle on the
identity-only full mesh of the same crystal (tests/symmetric_case.py), within
the FP64 parity bar of 1e-6 relative Frobenius.
"""
import numpy as np
import pytest

from respackdev_amd.chiqw import Chiqw
from symmetric_case import make

pytestmark = pytest.mark.gpu

Q = (0.0, 0.25, 0.25)


def relf(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def test_screw_symmetry_chi_matches_full_mesh_oracle():
    from oracle import pipeline
    wA, wB = make()
    ref = pipeline.run_q(wB, Q, ecut_for_eps=2.0, num_freq_grid=6)
    out = {}
    for name, w in (("irreducible", wA), ("full", wB)):
        g = Chiqw(0)
        try:
            info = g.open(w, ecut_for_eps=2.0, num_freq_grid=6)
            assert info["ntk"] == 64
            g.prepare_q(Q)
            out[name] = g.run_q(Q, "fp64", want_chi=True)
        finally:
            g.close()
    for name, r in out.items():
        ec = relf(r["chi"], ref["chi"])
        ee = relf(r["epsinv"], ref["epsinv"])
        print(f"{name}: chi {ec:.3e} epsinv {ee:.3e}", flush=True)
        assert ec <= 1e-6 and ee <= 1e-6
