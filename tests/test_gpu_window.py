"""Per-q windows of the batched configuration on the GPU (rp_chiqw_set_window):
a per-q chi cutoff and a per-q empty-band window give the oracle's chi and
eps^-1 for the same window (FP64 within 1e-6 relative Frobenius), q after q on
one opened system, and clearing the window restores the unwindowed result.
"""
import numpy as np
import pytest

from respackdev_amd import gen
from respackdev_amd.chiqw import Chiqw, from_system

pytestmark = pytest.mark.gpu


def relf(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def test_per_q_windows_match_oracle():
    from oracle import pipeline
    wl = gen.WORKLOADS["c1_plain_fp64"]
    s = gen.make_system(**{k: wl[k] for k in ("a", "nk", "nocc", "nvir", "ecut_psi", "seed")})
    w = from_system(s)
    cases = [((0.0, 0.0, 0.25), 5.0, 10), ((0.25, 0.0, 0.25), 6.0, 17), ((0.0, 0.5, 0.25), 4.0, 24)]
    g = Chiqw(0)
    try:
        g.open(w, ecut_for_eps=wl["ecut_eps"], num_freq_grid=wl["nen"])
        for q, ecut, nvir in cases:
            g.set_window(ecut, nvir)
            r = g.run_q(q, "fp64", want_chi=True)
            o = pipeline.run_q(w, q, ecut_for_eps=wl["ecut_eps"], num_freq_grid=wl["nen"],
                               ecut_eps_q=ecut, n_vir_window=nvir)
            ec, ee = relf(r["chi"], o["chi"]), relf(r["epsinv"], o["epsinv"])
            print(f"q={q} ecut={ecut} nvir={nvir} NG={r['chi'].shape[-1]}: chi {ec:.3e} "
                  f"epsinv {ee:.3e}", flush=True)
            assert r["chi"].shape == o["chi"].shape
            assert ec <= 1e-6 and ee <= 1e-6
        g.set_window(0.0, 0)
        q = cases[0][0]
        r = g.run_q(q, "fp64", want_chi=True)
        o = pipeline.run_q(w, q, ecut_for_eps=wl["ecut_eps"], num_freq_grid=wl["nen"])
        assert relf(r["chi"], o["chi"]) <= 1e-6
    finally:
        g.close()
