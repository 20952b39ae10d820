"""Parity of the gfx950 calc_w3d path (librespack_gpu.so through its C ABI)
against the oracle on the same seeded inputs: a 3^3 k mesh (time-reversal
partners present), 3 model Wannier functions and seeded eps^-1 matrices.

Tolerances: func, funcw, V and W within 1e-6 relative Frobenius of the oracle
(the pair densities are single-precision values on both sides, CMPLX in the
reference, so a float ulp of difference in M can reach ~1e-7 here), and every
V/W element within 1 meV.
"""
import numpy as np
import pytest

from respackdev_amd import gen
from respackdev_amd.w3d import AU, W3d
from test_oracle_w3d import _w3d_case

pytestmark = pytest.mark.gpu


def relf(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.mark.parametrize("kind", ["random", "identity"])
def test_w3d_matches_oracle(kind):
    from oracle import pipeline_w3d
    wl, s, w, W, em, sqi, eps = _w3d_case(kind)
    o = pipeline_w3d.run_w3d(w, W, sqi, eps, em, wl["ecut_eps"])
    g = W3d(0)
    try:
        info = g.open(w, wl["ecut_eps"], em)
        sk0, ng0 = g.kmesh()
        # the library's full mesh and G lists are the ones dat.wan was written for
        want_k, _ = gen.full_kmesh(s.ski)
        np.testing.assert_allclose(sk0, want_k, atol=1e-12)
        assert list(ng0) == [len(k) for k in gen.full_kg(s)]
        g.set_wannier(W)
        ng = g.set_q(sqi)
        assert list(ng) == o["nge"][:len(sqi)]
        for iq, e in enumerate(eps):
            g.set_epsinv(iq, e)
        func, funcw = g.funcs()
        v, wm, chead, cheadw = g.vw()
    finally:
        g.close()
    print(f"{kind}: func {relf(func, o['func']):.3e} funcw {relf(funcw, o['funcw']):.3e} "
          f"V {relf(v, o['vmat']):.3e} W {relf(wm, o['wmat']):.3e}", flush=True)
    assert abs(chead - o["chead"]) <= 1e-14 * abs(o["chead"])
    np.testing.assert_allclose(cheadw, o["cheadw"], rtol=1e-14, atol=0)
    assert relf(func, o["func"]) <= 1e-6
    assert relf(funcw, o["funcw"]) <= 1e-6
    assert relf(v, o["vmat"]) <= 1e-6
    assert relf(wm, o["wmat"]) <= 1e-6
    assert np.abs(v - o["vmat"]).max() * AU <= 1e-3
    assert np.abs(wm - o["wmat"]).max() * AU <= 1e-3


def test_j3d_matches_oracle():
    from oracle import pipeline_w3d
    rlist = [(0, 0, 0), (1, 0, 0), (0, -1, 1)]
    wl, s, w, W, em, sqi, eps = _w3d_case("random")
    o = pipeline_w3d.run_j3d(w, W, sqi, eps, em, wl["ecut_eps"], rlist)
    g = W3d(0)
    try:
        g.open(w, wl["ecut_eps"], em)
        g.set_wannier(W)
        g.set_q(sqi)
        for iq, e in enumerate(eps):
            g.set_epsinv(iq, e)
        x, j, chead, cheadw = g.j3d(rlist)
    finally:
        g.close()
    jo = np.moveaxis(o["J"], 1, 0)  # [ie][r][j][i]
    print(f"X {relf(x, o['X']):.3e} J {relf(j, jo):.3e}", flush=True)
    assert relf(x, o["X"]) <= 1e-6 and relf(j, jo) <= 1e-6
    assert np.abs(x - o["X"]).max() * AU <= 1e-3 and np.abs(j - jo).max() * AU <= 1e-3


def test_six_wannier_functions():
    """Six Wannier functions: W's contraction with six accumulators, and J's
    36 pair vectors in two workgroup groups of 32 (the second one partial)."""
    from oracle import pipeline_w3d
    wl, s, w, W, em, sqi, eps = _w3d_case("random", nwf=6, seed=11)
    o = pipeline_w3d.run_w3d(w, W, sqi, eps, em, wl["ecut_eps"])
    rlist = [(0, 0, 0), (0, 1, 0)]
    oj = pipeline_w3d.run_j3d(w, W, sqi, eps, em, wl["ecut_eps"], rlist)
    g = W3d(0)
    try:
        g.open(w, wl["ecut_eps"], em)
        g.set_wannier(W)
        g.set_q(sqi)
        for iq, e in enumerate(eps):
            g.set_epsinv(iq, e)
        func, funcw = g.funcs()
        v, wm, _, _ = g.vw()
        x, j, _, _ = g.j3d(rlist)
    finally:
        g.close()
    jo = np.moveaxis(oj["J"], 1, 0)
    print(f"nwf 6: funcw {relf(funcw, o['funcw']):.3e} W {relf(wm, o['wmat']):.3e} "
          f"X {relf(x, oj['X']):.3e} J {relf(j, jo):.3e}", flush=True)
    assert relf(funcw, o["funcw"]) <= 1e-6 and relf(wm, o["wmat"]) <= 1e-6
    assert relf(x, oj["X"]) <= 1e-6 and relf(j, jo) <= 1e-6
    assert np.abs(j - jo).max() * AU <= 1e-3
