"""FFT grids beyond one workgroup's LDS: a cell of the size of the reference's
SrVO3 sample (simple cubic a = 7.2613 bohr, Ecut_psi = 81 Ry,
sample/quantum-espresso/SrVO3.sc.6x6x6/SrVO3.scf.in:13,16), whose grid under
the reference's rule (chiqw.F:389-401) is 24^3 = 13824 points: the FP64 grid
(221 KB) does not fit the 160 KB of LDS, so the band grids are transformed
in three line passes over HBM, and calc_int's pair-density sums (the psi
sphere spans every z plane) too.  The chi sphere at the default cutoff
(Ecut_psi/10) spans few z planes, so the per-q pruned transform stays in LDS.

Reduced band set and k mesh (2^3 k, 2 occupied + 3 empty bands, ~4700 plane
waves): FP64 chi and eps^-1 against the oracle within 1e-6; calc_w3d's func
and funcw (two model Wannier functions, random eps^-1) within 1e-6.

The same cell with an eps cutoff of 32 Ry (NG_eps 1088-1189): calc_j3d's X and
J and calc_w3d's func and funcw against the oracle within 1e-6 (neither
contraction has a row limit: both take 16-row blocks of eps^-1 per pass).
"""
import numpy as np
import pytest

import oracle
from respackdev_amd import gen
from respackdev_amd.chiqw import Chiqw, from_system
from respackdev_amd.w3d import AU, W3d

pytestmark = pytest.mark.gpu

CELL = dict(a=7.2613, nk=2, nocc=2, nvir=3, ecut_psi=81.0, seed=1310)
Q = (0.0, 0.0, 0.5)


def relf(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.fixture(scope="module")
def srvo3_sized():
    return gen.make_system(**CELL)


def test_chiqw_on_a_24_cubed_grid(srvo3_sized):
    from oracle import pipeline
    w = from_system(srvo3_sized)
    o = pipeline.run_q(w, Q, num_freq_grid=16)
    assert tuple(o["setup"]["nw"]) == (24, 24, 24)
    g = Chiqw(0)
    try:
        info = g.open(w, num_freq_grid=16)
        assert tuple(info["nw"]) == (24, 24, 24)
        r = g.run_q(Q, "fp64", want_chi=True)
    finally:
        g.close()
    e_chi, e_eps = relf(r["chi"], o["chi"]), relf(r["epsinv"], o["epsinv"])
    t = r["timings"]
    K = o["npair"] * info["ntk"]
    nr, ng = 24 ** 3, r["chi"].shape[-1]
    gbs = 8.0 * (2 * nr + ng) * K / t["pairdens"] / 1e9
    print(f"24^3 grid, NG {ng}, {o['npair']} pairs: chi {e_chi:.3e} eps^-1 {e_eps:.3e}; pair "
          f"densities {t['pairdens'] * 1e3:.3f} ms ({gbs:.1f} GB/s compulsory)", flush=True)
    assert e_chi <= 1e-6 and e_eps <= 1e-6


def _w3d_inputs(s, nwf, nen, ecut_eps, seed):
    w = from_system(s)
    wan, _ = gen.make_wannier(s, nwf, 7)
    ntg = max(c.shape[1] for c in wan)
    W = np.zeros((len(wan), nwf, ntg), np.complex128)
    for k, c in enumerate(wan):
        W[k, :, :c.shape[1]] = c
    em = oracle.gen_grid(nen, 50.0 / 27.21151)
    sqi = np.array([[gen._fold(x) for x in k] for k in w.ski])
    rng = np.random.default_rng(seed)
    eps = []
    for q in sqi:
        n = len(oracle.make_kg0(s.bvec, ecut_eps, q))
        e = np.eye(n)[None] + 0.05 * (rng.standard_normal((nen, n, n)) +
                                      1j * rng.standard_normal((nen, n, n)))
        eps.append(e)
    return w, W, em, sqi, eps


def test_w3d_on_a_24_cubed_grid(srvo3_sized):
    from oracle import pipeline_w3d
    nwf, nen, ecut_eps = 2, 6, 8.1
    w, W, em, sqi, eps = _w3d_inputs(srvo3_sized, nwf, nen, ecut_eps, 3)
    o = pipeline_w3d.run_w3d(w, W, sqi, eps, em, ecut_eps)
    g = W3d(0)
    try:
        info = g.open(w, ecut_eps, em)
        assert (info["nwx"], info["nwy"], info["nwz"]) == (24, 24, 24)
        g.set_wannier(W)
        g.set_q(sqi)
        for iq, e in enumerate(eps):
            g.set_epsinv(iq, e)
        func, funcw = g.funcs()
    finally:
        g.close()
    ef, ew = relf(func, o["func"]), relf(funcw, o["funcw"])
    print(f"w3d on 24^3: func {ef:.3e} funcw {ew:.3e}", flush=True)
    assert ef <= 1e-6 and ew <= 1e-6


def test_w3d_and_j3d_with_more_than_1024_eps_rows(srvo3_sized):
    from oracle import pipeline_w3d
    nwf, nen, ecut_eps = 2, 2, 32.0
    rlist = [(0, 0, 0), (1, 0, 0)]
    w, W, em, sqi, eps = _w3d_inputs(srvo3_sized, nwf, nen, ecut_eps, 5)
    assert min(e.shape[-1] for e in eps) > 1024
    oj = pipeline_w3d.run_j3d(w, W, sqi, eps, em, ecut_eps, rlist)
    ow = pipeline_w3d.run_w3d(w, W, sqi, eps, em, ecut_eps)
    g = W3d(0)
    try:
        g.open(w, ecut_eps, em)
        g.set_wannier(W)
        g.set_q(sqi)
        for iq, e in enumerate(eps):
            g.set_epsinv(iq, e)
        x, j, _, _ = g.j3d(rlist)
        func, funcw = g.funcs()
    finally:
        g.close()
    jo = np.moveaxis(oj["J"], 1, 0)  # [ie][r][j][i]
    ex, ej = relf(x, oj["X"]), relf(j, jo)
    ef, ew = relf(func, ow["func"]), relf(funcw, ow["funcw"])
    print(f"NG_eps {max(e.shape[-1] for e in eps)}: X {ex:.3e} J {ej:.3e} func {ef:.3e} "
          f"funcw {ew:.3e}", flush=True)
    assert ex <= 1e-6 and ej <= 1e-6 and ef <= 1e-6 and ew <= 1e-6
    assert np.abs(j - jo).max() * AU <= 1e-3
