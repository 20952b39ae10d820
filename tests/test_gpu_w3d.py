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



def test_twelve_wannier_functions():
    """Twelve Wannier functions (two sets of six model functions with different
    centres): the pair sums with sixteen iw per thread, J's 144 pair vectors
    in five groups of 32 (the last partial), and W's contraction with one q
    per group of twelve vectors."""
    from oracle import pipeline_w3d
    wl, s, w, W6, em, sqi, eps = _w3d_case("random", nwf=6, seed=13)
    wan2, _ = gen.make_wannier(s, 6, 29)
    W = np.zeros((W6.shape[0], 12, W6.shape[2]), np.complex128)
    W[:, :6] = W6
    for k, c in enumerate(wan2):
        W[k, 6:, :c.shape[1]] = c
    o = pipeline_w3d.run_w3d(w, W, sqi, eps, em, wl["ecut_eps"])
    rlist = [(0, 0, 0)]
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
    print(f"nwf 12: func {relf(func, o['func']):.3e} funcw {relf(funcw, o['funcw']):.3e} "
          f"W {relf(wm, o['wmat']):.3e} X {relf(x, oj['X']):.3e} J {relf(j, jo):.3e}", flush=True)
    assert relf(func, o["func"]) <= 1e-6 and relf(funcw, o["funcw"]) <= 1e-6
    assert relf(v, o["vmat"]) <= 1e-6 and relf(wm, o["wmat"]) <= 1e-6
    assert relf(x, oj["X"]) <= 1e-6 and relf(j, jo) <= 1e-6
    assert np.abs(wm - o["wmat"]).max() * AU <= 1e-3 and np.abs(j - jo).max() * AU <= 1e-3

def test_w3d_wannier_subset_on_a_large_mesh():
    """The pair sums, W and J of two Wannier functions do not depend on which
    others are present.  On an 8^3 mesh (NTK 512) eight Wannier functions leave
    room in LDS for the band grids of at most four grid points per workgroup of
    the pair sums, J's sums run in two groups of jw, and the W contraction takes
    two q per group; the first two alone fit sixteen points and take up to
    eight q per group.  Same arithmetic per (q, pair, point) and per vector
    column, so the two runs agree to rounding."""
    wl = dict(gen.WORKLOADS["w3d_tiny"])
    wl["nk"] = 8
    s = gen.make_system(**{k: wl[k] for k in ("a", "nk", "nocc", "nvir", "ecut_psi", "seed")})
    from respackdev_amd.chiqw import from_system
    w = from_system(s)
    wan, _ = gen.make_wannier(s, 8, 5)
    ntg = max(c.shape[1] for c in wan)
    W8 = np.zeros((len(wan), 8, ntg), np.complex128)
    for k, c in enumerate(wan):
        W8[k, :, :c.shape[1]] = c
    em = np.linspace(0.0, 1.0, 4) + 0.01j
    sqi = np.array([[gen._fold(x) for x in k] for k in w.ski])
    rng = np.random.default_rng(3)
    out = {}
    for nwf in (8, 2):
        g = W3d(0)
        try:
            g.open(w, wl["ecut_eps"], em)
            g.set_wannier(np.ascontiguousarray(W8[:, :nwf]))
            ng = g.set_q(sqi)
            if not out:
                eps = [np.eye(n)[None] + 0.05 * (rng.standard_normal((len(em), n, n)) +
                                                1j * rng.standard_normal((len(em), n, n)))
                       for n in ng]
            for iq, e in enumerate(eps):
                g.set_epsinv(iq, e)
            assert g.info()["ntq"] == 512
            func, funcw = g.funcs()
            x, j, _, _ = g.j3d([(0, 0, 0), (1, 0, -1)])
        finally:
            g.close()
        out[nwf] = (func[:, :2, :2], funcw[:, :, :2, :2], x[:, :2, :2], j[:, :, :2, :2])
    for a, b, name in zip(out[8], out[2], ("func", "funcw", "X", "J")):
        err = relf(a, b)
        print(f"{name}: {err:.2e}", flush=True)
        assert err <= 1e-12, name


def test_sixteen_wannier_functions_on_an_11_cubed_mesh():
    """Sixteen Wannier functions on an 11^3 mesh (NTK 1331): the band grids of
    one grid point (16 x 1331 complex64) exceed the LDS, so the pair sums take
    the per-point form with sixteen iw per thread.  The first two Wannier
    functions alone fit the LDS-tiled form; their func, funcw, X and J agree
    with the sixteen-function run to rounding (the full mode multiplies the
    umklapp phase at a different place)."""
    wl = dict(gen.WORKLOADS["w3d_tiny"])
    wl["nk"] = 11
    s = gen.make_system(**{k: wl[k] for k in ("a", "nk", "nocc", "nvir", "ecut_psi", "seed")})
    from respackdev_amd.chiqw import from_system
    w = from_system(s)
    wans = [gen.make_wannier(s, 6, sd)[0] for sd in (5, 17, 29)]
    ntg = max(c.shape[1] for c in wans[0])
    W16 = np.zeros((len(wans[0]), 16, ntg), np.complex128)
    for k in range(len(wans[0])):
        c = np.concatenate([wn[k] for wn in wans])[:16]
        W16[k, :, :c.shape[1]] = c
    em = np.linspace(0.0, 1.0, 2) + 0.01j
    sqi = np.array([[gen._fold(x) for x in k] for k in w.ski])
    rng = np.random.default_rng(7)
    out = {}
    for nwf in (16, 2):
        g = W3d(0)
        try:
            g.open(w, wl["ecut_eps"], em)
            g.set_wannier(np.ascontiguousarray(W16[:, :nwf]))
            ng = g.set_q(sqi)
            if not out:
                eps = [np.eye(n)[None] + 0.05 * (rng.standard_normal((len(em), n, n)) +
                                                1j * rng.standard_normal((len(em), n, n)))
                       for n in ng]
            for iq, e in enumerate(eps):
                g.set_epsinv(iq, e)
            assert g.info()["ntq"] == 1331
            func, funcw = g.funcs()
            x, j, _, _ = g.j3d([(0, 1, 0)])
        finally:
            g.close()
        out[nwf] = (func[:, :2, :2], funcw[:, :, :2, :2], x[:, :2, :2], j[:, :, :2, :2])
    for a, b, name in zip(out[16], out[2], ("func", "funcw", "X", "J")):
        err = relf(a, b)
        print(f"{name}: {err:.2e}", flush=True)
        assert err <= 1e-12, name
