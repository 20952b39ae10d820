"""Pins the oracle where no compiled reference routine can (the inline chi
accumulation of calc_chiqw's MPI driver, chiqw.F:1397-1476) and adds closed
forms, library routines and invariants beside the compiled-reference tests
(test_oracle_vs_ref.py, which also holds the Dyson step calc_eps_rpa)."""
import numpy as np
import pytest

import oracle
from oracle import ptr

AU = 27.21151
B7 = np.eye(3) * (2 * np.pi / 7.0)


def test_zinv_matches_numpy_inverse():
    rng = np.random.default_rng(0)
    for n in (1, 2, 7, 40, 101):
        a = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)) + 3 * np.eye(n)
        np.testing.assert_allclose(oracle.zinv(a), np.linalg.inv(a), rtol=1e-10, atol=1e-12)


def test_zinv_pivots_and_detects_singularity():
    a = np.array([[0, 1], [1, 0]], np.complex128)  # needs a row swap
    np.testing.assert_allclose(oracle.zinv(a), a)
    with pytest.raises(np.linalg.LinAlgError):
        oracle.zinv(np.zeros((3, 3), np.complex128))


def test_eps_rpa_q0_is_the_small_q_limit():
    """The q = 0 head/wing assembly (ro_eps_rpa_q0, chiqw_sub.F:459-568) is
    the q -> 0 limit of the q != 0 one once the G = 0 matrix elements carry
    their 1/|q| (the head of chi scaled by |q|^2, the wings by |q|): at a tiny
    q along any direction both give the same eps^-1 (wd = 0).  A wing with two
    1/|q+G| factors, or a head with one, fails by orders of magnitude."""
    rng = np.random.default_rng(3)
    lg = oracle.make_kg0(B7, 3.0, (0, 0, 0))
    n = len(lg)
    g0 = int(np.where((lg == 0).all(1))[0][0])
    chi = (rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))) * 1e-3
    chi[g0, g0] = 0.02 + 0.003j
    b = [np.ascontiguousarray(B7[i]) for i in range(3)]
    out0 = np.zeros((n, n), np.complex128, order="F")
    chif = np.asfortranarray(chi)
    assert oracle.lib().ro_eps_rpa_q0(n, ptr(lg), ptr(b[0]), ptr(b[1]), ptr(b[2]), g0, 0.0,
                                      0.1 / 27.21151, 0.3, chif.ctypes.data,
                                      out0.ctypes.data) == 0
    for d in (np.array([0.0, 0.0, 1.0]), np.array([0.3, -0.5, 0.8])):
        q = 1e-7 * d / np.linalg.norm(d @ B7)  # |q| = 1e-7 bohr^-1
        qa = np.linalg.norm(q @ B7)
        c = chi.copy()
        c[g0, :] *= qa
        c[:, g0] *= qa
        out = np.zeros((n, n), np.complex128, order="F")
        assert oracle.lib().ro_eps_rpa(n, ptr(lg), ptr(np.ascontiguousarray(q)), ptr(b[0]),
                                       ptr(b[1]), ptr(b[2]), np.asfortranarray(c).ctypes.data,
                                       out.ctypes.data) == 0
        np.testing.assert_allclose(out, out0, rtol=0, atol=1e-6 * np.abs(out0).max())


def test_flat_band_tetrahedron_closed_form():
    # all four vertices degenerate -> x(k,w) = 1/(w - D + i d) - 1/(w + D - i d)
    nkb = (4, 4, 4)
    idx = oracle.mkidx(nkb, B7)
    em = oracle.gen_grid(16, 50 / AU)
    e1 = np.full(64, -3.0 / AU)
    e2 = np.full(64, 5.0 / AU)
    d = 0.1 / AU
    x = oracle.ttrhdrn_chi0(nkb, idx, e1, e2, 0.5 / AU, d, 1.0, em)
    x = oracle.ttrhdrn_chi0(nkb, idx, e1, e2, 0.5 / AU, d, -1.0, em, xo=x)
    dl = (5.0 + 3.0) / AU
    want = 1.0 / (em - dl + 1j * d) - 1.0 / (em + dl - 1j * d)
    np.testing.assert_allclose(x, np.broadcast_to(want, x.shape), rtol=1e-12)


def test_linear_band_tetrahedron_against_quadrature():
    # non-degenerate vertices: the tetrahedron integral of the linearly
    # interpolated 1/(w + g(r)), checked against brute-force quadrature on one
    # tetrahedron (closed forms vs numerical integration)
    rng = np.random.default_rng(4)
    g = -(rng.uniform(0.1, 0.6, 4)) + 0.08j
    em = np.array([0.0, 0.2, 0.35, 0.9], np.complex128)
    fk = np.ones(4, np.complex128)
    xo = np.zeros((4, len(em)), np.complex128)
    oracle.lib().ro_ttrhdrn_calc(1e-3, 1e-3, ptr(fk), ptr(g), len(em), ptr(em), ptr(xo), 1)
    # quadrature over the unit simplex (volume 1/6): weight of vertex v is its
    # barycentric coordinate; result scaled by 6 * (1/24 per-vertex normalisation)
    n = 2000000
    u = rng.random((n, 3))
    u.sort(axis=1)
    bary = np.stack([u[:, 0], u[:, 1] - u[:, 0], u[:, 2] - u[:, 1], 1 - u[:, 2]], 1)
    gg = bary @ g
    for ie, w in enumerate(em):
        f = 1.0 / (w + gg)
        q = (bary * f[:, None]).mean(0) / 6.0  # integral over volume-1/6 simplex / (1/6) * (1/6)
        np.testing.assert_allclose(xo[:, ie], q, rtol=1.5e-3)


def test_pairdens_single_plane_wave_closed_form():
    # psi_k = e^{i(k+G1)r}, psi_kq = e^{i(k+q+G2)r}: M(G) = delta(G2 = G1 + G + dG)
    kk = (0.0, 0.0, 0.25)
    kg = oracle.make_kg0(B7, 4.0, kk)
    lg = oracle.make_kg0(B7, 2.0, (0, 0, 0.25))
    nw = oracle.lib().ro_algn235(int(2 * np.sqrt(4.0) / B7[0, 0]) + 1)
    i1 = 3
    ck = np.zeros(len(kg), np.complex128); ck[i1] = 1.0
    for j in (1, 5):
        target = kg[i1] + lg[j] + np.array([0, 0, 1])
        hit = np.where((kg == target).all(1))[0]
        if len(hit) == 0:
            continue
        ckq = np.zeros(len(kg), np.complex128); ckq[hit[0]] = 1.0
        m = np.zeros(len(lg), np.complex128)
        sg = np.array([0, 0, 1], np.int32)
        oracle.lib().ro_pairdens(len(kg), ptr(kg), ptr(ck), len(kg), ptr(kg), ptr(ckq),
                                 nw, nw, nw, len(lg), ptr(lg), ptr(sg), ptr(m))
        want = np.zeros(len(lg)); want[j] = 1.0
        np.testing.assert_allclose(m, want, atol=1e-7)


def test_chi_of_plane_wave_states_closed_form():
    """The oracle's chi and eps^-1 (ro_chiqw_q, ro_eps_rpa) for bands made of
    one or two plane waves with flat energies, against the closed form of
    tests/plane_wave_case.py: pins 1/NTK, 1/V, the spin factor, 4 pi and the
    conj placement of chiqw.F:1397-1410."""
    import plane_wave_case as pw
    from oracle import pipeline
    w = pw.wfn()
    o = pipeline.run_q(w, pw.Q, ecut_for_eps=pw.ECUT_EPS, num_freq_grid=pw.NEN)
    delt = 0.1 / AU
    lg, chi, epsinv = pw.closed_form(o["em"], delt)
    np.testing.assert_array_equal(o["lg"], lg)
    e_chi = np.linalg.norm(o["chi"][0] - chi) / np.linalg.norm(chi)
    e_eps = np.linalg.norm(o["epsinv"][0] - epsinv) / np.linalg.norm(epsinv)
    # the two-plane-wave bands give off-diagonal elements (conj placement)
    off = chi - np.einsum("eii->ei", chi)[:, :, None] * np.eye(chi.shape[-1])[None]
    assert np.linalg.norm(off) > 0.3 * np.linalg.norm(chi)
    assert np.linalg.norm(np.conj(np.swapaxes(off, 1, 2)) - off) > 0.1 * np.linalg.norm(off)
    assert e_chi <= 1e-6 and e_eps <= 1e-6, (e_chi, e_eps)


@pytest.mark.parametrize("cls", ["normal", "pair", "triple", "two_pairs", "all_equal"])
def test_tetrahedron_series_against_quad_closed_forms(cls):
    """The series evaluation of calc's vertex weights (oracle/tet_forms.h)
    against the reference's closed forms of the same class evaluated in
    __float128, on one tetrahedron per class at frequencies that put
    r = max|g~ - c| / |e + c| between 1e-4 and 0.5.  There the double closed
    forms lose up to ~1e16 ulps (a relative error of order one at r = 1e-4)
    and the quad ones keep ~1e-18: every weight within 1e-14 of the largest."""
    rng = np.random.default_rng({"normal": 1, "pair": 2, "triple": 3, "two_pairs": 4,
                                 "all_equal": 5}[cls])
    delt = 0.004
    base = {"normal": [0.0, 0.031, 0.074, 0.12], "pair": [0.0, 2e-4, 0.05, 0.11],
            "triple": [0.0, 3e-4, 5e-4, 0.09], "two_pairs": [0.0, 4e-4, 0.08, 0.0806],
            "all_equal": [0.0, 1e-4, 3e-4, 6e-4]}[cls]
    g = -(0.3 + np.array(base) + rng.uniform(0, 1e-5, 4)) + 1j * delt
    fk = np.ones(4, np.complex128)
    # frequencies e with max_j |g_j - c| / |e + c| = r (c the mean vertex value)
    c = g.real.mean()
    dm = np.abs(g.real - c).max() + 1e-300
    r = np.geomspace(1e-4, 0.45, 24)
    em = (-c + dm / r).astype(np.complex128)
    x = np.zeros((4, len(em)), np.complex128)
    xq = np.zeros_like(x)
    assert oracle.lib().ro_ttrhdrn_calc(1e-3, 1e-3, ptr(fk), ptr(g), len(em), ptr(em), ptr(x), 1) == 0
    assert oracle.lib().ro_ttrhdrn_calc_quad(1e-3, 1e-3, ptr(fk), ptr(g), len(em), ptr(em),
                                             ptr(xq)) == 0
    err = np.abs(x - xq).max(axis=0) / np.abs(xq).max(axis=0)
    print(f"{cls}: series vs quad closed forms, worst {err.max():.1e} (relative per frequency)")
    assert err.max() <= 1e-14
