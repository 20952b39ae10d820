"""calc_int at its full workload size on the GPU: 6^3 k mesh (NTK = NTQ = 216,
112 irreducible q with time reversal), five d-like model Wannier functions,
chi cutoff 22.2 Ry (NG_eps 596-624, so the W contraction runs its second row
set and a partial last column pass, and the J contraction several row passes),
~1000 plane waves, 32 frequencies.

eps^-1 per irreducible q is synthetic and different for every q, frequency
and element: identity + 0.05 e^{i theta_q} R + 0.05 u v^H / NG with a seeded
random R shared by all q (sliced to NG) and seeded u, v per (q, w).

* func and funcw at four full-mesh q (Gamma, a time-reversal partner, two
  others) against the oracle (the same calc_intW.F:647-912 restatement the
  small tests pin): within 1e-6 relative Frobenius.
* Over all q: V Hermitian in R (V_ji(-R) = conj(V_ij(R) - chead) + chead),
  X_ii(0) = V_ii(0) and J_ii(0, w) = W_ii(0, w) (the exchange pair density of
  (i, i) at R = 0 is the direct one of i), to 1e-9 relative.
* eps^-1 = 1 on every q: W(R, w) = V(R) for every frequency.
"""
import numpy as np
import pytest

from respackdev_amd import _lib, gen
from respackdev_amd.chiqw import from_system
from respackdev_amd.w3d import W3d

pytestmark = pytest.mark.gpu

WL = gen.WORKLOADS["c4_calc_int"]


def relf(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def c4_case():
    s = gen.make_system(**{k: WL[k] for k in ("a", "nk", "nocc", "nvir", "ecut_psi", "seed")})
    w = from_system(s)
    wan, _ = gen.make_wannier(s, WL["nwf"], WL["seed"])
    ntg = max(c.shape[1] for c in wan)
    W = np.zeros((len(wan), WL["nwf"], ntg), np.complex128)
    for k, c in enumerate(wan):
        W[k, :, :c.shape[1]] = c
    em = np.zeros(WL["nen"], np.complex128)  # the reference's gen_grid rule, 0..50 eV
    _lib.check(_lib.lib().rp_gen_grid(WL["nen"], 50.0 / 27.21151, _lib.ptr(em)))
    sqi = np.array([[gen._fold(x) for x in k] for k in w.ski])
    return s, w, W, em, sqi


class SyntheticEps:
    """eps^-1 of irreducible q iq, [ie][jg][ig] complex128."""

    def __init__(self, nen, ngmax, seed=99):
        rng = np.random.default_rng(seed)
        self.r = (rng.standard_normal((nen, ngmax, ngmax)) +
                  1j * rng.standard_normal((nen, ngmax, ngmax))) / np.sqrt(2.0)
        self.nen, self.seed = nen, seed
        self.ng = {}

    def __call__(self, iq):
        n = self.ng[iq]
        rng = np.random.default_rng((self.seed, iq))
        th = rng.uniform(0, 2 * np.pi)
        u = rng.standard_normal((self.nen, n)) + 1j * rng.standard_normal((self.nen, n))
        v = rng.standard_normal((self.nen, n)) + 1j * rng.standard_normal((self.nen, n))
        e = 0.05 * np.exp(1j * th) * self.r[:, :n, :n]
        e += (0.05 / n) * u[:, :, None] * np.conj(v[:, None, :])
        e[:, np.arange(n), np.arange(n)] += 1.0
        return e


@pytest.fixture(scope="module")
def c4():
    s, w, W, em, sqi = c4_case()
    g = W3d(0)
    info = g.open(w, WL["ecut_eps"], em)
    g.set_wannier(W)
    ng = g.set_q(sqi)
    print(f"c4 opened: NTK {info['ntk']}, {len(ng)} irreducible q, NG_eps {ng.min()}-{ng.max()}",
          flush=True)
    yield g, (s, w, W, em, sqi), ng
    g.close()


def test_c4_func_funcw_match_oracle_at_sampled_q(c4):
    from oracle import pipeline_w3d
    g, (s, w, W, em, sqi), ng = c4
    eps = SyntheticEps(len(em), int(ng.max()))
    eps.ng = {i: int(n) for i, n in enumerate(ng)}
    for iq in range(len(ng)):
        g.set_epsinv(iq, eps(iq))
    func, funcw = g.funcs()
    v, wm, chead, cheadw = g.vw()
    x, j, _, _ = g.j3d([(0, 0, 0)])
    # the oracle at four q: Gamma, the first time-reversal partner, two others
    def pick(c):
        trs_q = [i for i, t in enumerate(c["trsq"]) if t == -1]
        return [0, 1, trs_q[0], len(c["trsq"]) - 1]
    o = pipeline_w3d.funcs_at(w, W, sqi, eps, em, WL["ecut_eps"], pick)
    for n, iq in enumerate(o["iqs"]):
        ef, ew = relf(func[iq], o["func"][n]), relf(funcw[iq], o["funcw"][n])
        print(f"c4 q {iq} (trs {o['trsq'][iq]}, NG_eps {o['nge'][iq]}): func {ef:.3e} "
              f"funcw {ew:.3e}", flush=True)
        assert ef <= 1e-6 and ew <= 1e-6
    # identities over the whole q mesh
    vr = (v - chead)[::-1, ::-1, ::-1]
    assert relf(np.conj(np.swapaxes(vr, -1, -2)) + chead, v) <= 1e-9
    na = [n // 2 for n in (WL["nk"],) * 3]
    v0 = v[na[2], na[1], na[0]]
    w0 = wm[:, na[2], na[1], na[0]]
    assert relf(np.diagonal(x[0]), np.diagonal(v0)) <= 1e-9
    jd = np.diagonal(j[:, 0], axis1=-2, axis2=-1)
    assert relf(jd, np.diagonal(w0, axis1=-2, axis2=-1)) <= 1e-9


def test_c4_identity_eps_gives_w_equal_v(c4):
    g, (s, w, W, em, sqi), ng = c4
    for iq, n in enumerate(ng):
        e = np.zeros((len(em), n, n), np.complex128)
        e[:, np.arange(n), np.arange(n)] = 1.0
        g.set_epsinv(iq, e)
    g.funcs()
    v, wm, chead, cheadw = g.vw()
    np.testing.assert_allclose(cheadw, chead, rtol=1e-14)
    for ie in range(len(em)):
        assert relf(wm[ie], v) <= 1e-12
