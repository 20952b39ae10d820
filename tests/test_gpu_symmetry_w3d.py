"""calc_int under crystal symmetry on the GPU: the 4_1 screw-axis crystal of
tests/symmetric_case.py given by its irreducible k list (four rotations with
fractional translations (0, 0, j/4), time reversal), so that calc_w3d and
calc_j3d rotate eps^-1 from each irreducible q to its star (make_eps,
calc_int_sub.F:222-311: the rotated G map and the phases of the fractional
translations, the transpose under time reversal).  func, funcw, V, W, X and J
against the oracle, whose make_eps matches the compiled reference bit for bit
(tests/test_oracle_w3d.py): within 1e-6 relative Frobenius and 1 meV.
"""
import numpy as np
import pytest

from respackdev_amd.w3d import AU, W3d
from symmetric_case import w3d_case

pytestmark = pytest.mark.gpu


def relf(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def test_screw_symmetry_w3d_and_j3d_match_oracle():
    from oracle import pipeline_w3d
    c = w3d_case()
    # the star of every irreducible q uses rotations and time reversal
    assert set(c["numrotq"]) == {0, 1, 2, 3} and -1 in c["trsq"]
    o = pipeline_w3d.run_w3d(c["w"], c["wan"], c["sqi"], c["eps"], c["em"], c["ecut_eps"])
    rlist = [(0, 0, 0), (1, 0, 0), (0, -1, 1)]
    oj = pipeline_w3d.run_j3d(c["w"], c["wan"], c["sqi"], c["eps"], c["em"], c["ecut_eps"], rlist)
    g = W3d(0)
    try:
        info = g.open(c["w"], c["ecut_eps"], c["em"])
        assert info["ntk"] == 64
        sk0, ng0 = g.kmesh()
        np.testing.assert_allclose(sk0, c["sk0"], atol=1e-12)
        np.testing.assert_array_equal(ng0, c["ng0"])
        g.set_wannier(c["wan"])
        ng = g.set_q(c["sqi"])
        assert list(ng) == o["nge"][:len(c["sqi"])]
        for iq, e in enumerate(c["eps"]):
            g.set_epsinv(iq, e)
        func, funcw = g.funcs()
        v, wm, chead, cheadw = g.vw()
        x, j, _, _ = g.j3d(rlist)
    finally:
        g.close()
    jo = np.moveaxis(oj["J"], 1, 0)
    errs = dict(func=relf(func, o["func"]), funcw=relf(funcw, o["funcw"]), V=relf(v, o["vmat"]),
                W=relf(wm, o["wmat"]), X=relf(x, oj["X"]), J=relf(j, jo))
    print("screw axis: " + " ".join(f"{k} {e:.3e}" for k, e in errs.items()), flush=True)
    assert max(errs.values()) <= 1e-6
    for a, b in ((v, o["vmat"]), (wm, o["wmat"]), (x, oj["X"]), (j, jo)):
        assert np.abs(a - b).max() * AU <= 1e-3
