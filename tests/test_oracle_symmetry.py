"""Symmetry-general input (nsym > 1, fractional translations) through the
oracle: the same crystal given as the irreducible k list of a 4_1 screw axis
and as the identity-only full mesh (tests/symmetric_case.py) yields the same
chi.  The unfolding count is checked against the reference's own est_NTK for
the four operations.
"""
import numpy as np
import pytest

import oracle
from oracle import pipeline, ptr
from symmetric_case import irreducible, make, ROT, _rg_rows

Q = (0.0, 0.25, 0.25)


def relf(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.mark.skipif(not oracle.have_ref(), reason="reference build absent")
def test_screw_unfolding_count_matches_reference():
    ski = np.ascontiguousarray(irreducible(4))
    rg = _rg_rows(ROT)
    n = 2 * len(ski) * len(ROT)
    sk0 = np.zeros((n, 3)); numirr = np.zeros(n, np.int32); numrot = np.zeros(n, np.int32)
    trs = np.zeros(n, np.int32); rw = np.zeros((n, 3), np.int32)
    ntk = oracle.lib().ro_unfold_k(len(ski), len(ROT), ptr(ski), ptr(rg), ptr(sk0), ptr(numirr),
                                   ptr(numrot), ptr(trs), ptr(rw))
    ntk_ref = np.zeros(1, np.int32)
    oracle.ref().rpref_est_ntk(len(ski), len(ROT), ptr(ski), ptr(rg), ptr(ntk_ref))
    assert ntk == ntk_ref[0] == 64
    assert len(ski) == 18              # time reversal alone leaves 36 of the 64 points
    assert set(numrot[:ntk]) == {0, 1, 2, 3} and set(trs[:ntk]) == {1, -1}


def test_screw_irreducible_input_gives_full_mesh_chi():
    wA, wB = make()
    ra = pipeline.run_q(wA, Q, ecut_for_eps=2.0, num_freq_grid=6)
    rb = pipeline.run_q(wB, Q, ecut_for_eps=2.0, num_freq_grid=6)
    # the unfolded wavefunctions are the same numbers in both forms (measured 1e-16)
    assert relf(ra["chi"], rb["chi"]) <= 1e-12
    assert relf(ra["epsinv"], rb["epsinv"]) <= 1e-12
    # a wrong fractional translation changes chi (measured 18%): the check is not vacuous
    wA.pg = np.array([(0, 0, 0), (0, 0, 2), (0, 0, 2), (0, 0, 3)], np.int32)
    rc = pipeline.run_q(wA, Q, ecut_for_eps=2.0, num_freq_grid=6)
    assert relf(rc["chi"], rb["chi"]) > 1e-2
