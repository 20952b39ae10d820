"""The batched configuration under graph capture (rp_chiqw_batch_plan/run):
every q of a list, each with its own chi cutoff and empty-band window, run as
one replayed HIP graph, against the eager per-q path of the same handle (same
kernels and launches: agreement to rounding of the atomics, 1e-12), on an
insulator with q = 0 in the list and on a metal (Fermi-surface cuts and the
Drude head inside the graph).  The graph is replayed several times with eager
work on the same handle in between, and every replay is checked; the eager
state of the handle is the same after planning as before.  With autotuning on,
the tuned launch is never slower than the heuristic's on its own timing.
"""
import numpy as np
import pytest

from respackdev_amd import gen
from respackdev_amd.chiqw import Chiqw, from_system

pytestmark = pytest.mark.gpu


def relf(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def _system(name):
    wl = gen.WORKLOADS[name]
    keys = ["a", "nk", "nocc", "nvir", "ecut_psi", "seed"] + (["n_cross"] if "n_cross" in wl else [])
    return wl, from_system(gen.make_system(**{k: wl[k] for k in keys}))


CASES = {
    "c1_plain_fp64": [((0.0, 0.0, 0.25), 6.9, 0), ((0.25, 0.0, 0.25), 5.0, 12),
                      ((0.0, 0.0, 0.0), 6.0, 20), ((0.0, 0.5, 0.25), 4.0, 24),
                      ((0.5, 0.5, 0.5), 6.9, 0)],
    "metal_tiny": [((0.0, 0.0, 1.0 / 3.0), 2.0, 0), ((0.0, 0.0, 0.0), 2.0, 0),
                   ((1.0 / 3.0, 0.0, 1.0 / 3.0), 1.5, 3)],
}


@pytest.mark.parametrize("name,prec", [("c1_plain_fp64", "bf16x3"), ("c1_plain_fp64", "fp64"),
                                       ("metal_tiny", "fp64")])
def test_batch_graph_matches_eager(name, prec):
    wl, w = _system(name)
    cases = CASES[name]
    g = Chiqw(0)
    try:
        g.open(w, ecut_for_eps=wl["ecut_eps"], num_freq_grid=wl["nen"])
        # eager reference, q by q with the same windows
        eager = []
        for q, ecut, nvir in cases:
            g.set_window(ecut, nvir)
            eager.append(g.run_q(q, prec)["epsinv"])
        g.set_window(0.0, 0)
        q_before = wl["q"]
        before = g.run_q(q_before, prec)["epsinv"]
        g.batch_plan([c[0] for c in cases], [c[1] for c in cases], [c[2] for c in cases], prec)
        # the handle's own q state is untouched by the plan
        np.testing.assert_array_equal(g.run_q(q_before, prec)["epsinv"].shape, before.shape)
        first = None
        for rep in range(3):
            t = g.batch_run()
            got = [g.batch_result(i)["epsinv"] for i in range(len(cases))]
            errs = [relf(a, b) for a, b in zip(got, eager)]
            print(f"{name} {prec} replay {rep}: {t * 1e3:.2f} ms, max rel err vs eager "
                  f"{max(errs):.2e}", flush=True)
            for (q, _, _), a, b in zip(cases, got, eager):
                assert a.shape == b.shape, q
            assert max(errs) <= 1e-12
            if first is None:
                first = got
            else:
                assert max(relf(a, b) for a, b in zip(got, first)) <= 1e-12
            # eager work on the same handle between replays
            e = g.run_q(q_before, prec)["epsinv"]
            assert relf(e, before) <= 1e-12
        g.batch_free()
    finally:
        g.close()


def test_autotune_picks_no_slower_launch():
    wl, w = _system("c1_plain_fp64")
    g = Chiqw(0)
    try:
        g.open(w, ecut_for_eps=wl["ecut_eps"], num_freq_grid=wl["nen"])
        ref = g.run_q(wl["q"], "bf16x3")["epsinv"]
        g.autotune(True)
        g.batch_plan([wl["q"], (0.0, 0.0, 0.0)], precision="bf16x3")
        table = g.tune_table()
        print(table, flush=True)
        assert len(table) == 2 and {t["ndir"] for t in table} == {1, 3}
        for t in table:
            assert t["candidates"] >= 2
            assert t["ms"] <= t["heuristic_ms"] * 1.0001
        g.batch_run()
        got = g.batch_result(0)["epsinv"]
        # another K split changes only the summation order of the contraction
        assert relf(got, ref) <= 1e-6
    finally:
        g.close()
