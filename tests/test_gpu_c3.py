"""The batched workload at full size on the GPU: every q of a 3^3 mesh (27 q,
Gamma included) of a 3^3-k system with 8 occupied and 200 empty bands
(~1000 plane waves, 15^3 grid), each q with its own chi cutoff of 17.5-24.5 Ry
(NG 434-705) and its own window of the lowest 150-200 empty bands, 32
frequencies, run as one captured HIP graph.

* The FP64 graph replay equals the eager FP64 path q by q (1e-12).
* The split-BF16 graph's eps^-1 is within 1e-5 relative Frobenius of the eager
  FP64 eps^-1 at every q (the mixed path's bar, held at the near-singular q
  too).
* Three q against the oracle with narrow empty-band windows (the same G
  spheres, k mesh and grid): FP64 within 1e-6.
"""
import numpy as np
import pytest

from respackdev_amd import gen
from respackdev_amd.chiqw import Chiqw, from_system

pytestmark = pytest.mark.gpu

WL = gen.WORKLOADS["c3_batch"]


def relf(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.fixture(scope="module")
def c3():
    s = gen.make_system(**{k: WL[k] for k in ("a", "nk", "nocc", "nvir", "ecut_psi", "seed")})
    w = from_system(s)
    g = Chiqw(0)
    info = g.open(w, ecut_for_eps=WL["ecut_eps"], num_freq_grid=WL["nen"])
    print(f"c3 opened: NTK {info['ntk']}, {info['n_occ']} occ + {info['n_vir']} empty", flush=True)
    yield g, w
    g.close()


def test_c3_graph_batches_against_eager_fp64(c3):
    g, _ = c3
    cases = gen.c3_windows(WL["nk"])
    qs, ecs, nvs = [c[0] for c in cases], [c[1] for c in cases], [c[2] for c in cases]
    g.autotune(True)
    try:
        g.batch_plan(qs, ecs, nvs, "fp64")
        g.batch_run()
        fp64_graph = [g.batch_result(i)["epsinv"] for i in range(len(qs))]
        g.batch_plan(qs, ecs, nvs, "bf16x3")
        g.batch_run()
        worst = 0.0
        for i, (q, ec, nv) in enumerate(cases):
            g.set_window(ec, nv)
            ref = g.run_q(q, "fp64")["epsinv"]
            e_graph = relf(fp64_graph[i], ref)
            mixed = g.batch_result(i)["epsinv"]
            e_mixed = relf(mixed, ref)
            worst = max(worst, e_mixed)
            print(f"c3 q {i} {tuple(round(x, 3) for x in q)} NG {ref.shape[-1]} nvir {nv}: "
                  f"fp64 graph {e_graph:.2e}, bf16x3 graph {e_mixed:.2e}", flush=True)
            assert e_graph <= 1e-12
            assert e_mixed <= 1e-5
        print(f"c3 worst split-BF16 eps^-1 error {worst:.3e}", flush=True)
    finally:
        g.set_window(0.0, 0)
        g.batch_free()
        g.autotune(False)


@pytest.mark.parametrize("iq,nvir", [(0, 2), (5, 2), (26, 1)])
def test_c3_q_matches_oracle(c3, iq, nvir):
    from oracle import pipeline
    g, w = c3
    q, ec, _ = gen.c3_windows(WL["nk"])[iq]
    g.set_window(ec, nvir)
    try:
        r = g.run_q(q, "fp64", want_chi=True)
    finally:
        g.set_window(0.0, 0)
    o = pipeline.run_q(w, q, ecut_for_eps=WL["ecut_eps"], num_freq_grid=WL["nen"],
                       ecut_eps_q=ec, n_vir_window=nvir)
    e_chi, e_eps = relf(r["chi"], o["chi"]), relf(r["epsinv"], o["epsinv"])
    print(f"c3 q {iq} NG {r['chi'].shape[-1]} nvir {nvir}: chi {e_chi:.3e} eps^-1 {e_eps:.3e}",
          flush=True)
    assert r["chi"].shape == o["chi"].shape
    assert e_chi <= 1e-6 and e_eps <= 1e-6
