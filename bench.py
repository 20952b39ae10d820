"""chi(q,w) seconds per q-point on one MI355X: split-BF16 mixed path vs plain FP64.

Workload (the precision configuration): simple-cubic a = 7.0 bohr, 6x6x6
Gamma-centred k mesh (216 k), 20 occupied + 180 empty bands (3600 pairs),
~1000 plane waves per k (reference FFT grid 15^3), chi cutoff 22.2 Ry
(~600 G vectors), 32 frequencies from the reference's gen_grid, q = (0,0,1/6).
Seeded synthetic wavefunctions (see respackdev_amd/gen.py).

A step is one q-point of the hot path with the wavefunctions resident in HBM:
pair densities, tetrahedron weights, the chi contraction and the Dyson step
(eps^-1 for all 32 frequencies), i.e. everything chiqw does per q.

Prints ONE JSON line (rank 0).  --gpus N > 1 is not supported (one GPU).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "chi(q,w) s per q-point: mixed/batched vs plain FP64 GPU at <=1e-5 rel err"
ERR_BAR = 1e-5  # chi and eps^-1 relative Frobenius error of the mixed path against FP64
PEAK = {"fp64": 78.6e12, "fp32": 157.3e12, "bf16x3": 2.5e15}  # dense MFMA, MI355X
KERNEL = {"fp64": "k_chi_gemm<0>", "fp32": "k_chi_gemm<1>", "bf16x3": "k_chi_bf16x3"}
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "hbm_traffic.txt")


def traffic_per_launch(workload: str, kernel: str):
    """HBM bytes per launch of `kernel` on `workload` from the committed rocprofv3
    FETCH_SIZE/WRITE_SIZE summary (lines 'traffic <workload> <kernel> <bytes>'),
    or None when it has not been measured."""
    try:
        with open(TRAFFIC_FILE) as f:
            for line in f:
                p = line.split()
                if len(p) == 4 and p[0] == "traffic" and p[1] == workload and p[2] == kernel:
                    return float(p[3])
    except OSError:
        pass
    return None


def copy_probe(mib: int, reps: int) -> int:
    """Elementwise copy of `mib` MiB (read mib, write mib) for calibrating the
    FETCH_SIZE / WRITE_SIZE counters."""
    import torch
    n = mib * (1 << 20) // 4
    src = torch.ones(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(reps):
        torch.mul(src, 1.0, out=dst)
    torch.cuda.synchronize()
    log(f"copy probe: {reps} x {mib} MiB read + {mib} MiB written")
    return 0


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def cpu_baseline(w, wl, npair_total: int, sample_pairs: int, threads: int) -> dict:
    """The oracle's FP64 OpenMP restatement (same loop order as chiqw.F:1319-1410)
    on `sample_pairs` band pairs spread evenly over the list, scaled to the full
    pair count."""
    import ctypes as C
    from oracle import lib, ptr, pipeline
    s = pipeline.setup(w, ecut_for_eps=wl["ecut_eps"], num_freq_grid=wl["nen"])
    L = lib()
    b = s["b"]
    sq = np.zeros(3)
    qin = np.ascontiguousarray(wl["q"], np.float64)
    L.ro_search_list(ptr(qin), ptr(sq))
    tmp = np.zeros((200000, 3), np.int32)
    ng = L.ro_make_kg0(200000, ptr(b[0]), ptr(b[1]), ptr(b[2]), s["ecut_eps"], sq[0], sq[1], sq[2], ptr(tmp))
    lg = np.ascontiguousarray(tmp[:ng])
    ntk, ntg = s["ntk"], s["ntg"]
    ikq = np.zeros(ntk, np.int32)
    shift = np.zeros((ntk, 3), np.int32)
    for k in range(ntk):
        sg = np.zeros(3, np.int32)
        ikq[k] = L.ro_search_kq(ntk, ptr(s["sk0"]), sq[0], sq[1], sq[2], k, ptr(sg))
        shift[k] = sg
    # pairs spread evenly over the list (every occupied band, empty bands across
    # the window), so that the sample meets the same mix of tetrahedron classes
    allp = [(j, s["n_docc"] + i) for j in range(s["n_occ"]) for i in range(s["n_vir"])]
    pick = np.unique(np.linspace(0, len(allp) - 1, min(sample_pairs, len(allp))).round().astype(int))
    pocc = np.array([allp[x][0] for x in pick], np.int32)
    pvir = np.array([allp[x][1] for x in pick], np.int32)
    nb = s["coef"].shape[0]
    coef = np.ascontiguousarray(s["coef"])
    pband = np.zeros((nb, ntk), np.complex128)
    nen = len(s["em"])
    chi = np.zeros((nen, ng, ng), np.complex128)
    nkb = [int(x) for x in s["nkb"]]
    t0 = time.perf_counter()
    L.ro_chiqw_q(ntk, ntg, ptr(s["ng0"]), ptr(s["kg0"]), nb, ptr(coef), nb, ptr(coef),
                 ptr(s["eband"]), ptr(s["eband"]), ptr(pband), ptr(pband), len(pocc), ptr(pocc),
                 ptr(pvir), ptr(ikq), ptr(shift), s["nw"][0], s["nw"][1], s["nw"][2], ng, ptr(lg),
                 nkb[0], nkb[1], nkb[2], ptr(s["imt"]), ptr(s["index_kpt"]), s["ef"], s["delt"],
                 s["dmna"], s["dmnr"], nen, ptr(s["em"]), s["volume"], ptr(chi), None, None,
                 threads)
    dt = time.perf_counter() - t0
    return {"value": dt * npair_total / len(pocc), "unit": "s/q-point", "cores": threads,
            "kind": "port",
            "sample": f"{len(pocc)} of {npair_total} band pairs (evenly spaced over the list) "
                      f"x {ntk} k x {nen} w at NG={ng} ({dt:.1f} s measured, scaled linearly "
                      f"in pairs)"}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="c2_precision")
    ap.add_argument("--mixed", default="bf16x3", choices=["bf16x3", "fp32"])
    ap.add_argument("--cpu-pairs", type=int, default=16)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--only", default="", help="run a single precision (profiling)")
    ap.add_argument("--autotune", action="store_true",
                    help="time the contraction's launch candidates in an untimed pass first")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run an elementwise copy of MIB MiB (counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py runs on one GPU only: this build has no multi-GPU path "
              "(q-points would shard across GPUs with no communication)", file=sys.stderr)
        return 2
    if args.copy_probe:
        return copy_probe(args.copy_probe, 3)

    from respackdev_amd import gen
    from respackdev_amd.chiqw import Chiqw, from_system

    if args.workload == "c4_calc_int":
        return calc_int_bench(args)
    if args.workload == "c3_batch":
        return c3_bench(args)
    if args.workload == "c5_closure":
        return closure_bench(args)
    wl = gen.WORKLOADS[args.workload]
    t0 = time.perf_counter()
    s = gen.make_system(**{k: wl[k] for k in ("a", "nk", "nocc", "nvir", "ecut_psi", "seed")})
    w = from_system(s)
    log(f"system generated in {time.perf_counter() - t0:.1f} s: {len(w.ski)} irreducible k, "
        f"{w.ntb} bands, NG_psi {w.ngi.min()}-{w.ngi.max()}")
    c = Chiqw(0)
    t0 = time.perf_counter()
    info = c.open(w, ecut_for_eps=wl["ecut_eps"], num_freq_grid=wl["nen"])
    log(f"opened in {time.perf_counter() - t0:.1f} s: NTK {info['ntk']} grid {info['nw']} "
        f"nen {info['nen']}")
    lg = c.prepare_q(wl["q"])
    ng, npair, ntk, nen = len(lg), c.npair, info["ntk"], info["nen"]
    K = npair * ntk
    flops = 8.0 * ng * ng * nen * K
    log(f"NG {ng}, pairs {npair}, K {K}, contraction {flops / 1e12:.1f} TFLOP per q")

    precs = [args.only] if args.only else ["fp64", args.mixed]
    results = {}
    if args.autotune:
        c.autotune(True)
    for prec in precs:
        if args.autotune:
            # the tuning pass (every candidate launch timed), outside the timed steps
            c.run_q(wl["q"], prec, want_chi=False)
            log(f"{prec} tuned: {c.tune_table()}")
        for i in range(args.warmup):
            r = c.run_q(wl["q"], prec, want_chi=False)
            log(f"{prec} warmup {i}: {r['timings']['device_total']:.3f} s")
        # timed steps: wavefunctions resident in HBM, eps^-1 left in HBM
        times, stages = [], []
        t_all = time.perf_counter()
        for i in range(args.steps):
            r = c.run_q(wl["q"], prec, want_eps=False)
            times.append(r["timings"]["device_total"])
            stages.append(r["timings"])
            log(f"{prec} step {i}: {r['timings']}")
        wall = (time.perf_counter() - t_all) / args.steps
        # one more step with eps^-1 (and chi) copied to the host over PCIe: the
        # rate a file-writing host sees, and the data for the error check
        t0 = time.perf_counter()
        r = c.run_q(wl["q"], prec, want_chi=True, want_eps=True)
        pcie = time.perf_counter() - t0
        results[prec] = dict(device=float(np.mean(times)), wall=wall, pcie=pcie, chi=r["chi"],
                             eps=r["epsinv"],
                             stages={k: float(np.mean([st[k] for st in stages])) for k in stages[0]})
    c.close()

    main_prec = precs[-1]
    res = results[main_prec]
    rel = lambda a, b: float(np.linalg.norm(a - b) / np.linalg.norm(b))
    out = {
        "metric": METRIC,
        "value": res["wall"],
        "unit": "s/q-point",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": res["wall"] * 1e3,
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": f"{main_prec} contraction (fp32 accumulation, fp64 merge); fp64 pair densities, "
                 f"tetrahedron weights and Dyson" if main_prec != "fp64" else "fp64",
        "data": "synthetic (seeded Gaussian plane-wave coefficients, Gram-Schmidt per k)",
        "config": {"workload": f"{args.workload}: a=7.0 bohr sc, {info['nkb'][0]}^3 k (NTK {ntk}), "
                               f"{info['n_occ']} occ + {info['n_vir']} empty bands, grid "
                               f"{info['nw'][0]}^3, NG {ng}, nen {nen}, q={tuple(wl['q'])}",
                   "ntk": ntk, "npair": npair, "ng": ng, "nen": nen, "K": K},
    }
    out["pcie_inclusive_s_per_q"] = res["pcie"]
    if "fp64" in results and main_prec != "fp64":
        out["fp64_s_per_q"] = results["fp64"]["wall"]
        out["speedup_vs_fp64"] = results["fp64"]["wall"] / res["wall"]
        out["chi_rel_err_vs_fp64"] = rel(res["chi"], results["fp64"]["chi"])
        out["epsinv_rel_err_vs_fp64"] = rel(res["eps"], results["fp64"]["eps"])
        out["within_error_bar"] = bool(max(out["chi_rel_err_vs_fp64"],
                                           out["epsinv_rel_err_vs_fp64"]) <= ERR_BAR)
        if not out["within_error_bar"]:
            log(f"WARNING: mixed-precision error above the {ERR_BAR:g} bar")
    out["stage_s"] = {p: results[p]["stages"] for p in results}
    # dominant kernel: the main contraction launch, timed with HIP events on the
    # library's stream around each launch (one launch per q away from Gamma)
    st = res["stages"]
    t_launch = st["gemm_kernel"] / max(1, st["gemm_launches"])
    achieved = flops / t_launch
    rf = {"kernel": KERNEL[main_prec], "bound": "mfma", "achieved": achieved / 1e12,
          "peak": PEAK[main_prec] / 1e12, "unit": "TFLOP/s", "frac": achieved / PEAK[main_prec],
          "traffic": traffic_per_launch(args.workload, KERNEL[main_prec]),
          "algorithmic_flops_per_launch": flops, "avg_launch_ms": t_launch * 1e3}
    if main_prec == "bf16x3":
        # three bf16 products per real product: executed MFMA work against the same peak
        rf["executed_frac"] = 3.0 * achieved / PEAK[main_prec]
    out["roofline"] = rf
    # the other stages of the q (device times from events around each stage):
    # pair densities against HBM (per (pair, k): the two complex64 band grids
    # of N_r points read, M(G) written), the FP64 Dyson step against the FP64
    # peak (8 NG^3 per frequency), tetrahedron evaluations per second
    nr = int(np.prod(info["nw"]))
    pd_bytes = 8.0 * (2 * nr + ng) * K
    dy_flops = 8.0 * float(ng) ** 3 * nen
    n_eval = npair * 6 * ntk * 2 * nen
    out["rooflines"] = {
        "k_pair_densities_pruned": {"bound": "hbm", "achieved": pd_bytes / st["pairdens"] / 1e9,
                                    "peak": 8000.0, "unit": "GB/s",
                                    "frac": pd_bytes / st["pairdens"] / 8.0e12,
                                    "algorithmic_bytes_per_launch": pd_bytes,
                                    "avg_launch_ms": st["pairdens"] * 1e3,
                                    "traffic": traffic_per_launch(args.workload,
                                                                  "k_pair_densities_pruned")},
        "dyson": {"bound": "fp64", "achieved": dy_flops / st["dyson"] / 1e12, "peak": 78.6,
                  "unit": "TFLOP/s", "frac": dy_flops / st["dyson"] / 78.6e12,
                  "algorithmic_flops": dy_flops, "stage_ms": st["dyson"] * 1e3},
        "tetrahedron": {"evaluations_per_s": n_eval / st["weights"], "stage_ms": st["weights"] * 1e3,
                        "unit": "(pair, tetrahedron, dsgn, w) per s"},
    }
    if "fp64" in results:
        s64 = results["fp64"]["stages"]
        a64 = flops / (s64["gemm_kernel"] / max(1, s64["gemm_launches"]))
        out["roofline_fp64"] = {"kernel": KERNEL["fp64"], "achieved": a64 / 1e12,
                                "peak": PEAK["fp64"] / 1e12, "unit": "TFLOP/s",
                                "frac": a64 / PEAK["fp64"]}
    if not args.no_cpu:
        threads = min(os.cpu_count() or 1, 16)
        log(f"cpu baseline: oracle on {args.cpu_pairs} pairs with {threads} threads")
        out["cpu_baseline"] = cpu_baseline(w, wl, npair, args.cpu_pairs, threads)
    print(json.dumps(out), flush=True)
    return 0


def c3_bench(args) -> int:
    """The batched configuration (--workload c3_batch; not the default line):
    every q of a 3^3 mesh on one opened system (3^3 k, 8 occupied + 200 empty
    bands, ~1000 plane waves) with per-q chi cutoffs 17.5-24.5 Ry (NG ~420-700)
    and per-q windows of 150-200 empty bands, with the wavefunctions resident.
    The q list is planned once per precision (the contraction autotuned per
    shape) and captured into one HIP graph; a step is one replay of the whole
    list.  value is seconds per q of the mixed path's replay, with the FP64
    replay, the eager q-by-q loop and the per-q errors beside it."""
    from respackdev_amd import gen
    from respackdev_amd.chiqw import Chiqw, from_system

    wl = gen.WORKLOADS["c3_batch"]
    s = gen.make_system(**{k: wl[k] for k in ("a", "nk", "nocc", "nvir", "ecut_psi", "seed")})
    w = from_system(s)
    cases = gen.c3_windows(wl["nk"])
    c = Chiqw(0)
    try:
        info = c.open(w, ecut_for_eps=wl["ecut_eps"], num_freq_grid=wl["nen"])
        log(f"c3: NTK {info['ntk']}, {len(cases)} q, {info['n_occ']} occ + {info['n_vir']} empty")
        res = {}
        qs = [cs[0] for cs in cases]
        ecs = [cs[1] for cs in cases]
        nvs = [cs[2] for cs in cases]
        c.autotune(True)
        graph = {}
        for prec in ("fp64", args.mixed):
            t0 = time.perf_counter()
            c.batch_plan(qs, ecs, nvs, prec)
            plan_s = time.perf_counter() - t0
            for i in range(args.warmup):
                c.batch_run()
            walls, dev = [], []
            for step in range(args.steps):
                t0 = time.perf_counter()
                dev.append(c.batch_run())
                walls.append(time.perf_counter() - t0)
                log(f"{prec} graph step {step}: {walls[-1]:.3f} s wall, {dev[-1]:.3f} s device "
                    f"for {len(cases)} q")
            graph[prec] = dict(wall=float(np.mean(walls)), device=float(np.mean(dev)), plan=plan_s)
            c.batch_free()
        tune = c.tune_table()
        for prec in ("fp64", args.mixed):
            for i in range(args.warmup):
                c.set_window(cases[0][1], cases[0][2])
                c.run_q(cases[0][0], prec, want_eps=False)
            walls, dev, ngs, flops = [], [], [], 0.0
            stage_sum = {}
            for step in range(args.steps):
                t0 = time.perf_counter()
                devs = 0.0
                for q, ecut, nvir in cases:
                    c.set_window(ecut, nvir)
                    r = c.run_q(q, prec, want_eps=False)
                    devs += r["timings"]["device_total"]
                    for k_ in ("pairdens", "weights", "contraction", "dyson"):
                        stage_sum[k_] = stage_sum.get(k_, 0.0) + r["timings"][k_] / args.steps
                    if step == 0:
                        ng = int(len(r["lg"]))
                        ngs.append(ng)
                        flops += 8.0 * ng * ng * info["nen"] * c.npair * info["ntk"] * (3 if q == (0.0, 0.0, 0.0) else 1)
                walls.append(time.perf_counter() - t0)
                dev.append(devs)
                log(f"{prec} step {step}: {walls[-1]:.3f} s wall, {devs:.3f} s device for {len(cases)} q")
            res[prec] = dict(wall=float(np.mean(walls)), device=float(np.mean(dev)), ngs=ngs,
                             flops=flops, stages=stage_sum)
        # errors of the mixed path against FP64, q by q (one pair of results held at a time)
        rel = lambda a, b: float(np.linalg.norm(a - b) / np.linalg.norm(b))
        m = args.mixed
        err_chi = err_eps = 0.0
        for q, ecut, nvir in cases:
            c.set_window(ecut, nvir)
            r64 = c.run_q(q, "fp64", want_chi=True, want_eps=True)
            rmx = c.run_q(q, m, want_chi=True, want_eps=True)
            err_chi = max(err_chi, rel(rmx["chi"], r64["chi"]))
            err_eps = max(err_eps, rel(rmx["epsinv"], r64["epsinv"]))
    finally:
        c.close()
    nq = len(cases)
    cpu = None
    if not args.no_cpu:
        # the oracle on a sample of the pairs of q 1 (its cutoff and window)
        q1, ec1, nv1 = cases[1]
        threads = min(os.cpu_count() or 1, 16)
        cpu = cpu_baseline(w, dict(wl, q=q1, ecut_eps=ec1), info["n_occ"] * nv1, args.cpu_pairs,
                           threads)
        cpu["sample"] += f"; q {tuple(round(x, 4) for x in q1)}, {nv1} empty bands"
    out = {"metric": "chi(q,w) s per q-point over a batch of uneven q (3^3 mesh)",
           "value": graph[m]["wall"] / nq, "unit": "s/q-point", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "higher_is_better": False, "dtype": m,
           "fp64_s_per_q": graph["fp64"]["wall"] / nq,
           "speedup_vs_fp64": graph["fp64"]["wall"] / graph[m]["wall"],
           "device_s_per_q": graph[m]["device"] / nq,
           "graph": {"plan_s": {p: graph[p]["plan"] for p in graph},
                     "eager_s_per_q": {p: res[p]["wall"] / nq for p in res},
                     "eager_device_s_per_q": {p: res[p]["device"] / nq for p in res}},
           "autotune": {p: {"shapes": len(ts),
                            "changed": sum(t["nsplit"] != t["heuristic_nsplit"] for t in ts),
                            "contraction_ms_tuned": sum(t["ms"] for t in ts),
                            "contraction_ms_heuristic": sum(t["heuristic_ms"] for t in ts)}
                        for p, ts in (("fp64", [t for t in tune if t["precision"] == 0]),
                                      (m, [t for t in tune if t["precision"] != 0]))},
           "chi_rel_err_vs_fp64_max": err_chi, "epsinv_rel_err_vs_fp64_max": err_eps,
           "within_error_bar": bool(max(err_chi, err_eps) <= ERR_BAR),
           "eager_stage_s_per_q": {p: {k_: v_ / nq for k_, v_ in res[p]["stages"].items()}
                                   for p in res},
           "cpu_baseline": cpu,
           "contraction_tflop_total": res[m]["flops"] / 1e12,
           "data": "synthetic (seeded Gaussian plane-wave coefficients, Gram-Schmidt per k)",
           "config": {"workload": "c3_batch: a=7.0 bohr sc, 3^3 k (NTK 27), 8 occ + 200 empty, "
                                  "every q of the mesh, per-q chi cutoff 17.5-24.5 Ry and "
                                  "150-200 empty bands", "nq": nq,
                      "ng_range": [min(res[m]["ngs"]), max(res[m]["ngs"])]}}
    print(json.dumps(out), flush=True)
    return 0


def closure_bench(args) -> int:
    """The learned band-tail closure (--workload c5_closure; not the default
    line) at the closure configuration's scale: trains the network on chi of 16
    members of the precision configuration's family (lattice 6.5-8.0 bohr, gap
    1-4 eV; 6^3 k, 20 occupied bands, ~1000 plane waves, chi cutoff 22.2 Ry so
    NG ~ 480-900, 32 frequencies; 50 vs 400 empty bands at 3 irreducible q drawn
    per system) with the library's own kernels, then reports on 4 held-out
    systems the static W_ij error against the 400-band result (eps^-1 of all 112
    irreducible q at w = 0, calc_w3d with 5 model Wannier functions) for plain
    truncation at 50, 100 and 200 empty bands and for 50 bands plus the closure."""
    from respackdev_amd import closure
    t0 = time.perf_counter()
    r = closure.run(n_train=16, n_test=4, nk=6, nocc=20, n_trunc=50, n_full=400,
                    plain=(50, 100, 200), ecut_psi=30.97, ecut_eps=22.2, nen=32, steps=6000,
                    q_per_system=3, test_nen=1, log=log)
    out = {"metric": "static W_ij error (meV, max over i, j, R at w = 0; mean over held-out "
                     "systems) vs the 400-empty-band result",
           "value": r["w_err_mev"].get("closure"), "unit": "meV", "n_gpus": 1,
           "higher_is_better": False, "wall_s": time.perf_counter() - t0, **r,
           "data": "synthetic family (seeded)",
           "config": {"workload": "c5_closure: 16 train + 4 held-out systems, 6^3 k, 20 occ, "
                                  "50 vs 400 empty bands, ~1000 PW, chi cutoff 22.2 Ry, 32 w "
                                  "(3 q per training system; held-out W at w = 0 over all 112 q); "
                                  "MLP 8-128-128-128-2 GELU"}}
    print(json.dumps(out), flush=True)
    return 0


def calc_int_bench(args) -> int:
    """calc_w3d and calc_j3d at the calc_int size (--workload c4_calc_int; not
    the default line): 6^3 mesh (NTK = NTQ = 216, 112 irreducible q), 5 d-like
    Wannier functions, chi cutoff 22.2 Ry (NG ~ 600), 32 frequencies.  eps^-1
    of every irreducible q comes from this build's chi on the same system
    (split-BF16 contraction, FP64 Dyson; at Gamma the mean of the three
    directions, as calc_int's reader forms it), resident in HBM as complex64.
    A step is one rp_w3d_funcs + rp_w3d_vw (V and W over all q), and one
    rp_j3d_run for R = 0 (X and J of one lattice vector).

    Also reported: device time per stage (events on the library's stream), the
    HBM rooflines of the W and J contractions (bytes per w: 8 NG_eps^2 per
    eps^-1 read, once per group of q sharing an irreducible eps^-1 for W and
    once per q for J, + 16 NG_psi NV per q, NV the vectors: NWF for W, NWF^2
    pair vectors for J), V
    and W against the oracle over every q in meV (calc_intW.F:942-985),
    X_ii(0) - V_ii(0) and J_ii(0, w) - W_ii(0, w) in meV, and the oracle's
    calc_w3d timed on the host as the CPU baseline."""
    from respackdev_amd import gen
    from respackdev_amd.chiqw import Chiqw, from_system
    from respackdev_amd.w3d import AU, W3d

    wl = gen.WORKLOADS["c4_calc_int"]
    t0 = time.perf_counter()
    s = gen.make_system(**{k: wl[k] for k in ("a", "nk", "nocc", "nvir", "ecut_psi", "seed")})
    w = from_system(s)
    wan, _ = gen.make_wannier(s, wl["nwf"], wl["seed"])
    ntg = max(c.shape[1] for c in wan)
    W = np.zeros((len(wan), wl["nwf"], ntg), np.complex128)
    for k, c in enumerate(wan):
        W[k, :, :c.shape[1]] = c
    sqi = np.array([[gen._fold(x) for x in k] for k in w.ski])
    # eps^-1 of every irreducible q from chi (untimed)
    c = Chiqw(0)
    try:
        cinfo = c.open(w, ecut_for_eps=wl["ecut_eps"], num_freq_grid=wl["nen"])
        em = np.asarray(cinfo["em"], np.complex128)
        eps = []
        for q in sqi:
            e = c.run_q(q, "bf16x3")["epsinv"]
            eps.append((e.mean(axis=0) if e.shape[0] > 1 else e[0]).astype(np.complex64))
    finally:
        c.close()
    log(f"eps^-1 of {len(sqi)} irreducible q from chi in {time.perf_counter() - t0:.1f} s")
    g = W3d(0)
    try:
        info = g.open(w, wl["ecut_eps"], em)
        g.set_wannier(W)
        ng = g.set_q(sqi)
        assert [e.shape[-1] for e in eps] == list(ng), "chi sphere differs from calc_int's LG0"
        for iq, e in enumerate(eps):
            g.set_epsinv(iq, e)
        info = g.info()
        log(f"calc_int set up: NTK {info['ntk']}, NTQ {info['ntq']}, {len(ng)} irreducible q, "
            f"NG {ng.min()}-{ng.max()}, NWF {info['nwf']}")
        for i in range(args.warmup):
            g.funcs(); g.vw(); g.j3d([(0, 0, 0)])
        tf, tv, tj, st = [], [], [], []
        for i in range(args.steps):
            t0 = time.perf_counter(); g.funcs(); tf.append(time.perf_counter() - t0)
            sw = g.stage_times()
            t0 = time.perf_counter(); v, wm, _, _ = g.vw(); tv.append(time.perf_counter() - t0)
            t0 = time.perf_counter(); x, j, _, _ = g.j3d([(0, 0, 0)]); tj.append(time.perf_counter() - t0)
            sj = g.stage_times()
            st.append({**{k: sw[k] for k in sw if k.startswith("w_")},
                       **{k: sj[k] for k in sj if k.startswith("j_")}})
            log(f"step {i}: funcs {tf[-1]:.4f} s, vw {tv[-1]:.4f} s, j3d(R=0) {tj[-1]:.4f} s; "
                f"stages {st[-1]}")
    finally:
        g.close()
    stage = {k: float(np.mean([d[k] for d in st])) for k in st[0]}
    nwf, nen, ntq = info["nwf"], info["nen"], info["ntq"]
    out = {"metric": "calc_w3d s per run (V and W over all q) and calc_j3d s per lattice vector",
           "value": float(np.mean(tf) + np.mean(tv)), "unit": "s", "n_gpus": 1,
           "steps": args.steps, "warmup": args.warmup, "higher_is_better": False,
           "w3d_funcs_s": float(np.mean(tf)), "w3d_vw_s": float(np.mean(tv)),
           "j3d_per_r_s": float(np.mean(tj)), "stage_s": stage,
           "dtype": "fp64 (eps^-1 complex64 as the reference)",
           "data": "synthetic Wannier functions; eps^-1 from this build's chi of the same system",
           "config": {"workload": "c4_calc_int: a=7.0 bohr sc, 6^3 k (NTK = NTQ 216), 5 Wannier "
                                  "functions, NG_eps ~600, nen 32",
                      "ntk": info["ntk"], "ntq": ntq, "nwf": nwf, "nen": nen,
                      "nq_irr": int(len(ng)), "ng_eps": [int(ng.min()), int(ng.max())]}}
    # HBM rooflines of the two contractions: per frequency, eps^-1 of every
    # full-mesh q (J) or of every group of up to 16 // NWF q sharing one
    # irreducible eps^-1 (W), read once, plus the vectors
    from oracle import pipeline_w3d
    cset = pipeline_w3d.setup_w3d(w, W, sqi, lambda iq: eps[iq], em, wl["ecut_eps"])
    nge = np.array([cset["nge"][iq] for iq in range(ntq)], np.float64)
    ngp = np.array([cset["ngp"][iq] for iq in range(ntq)], np.float64)
    star = np.bincount(np.asarray(cset["numirrq"][:ntq]), minlength=len(ng))
    ngroups = np.ceil(star / (16 // nwf))
    e_w = float(np.sum(ngroups * 8.0 * np.asarray(ng, np.float64) ** 2))
    for name, nv, key, e_bytes in (("k_w_contract", nwf, "w_contract", e_w),
                                   ("k_j_contract", nwf * nwf, "j_contract", float(np.sum(8.0 * nge ** 2)))):
        byt = float(nen * (e_bytes + np.sum(16.0 * ngp * nv)))
        out.setdefault("rooflines", {})[name] = {
            "bound": "hbm", "achieved": byt / stage[key] / 1e9, "peak": 8000.0, "unit": "GB/s",
            "frac": byt / stage[key] / 8.0e12, "algorithmic_bytes_per_launch": byt,
            "avg_launch_ms": stage[key] * 1e3,
            "traffic": traffic_per_launch("c4_calc_int", name)}
    # parity: V and W over every q against the oracle (its calc_w3d on the
    # host is also the CPU baseline), and the exchange identities at R = 0
    if not args.no_cpu:
        threads = min(os.cpu_count() or 1, 16)
        log(f"oracle calc_w3d over {ntq} q on the host ({threads} threads)")
        t0 = time.perf_counter()
        o = pipeline_w3d.run_w3d(w, W, sqi, lambda iq: eps[iq], em, wl["ecut_eps"])
        tcpu = time.perf_counter() - t0
        out["v_err_meV"] = float(np.abs(v - o["vmat"]).max() * AU * 1e3)
        out["w_err_meV"] = float(np.abs(wm - o["wmat"]).max() * AU * 1e3)
        out["v_rel_err"] = float(np.linalg.norm(v - o["vmat"]) / np.linalg.norm(o["vmat"]))
        out["w_rel_err"] = float(np.linalg.norm(wm - o["wmat"]) / np.linalg.norm(o["wmat"]))
        out["cpu_baseline"] = {"value": tcpu, "unit": "s per calc_w3d run", "cores": threads,
                               "kind": "port", "sample": f"the oracle's calc_w3d over all {ntq} q "
                               "(pair densities in threads over k, funcw in threads over w)"}
    na = (v.shape[0] - 1) // 2
    out["x_minus_v_diag_meV"] = float(np.abs(np.diagonal(x[0]) -
                                             np.diagonal(v[na, na, na])).max() * AU * 1e3)
    out["j_minus_w_diag_meV"] = float(np.abs(np.diagonal(j[:, 0], axis1=-2, axis2=-1) -
                                             np.diagonal(wm[:, na, na, na], axis1=-2,
                                                         axis2=-1)).max() * AU * 1e3)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
