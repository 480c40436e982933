#!/usr/bin/env python3
"""Benchmark: integrand evaluations per second of the plain GPU method on the
8-D workloads (oscillatory + product peak on [0,1]^8, FP64, Genz-Malik
degree-7/5, 401 points per region, epsrel 1e-6, maxeval 1e9).

One step = one cuhre_integrate call per 8-D integrand, from the initial grid
to the stop (global test met, or the evaluation budget reached).  The value
is whole-step evals/s = sum(neval) / wall time.  Inputs (the integrand
parameters) are a few hundred bytes; the region list is created on the GPU.

Also reported: the roofline of the dominant kernel (one-region-per-lane rule
evaluation of the product peak, FP64-bound), the CPU baseline (the oracle
port, OpenMP over the host cores) on the same workload, and the time to
epsrel 1e-6 on the 8-D product peak (maxeval 1e12) of the plain method
against the learned split guide, trained in the same run on product peaks
with other centres (learned_vs_plain_8d), and 4096 independent 5-D
integrals in one batched call against 4096 plain calls (batched).

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "old-gpuintegration_amd"))

import workloads as W  # noqa: E402

FP64_PEAK_TFLOPS = 78.6
# Executed FP64 flops per region of the rule evaluation, from the
# SQ_INSTS_VALU_{ADD,MUL,FMA,TRANS}_F64 counters (add/mul/transcendental 1,
# fma 2) over every evaluation launch of a run of the workload
# (profiles/eval_product_peak8d.md).  These are the roofline's work
# count: the evaluation shares an integrand's per-coordinate partial
# results between points with equal coordinates, so it executes fewer flops
# than a point-by-point evaluation (region_flops, the model, reported beside
# it as model_flops_per_region).
EXEC_FLOPS_PER_REGION_PP8 = 8240  # 8-D product peak, bench workload (model 24,426)
# Issue ceiling of that evaluation: per region it issues 5,158 full-rate
# VALU instructions and 402 v_rcp_f64 (one reciprocal per point), and the
# reciprocal issues at about a quarter of the FMA rate (flopcost's issue-rate
# kernels: 4.24 against 1.24 CU-cycles per wave instruction,
# profiles/flop_costs_gfx950.md), so at most 8,240 flops per
# 2 * (5,158 + 4 * 402) issue slots, 0.61 of the FP64 peak, is reachable.
ISSUE_CEILING_PP8 = 8240 / (2 * (5158 + 4 * 402))
EXEC_FLOPS_PER_REGION_T6 = 4754   # table-driven 6-D integrand, epsrel 1e-5 (model 16,322)
# the batched evaluation (k_batch_eval, the three 5-D families of the
# batch workload) on the autotuner's sample: executed FP64 flops per region,
# same counters (profiles/batch_c4_5d.md)
EXEC_FLOPS_PER_REGION_BATCH5 = 2988
# measured HBM traffic per rule-evaluation launch of the 8-D product peak on
# this workload (rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE averaged over the
# 13 launches of a run that evaluate regions, FETCH_SIZE scaled x2 and
# WRITE_SIZE x1 as a 32 MiB copy kernel reads)
TRAFFIC_PER_LAUNCH = 2 * 10142.1 * 1024 + 4465.5 * 1024


def flop_costs():
    path = os.path.join(ROOT, "old-gpuintegration_amd", "flop_costs.json")
    with open(path) as fh:
        return json.load(fh)["fp64_flops"]


def ipow_flops(e, d):
    """Multiplies of b^e (unrolled for Genz's exponent d + 1, else the loop)."""
    if e <= 0:
        return 0
    sq = e.bit_length() - 1
    bits = bin(e).count("1")
    return sq + bits - 1 if e == d + 1 else sq + bits


def family_flops(f, d, c):
    """FP64 flops of one integrand evaluation (add/sub/mul 1, fma 2,
    device-math calls charged their counted instruction flops)."""
    fam = f.family
    if fam == W.GAUSSIAN:
        return 4 * d + c["exp"]
    if fam == W.PRODUCT_PEAK:
        return 4 * d - 1 + c["recip"]
    if fam == W.OSCILLATORY:
        return 2 * d + 1 + c["cos"]
    if fam == W.CORNER_PEAK:
        return 2 * d + ipow_flops(f.exponent, d) + c["recip"]
    if fam == W.TABLE6:
        return 20 + c["exp"] + c["sqrt"] + 2 * c["div"]
    raise ValueError(fam)


def region_flops(f, d, c):
    """Per region: integrand + d coordinate FMAs at every point, three sum
    FMAs at centre/axis/pair points and two at corners, plus setup."""
    p = W.gm_npoints(d)
    corners = 1 << d
    return p * (family_flops(f, d, c) + 2 * d) + 6 * (p - corners) + 4 * corners + (8 * d + 12)


def check_against_exact(problems, per, conv):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    out = {}
    for p, item in zip(problems, per):
        ex, _ = oracle.exact(p.f, p.lows, p.highs)
        out[p.name] = {"exact": ex, "rel_err": abs(item["integral"] - ex) / abs(ex),
                       "reported_rel_error": item["error"] / abs(item["integral"])}
    if conv:
        ex, _ = oracle.exact(problems[1].f, problems[1].lows, problems[1].highs)
        out["product_peak8d_to_convergence"] = {"exact": ex, "rel_err": abs(conv["integral"] - ex) / abs(ex),
                                                "reported_rel_error": conv["error"] / abs(conv["integral"])}
    return out


def host_info():
    """CPU model, logical CPUs of the machine and of this process."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        allowed = len(os.sched_getaffinity(0))
    except AttributeError:
        allowed = os.cpu_count() or 1
    return {"cpu_model": model, "nproc": os.cpu_count(), "cpus_allowed": allowed}


def _timed(fn):
    t0 = time.perf_counter()
    r = fn()
    return r, time.perf_counter() - t0


def cpu_baseline(problems):
    """The oracle on the host cores, on the full benchmark workload (both
    8-D integrands, maxeval 1e9): its breadth-first loop (the GPU's
    algorithm, C + OpenMP over the regions of an iteration) on the
    process's share of the host (at most 16 threads, the CPU share of one
    GPU on these machines) and single-threaded, and its Cuhre-style
    global-adaptive heap (largest error first, one region at a time,
    single-threaded by nature).  value = the multi-threaded breadth-first
    figure."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    info = host_info()
    threads = max(1, min(16, info["cpus_allowed"]))
    flavours = {}
    for name, fn in (("bfs_threads", lambda p: oracle.integrate_bfs(p, nthreads=threads)),
                     ("bfs_1thread", lambda p: oracle.integrate_bfs(p, nthreads=1)),
                     ("heap_1thread", lambda p: oracle.integrate_heap(p))):
        ne, dt, per = 0, 0.0, []
        for p in problems:
            r, t = _timed(lambda: fn(p))
            ne += r["neval"]
            dt += t
            per.append({"integrand": p.name, "rc": r["rc"], "integral": r["integral"], "error": r["error"],
                        "nregions": r["nregions"], "seconds": t})
        flavours[name] = {"evals_per_s": ne / dt, "neval": ne, "seconds": dt, "per_integrand": per}
    main = flavours["bfs_threads"]
    return {"value": main["evals_per_s"], "unit": "evals/s", "cores": threads, "kind": "port",
            "sample": f"full workload: the oracle's breadth-first GM7/5 loop on both 8-D integrands "
                      f"(maxeval 1e9), {main['neval']} evals in {main['seconds']:.2f} s on {threads} "
                      f"threads (OpenMP over regions); single-threaded and heap flavours below",
            "host": info, "threads": threads,
            "bfs_1thread_evals_per_s": flavours["bfs_1thread"]["evals_per_s"],
            "heap_1thread_evals_per_s": flavours["heap_1thread"]["evals_per_s"],
            "flavours": flavours}


def bench_batched(ctx, count=4096):
    """4096 independent 5-D integrals (Gaussian / product peak / corner peak,
    epsrel 1e-5 each): one batched call vs one plain call per integral, with
    the plain per-region test (results equal the plain calls') and with the
    shared-allotment finish policy (finish_budget 0.25)."""
    probs = W.c4_batch(count=count)
    cfg, table = ctx.batch_autotune(probs)
    ctx.integrate_batch(probs)  # warm-up (graph capture)
    t0 = time.perf_counter()
    res, info, rc = ctx.integrate_batch(probs)
    tb = time.perf_counter() - t0
    ctx.integrate_batch(probs, finish_budget=0.25)
    ta = []
    for _ in range(3):
        t0 = time.perf_counter()
        ares, ainfo, arc = ctx.integrate_batch(probs, finish_budget=0.25)
        ta.append(time.perf_counter() - t0)
    ta = min(ta)
    # the persistent form: one launch runs the instances' loops in
    # per-workgroup region queues, the largest are handed over
    pers = {}
    for budget in (0.0, 0.25):
        ctx.integrate_batch(probs, finish_budget=budget, persistent=True)
        best = None
        for _ in range(3):
            t0 = time.perf_counter()
            pres, pinfo, prc = ctx.integrate_batch(probs, finish_budget=budget, persistent=True)
            dt = time.perf_counter() - t0
            if best is None or dt < best[0]:
                best = (dt, pres, pinfo, prc)
        dt, pres, pinfo, prc = best
        ref = res if budget == 0.0 else ares
        pers["plain_policy" if budget == 0.0 else "allotment"] = {
            "batched_s": dt, "rc": prc, "regions": pinfo["total_regions"],
            "evals_per_s": pinfo["total_neval"] / dt, "not_converged": pinfo["not_converged"],
            "finished_in_persistent_launch": pinfo["persistent_finished"],
            "same_as_replayed": sum(1 for a, b in zip(pres, ref) if a["nregions"] == b["nregions"]
                                    and a["iterations"] == b["iterations"]
                                    and abs(a["integral"] - b["integral"]) <= 1e-12 * abs(b["integral"]))}
    t0 = time.perf_counter()
    plain = [ctx.integrate(p) for p in probs]
    tp = time.perf_counter() - t0
    for v in pers.values():
        v["wall_vs_plain_calls"] = v["batched_s"] / tp
    same = sum(1 for a, b in zip(res, plain) if a["nregions"] == b["nregions"]
               and abs(a["integral"] - b["integral"]) <= 1e-12 * abs(b["integral"]))
    regions = sum(r["nregions"] for r in res)
    # roofline of the batched evaluation: the autotuner times one evaluation
    # of its sample (every instance's initial grid refined uniformly until
    # the list holds 2^19 regions or more) per launch shape with HIP events
    d = probs[0].f.dim
    k = 4 if d <= 4 else 2
    n_sample = count * k ** d
    while n_sample < (1 << 19):
        n_sample *= 2
    best_ms = float(min(table))
    eval_tf = EXEC_FLOPS_PER_REGION_BATCH5 * n_sample / (best_ms * 1e-3) / 1e12
    return {"instances": count, "batched_s": tb, "plain_sequential_s": tp, "wall_ratio": tb / tp,
            "eval_roofline": {"bound": "fp64", "achieved_tflops": eval_tf, "frac": eval_tf / FP64_PEAK_TFLOPS,
                              "flops_per_region": EXEC_FLOPS_PER_REGION_BATCH5, "sample_regions": n_sample,
                              "ms_per_evaluation": best_ms,
                              "note": "k_batch_eval over the autotuner's sample at its chosen shape (HIP "
                                      "events); executed FP64 flops per region from the VALU counters"},
            "batched_evals_per_s": info["total_neval"] / tb,
            "plain_evals_per_s": sum(r["neval"] for r in plain) / tp,
            "regions": regions, "regions_plain": sum(r["nregions"] for r in plain),
            "identical_instances": same, "not_converged": info["not_converged"],
            "iterations": info["iterations"], "graph_used": info["graph_used"], "autotuned": cfg,
            "autotune_ms": table.tolist(),
            "per_instance_regions_min_max": [min(r["nregions"] for r in res), max(r["nregions"] for r in res)],
            "allotment": {"finish_budget": 0.25, "batched_s": ta, "rc": arc,
                          "regions": ainfo["total_regions"], "evals_per_s": ainfo["total_neval"] / ta,
                          "regions_vs_plain_calls": ainfo["total_regions"] / sum(r["nregions"] for r in plain),
                          "wall_vs_plain_calls": ta / tp, "not_converged": ainfo["not_converged"],
                          "iterations": ainfo["iterations"],
                          "per_instance_regions_min_max": [min(r["nregions"] for r in ares),
                                                           max(r["nregions"] for r in ares)]},
            "persistent": pers,
            "_results": {"plain_policy": res, "allotment": ares}}


def _learned_family(ctx, dim, train, held, maxeval, train_maxeval, samples_per_run, policies, rounds=3):
    """Collect on the training problems (round 0 under the plain axes, later
    rounds under the network's own axes, every sample labelled by the
    look-ahead), train on the GPU after each round, then run each policy on
    the held-out problems (best of two runs)."""
    import cuhre
    g = cuhre.Guide(ctx, seed=11)
    t0 = time.perf_counter()
    log = []
    for rnd in range(rounds):
        for p in train:
            g.collect(p, max_samples=samples_per_run, maxeval=train_maxeval, axis_mode=0 if rnd == 0 else 1,
                      axis_confidence=0.6)
        for ep, lr in ((2, 2e-3), (2, 1e-3), (1, 3e-4)):
            st = g.train(dim, epochs=ep, batch=8192, lr=lr, seed=10 * rnd + ep)
        log.append(dict(st, round=rnd))
    t_learn = time.perf_counter() - t0
    rows = []
    for p in held:
        row = {"integrand": p.name}
        for name, pol in policies.items():
            best = None
            for _ in range(2):
                r = ctx.integrate(p, maxeval=maxeval) if pol is None else g.integrate(p, maxeval=maxeval, **pol)
                if best is None or r["ms_total"] < best["ms_total"]:
                    best = r
            row[name] = {"rc": best["rc"], "nregions": best["nregions"], "ms": best["ms_total"],
                         "integral": best["integral"], "error": best["error"],
                         "fp64_err_bound": best["fp64_err_bound"],
                         "retired": best.get("guide", {}).get("retired", 0),
                         "overridden": best.get("guide", {}).get("overridden", 0)}
        for name in policies:
            if name != "plain":
                row[name]["regions_vs_plain"] = row[name]["nregions"] / row["plain"]["nregions"]
                row[name]["time_vs_plain"] = row[name]["ms"] / row["plain"]["ms"]
        rows.append(row)
    n = g.samples()
    g.close()
    return {"samples": n, "collect_and_train_s": t_learn, "train": log, "held_out": rows, "maxeval": maxeval}


PEAK8_POLICIES = {
    "plain": None,
    "shared_allotment": dict(axis_mode=0, finish_mode=3, budget=0.5),
    "learned_axis": dict(axis_mode=1, finish_mode=0, axis_confidence=0.6),
    "learned_axis_and_allotment": dict(axis_mode=1, finish_mode=3, budget=0.5, axis_confidence=0.6)}


def bench_learned_peak8(ctx, policies=PEAK8_POLICIES):
    """8-D product peaks, epsrel 1e-6: the guide trained on centres 200..211
    (maxeval 1e10, 400k samples a run), held out: the 8-D product peak of
    the main benchmark, run to convergence (maxeval 1e12)."""
    return _learned_family(ctx, 8, [W.c2_product_peak8(s) for s in range(200, 212)], [W.c2_product_peak8()],
                           10**12, 10**10, 400_000, policies)


def bench_learned(ctx):
    """Learned split guide, epsrel 1e-6, trained and tested on the GPU:
    * 7-D corner peaks: training seeds 100..115 (maxeval 1e11, 500k samples
      a run), held-out seeds 6, 7, 8;
    * 8-D product peaks (bench_learned_peak8)."""
    corner = _learned_family(
        ctx, 7, W.c5_train_problems(100, 16), [W.c5_corner7(s) for s in (6, 7, 8)], 10**11, 10**11, 500_000,
        {"plain": None,
         "learned_axis": dict(axis_mode=1, finish_mode=0, axis_confidence=0.6),
         "learned_axis_and_finish": dict(axis_mode=1, finish_mode=1, axis_confidence=0.6, budget=0.25),
         "volume_guard": dict(axis_mode=0, finish_mode=2, budget=0.25),
         "shared_allotment": dict(axis_mode=0, finish_mode=3, budget=0.5)})
    return {"corner_peak7": corner, "product_peak8": bench_learned_peak8(ctx)}


def bench_reference(ctx):
    """The reference's own method on MI355X: degree-9 rule (KEY 9), Phase 1
    breadth-first to 2^14 regions, Phase 2 per region (2048 bisections max),
    on the 8-D workload and the 7-D corner peak, against the default plain
    GM7/5 method on the same integrals."""
    rows = []
    keys = ("rc", "fail", "integral", "error", "nregions", "neval", "iterations")
    for p, me in ((W.c2_oscillatory8(), 10**9), (W.c2_product_peak8(), 10**12), (W.c5_corner7(), 10**11)):
        ref = min((ctx.integrate_two_phase(p, key=9) for _ in range(3)), key=lambda r: r["ms_total"])
        # the same two phases with the GM7/5 rule, whose |I7 - I5| error is
        # the certifiable one
        ref7 = min((ctx.integrate_two_phase(p, key=7) for _ in range(2)), key=lambda r: r["ms_total"])
        plain = min((ctx.integrate(p, maxeval=me) for _ in range(3)), key=lambda r: r["ms_total"])
        plain9 = min((ctx.integrate(p, maxeval=me, key=9) for _ in range(3)), key=lambda r: r["ms_total"])
        rows.append({"integrand": p.name,
                     "reference_method": {k: ref[k] for k in keys} | {"ms": ref["ms_total"]},
                     "two_phase_gm75": {k: ref7[k] for k in keys} | {"ms": ref7["ms_total"]},
                     "plain_gm75": {k: plain[k] for k in keys} | {"ms": plain["ms_total"], "maxeval": me},
                     "plain_degree9": {k: plain9[k] for k in keys} | {"ms": plain9["ms_total"], "maxeval": me}})
    return rows


def bench_table6(ctx):
    """The table-driven 6-D integrand (256 x 256 FP64 bilinear field) at its
    stated epsrel 1e-5 (maxeval 1e9): wall time, evals/s and the rule
    evaluation's roofline from HIP events (the integrand's flops as counted
    in region_flops plus the 4 table values read per point, 32 B, served by
    the cache: the 512 KiB table stays resident); and at epsrel 1e-3 the
    plain per-region test against the normalised allotment (guided loop,
    finish_mode 4, no network), which certifies it."""
    import cuhre
    p = W.c3_table6()
    ctx.integrate(p)  # warm-up
    best = min((ctx.integrate(p) for _ in range(3)), key=lambda r: r["ms_total"])
    ctx.set(True, 0)
    prof = ctx.integrate(p)
    ctx.set(False, 0)
    tf = EXEC_FLOPS_PER_REGION_T6 * prof["nregions"] / (prof["ms_eval"] * 1e-3) / 1e12
    pts = W.gm_npoints(6)
    row = {"epsrel": 1e-5, "maxeval": 10**9, "rc": best["rc"], "integral": best["integral"],
           "error": best["error"], "nregions": best["nregions"], "neval": best["neval"],
           "iterations": best["iterations"], "ms": best["ms_total"], "evals_per_s": best["neval"] / best["ms_total"] * 1e3,
           "eval_roofline": {"bound": "fp64", "achieved_tflops": tf, "frac": tf / FP64_PEAK_TFLOPS,
                             "ms_eval": prof["ms_eval"], "eval_launches": prof["eval_launches"],
                             "flops_per_region": EXEC_FLOPS_PER_REGION_T6,
                             "model_flops_per_region": prof["eval_flops"] / prof["nregions"],
                             "table_bytes_per_region": 32 * pts,
                             "table_GBps_from_cache": prof["nregions"] * 32 * pts / (prof["ms_eval"] * 1e-3) / 1e9}}
    g = cuhre.Guide(ctx, seed=1)
    plain3 = ctx.integrate(p, epsrel=1e-3)
    norm = min((g.integrate(p, axis_mode=0, finish_mode=4, budget=0.25, epsrel=1e-3) for _ in range(2)),
               key=lambda r: r["ms_total"])
    g.close()
    keys = ("rc", "integral", "error", "nregions", "iterations")
    return {"table6d_1e-5": row,
            "table6d_1e-3": {"plain": {k: plain3[k] for k in keys} | {"ms": plain3["ms_total"]},
                             "normalised_allotment_b0.25": {k: norm[k] for k in keys} | {"ms": norm["ms_total"]}}}


SWEEP_SIZES = (256, 4096, 16384, 65536, 262144, 1048576)
SWEEP_SEGMENTS = (1, 2, 4, 8, 16, 32)


def bench_sweep(ctx, reps=3):
    """Point-segment sweep of the rule evaluation: random 8-D regions of the
    product peak and the oscillatory integrand, every list size with every
    segment count (32 = the half-wavefront form, run up to 8192 regions),
    through cuhre_eval_regions.  The kernel times come from a profiler
    around this run; the list records the launches in order."""
    import numpy as np
    rng = np.random.default_rng(2024)
    order = []
    for p in (W.c2_product_peak8(), W.c2_oscillatory8()):
        for n in SWEEP_SIZES:
            L = rng.uniform(0.0, 0.5, size=(n, 8))
            Wd = rng.uniform(0.05, 0.5, size=(n, 8))
            for S in SWEEP_SEGMENTS:
                if S == 32 and n > 8192:
                    continue
                ctx.set(False, S)
                for _ in range(reps):
                    ctx.eval_regions(p.f, L, Wd)
                order.append({"integrand": p.name, "regions": n, "segments": S, "launches": reps})
    ctx.set(False, 0)
    return order


def lclt_tables():
    with open(os.path.join(ROOT, "tests", "golden", "lc_lt_t_tables.json")) as fh:
        return json.load(fh)


def bench_lclt(ctx):
    """The reference driver's run (cudaCuhre.cu:43-79): its LC_LT_t integrand
    over its Volume at epsrel 1e-12 then 1e-7 (epsabs 1e-12), by the
    reference's method (two phases, degree 9) and by the plain GM7/5 loop,
    best of 5."""
    t = lclt_tables()
    rows = []
    for eps in (1e-12, 1e-7):
        p = W.lclt_driver(t, eps)
        row = {"epsrel": eps}
        for name, run in (("reference_method", lambda: ctx.integrate_two_phase(p, key=9)),
                          ("plain_gm75", lambda: ctx.integrate(p)),
                          ("plain_degree9", lambda: ctx.integrate(p, key=9))):
            r = min((run() for _ in range(5)), key=lambda r: r["ms_total"])
            row[name] = {k: r[k] for k in ("rc", "fail", "integral", "error", "nregions", "neval",
                                           "iterations")} | {"ms": r["ms_total"]}
        rows.append(row)
    return rows


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--batched", action="store_true",
                    help="the batched comparison below (run by default; the flag is kept for older command lines)")
    ap.add_argument("--no-batched", action="store_true",
                    help="skip the default timing of the 4096-integral batched form against sequential plain calls")
    ap.add_argument("--learned", action="store_true",
                    help="also train the split guide and compare it with the plain method on 7-D corner peaks "
                         "(and 8-D product peaks, which the default run includes)")
    ap.add_argument("--no-learned8", action="store_true",
                    help="skip the default learned-vs-plain comparison on the 8-D product peak")
    ap.add_argument("--reference", action="store_true",
                    help="the reference's own method beside the plain one (run by default; the flag is kept for "
                         "older command lines)")
    ap.add_argument("--no-reference", action="store_true",
                    help="skip the default run of the reference's own method (degree-9 rule, two phases)")
    ap.add_argument("--no-table", action="store_true",
                    help="skip the default run of the table-driven 6-D integrand")
    ap.add_argument("--sweep", action="store_true",
                    help="also run the point-segment sweep of the rule evaluation (for a kernel profiler)")
    ap.add_argument("--lclt", action="store_true",
                    help="also run the reference driver's LC_LT_t integrand (its tables, its Volume, its tolerances)")
    ap.add_argument("--converge", action="store_true",
                    help="also run the product peak with maxeval 1e12 to reach epsrel 1e-6")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the path runs on a single MI355X "
              "(multi-GPU sharding is a design note, not built)", file=sys.stderr)
        sys.exit(2)

    import cuhre
    problems = [W.c2_oscillatory8(), W.c2_product_peak8()]
    ctx = cuhre.Context(0)
    for _ in range(args.warmup):
        for p in problems:
            ctx.integrate(p)
    # timed region: plain steps, no profiling events
    t0 = time.perf_counter()
    runs = []
    for _ in range(args.steps):
        runs.append([ctx.integrate(p) for p in problems])
    dt = time.perf_counter() - t0
    neval = sum(r["neval"] for step in runs for r in step)
    value = neval / dt

    # the same steps with the loop queueing one iteration at a time (not
    # the metric: the loop above queues each iteration before the host has
    # read the previous state; results are identical)
    ctx.queue_ahead(False)
    t1 = time.perf_counter()
    serial = [[ctx.integrate(p) for p in problems] for _ in range(args.steps)]
    dt_serial = time.perf_counter() - t1
    ctx.queue_ahead(True)
    same_serial = all(a["integral"] == b["integral"] and a["nregions"] == b["nregions"]
                      for sa, sb in zip(runs, serial) for a, b in zip(sa, sb))

    # kernel timing with HIP events on the engine's stream over the same steps
    ctx.set(True, 0)
    prof = [[ctx.integrate(p) for p in problems] for _ in range(args.steps)]
    ctx.set(False, 0)
    costs = flop_costs()
    pp = [step[1] for step in prof]
    # dominant kernel: rule evaluation of the product peak, every launch of
    # the timed steps (HIP events on the engine's stream)
    ev_ms = sum(r["ms_eval"] for r in pp)
    ev_launches = sum(r["eval_launches"] for r in pp)
    ev_regions = sum(r["nregions"] for r in pp)
    fl = region_flops(problems[1].f, 8, costs)
    xfl = EXEC_FLOPS_PER_REGION_PP8
    avg_ms = ev_ms / ev_launches
    achieved = (xfl * ev_regions / ev_launches) / (avg_ms * 1e-3) / 1e12
    model_tf = (fl * ev_regions / ev_launches) / (avg_ms * 1e-3) / 1e12
    big = max(pp, key=lambda r: r["largest_regions"])
    big_tf = xfl * big["largest_regions"] / (big["ms_eval_largest"] * 1e-3) / 1e12

    last = runs[-1]
    per = []
    for p, r, pr in zip(problems, last, prof[-1]):
        item = {"integrand": p.name, "integral": r["integral"], "error": r["error"],
                "fp64_err_bound": r["fp64_err_bound"], "nregions": r["nregions"],
                "neval": r["neval"], "fail": r["fail"], "iterations": r["iterations"],
                "peak_regions": r["peak_regions"], "ms": r["ms_total"],
                "ms_eval": pr["ms_eval"], "ms_split": pr["ms_split"]}
        per.append(item)

    out = {
        "metric": "integrand evals/s to epsrel 1e-6 (8-D), plain GPU method",
        "value": value,
        "unit": "evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": dt / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded SplitMix64 integrand parameters)",
        "config": {"workload": "oscillatory8d + product_peak8d on [0,1]^8, GM7/5 (401 pts/region), "
                               "epsrel 1e-6, maxeval 1e9 each",
                   "parallelism": "single GPU"},
        "roofline": {"bound": "fp64", "achieved": achieved, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / FP64_PEAK_TFLOPS, "traffic": TRAFFIC_PER_LAUNCH,
                     "issue_ceiling_frac": ISSUE_CEILING_PP8,
                     "frac_of_issue_ceiling": achieved / FP64_PEAK_TFLOPS / ISSUE_CEILING_PP8,
                     "flops_per_region": xfl,
                     "flops_note": "executed FP64 flops per region counted by SQ_INSTS_VALU_*_F64 "
                                   "(profiles/eval_product_peak8d.md); the point-by-point model "
                                   "(model_flops_per_region) charges every point the whole integrand, which "
                                   "this evaluation does not execute: it shares the integrand's "
                                   "per-coordinate partial results between points",
                     "model_flops_per_region": fl, "model_equivalent_tflops": model_tf,
                     "traffic_note": "HBM bytes per launch of this kernel on this workload, FETCH_SIZE x2 + "
                                     "WRITE_SIZE x1 (copy-kernel calibration), profiles/eval_product_peak8d.md; "
                                     "algorithmic 153 B/region x 159,134 regions = 24.3 MB",
                     "kernel": "rule evaluation of product_peak8d, every launch: k_eval_half<8, product_peak, 7> "
                               "(lists up to 8192 regions), k_eval_seg<8, product_peak> (2..16 segments), "
                               "k_eval_one<8, product_peak> (one segment, from 64k regions)",
                     "launches": ev_launches, "avg_launch_ms": avg_ms,
                     "regions_per_launch": ev_regions / ev_launches,
                     "largest_launch": {"regions": big["largest_regions"], "ms": big["ms_eval_largest"],
                                        "tflops": big_tf, "frac": big_tf / FP64_PEAK_TFLOPS,
                                        "frac_of_issue_ceiling": big_tf / FP64_PEAK_TFLOPS / ISSUE_CEILING_PP8}},
        "per_integrand": per,
        "queue_ahead": {"ms_per_step": dt / args.steps * 1e3,
                        "ms_per_step_one_iteration_at_a_time": dt_serial / args.steps * 1e3,
                        "speedup": dt_serial / dt, "results_identical": same_serial},
    }
    if args.converge:
        p = W.c2_product_peak8()
        ctx.integrate(p, maxeval=10**12)  # warm-up: the lists grow to their peak once
        t1 = time.perf_counter()
        r = ctx.integrate(p, maxeval=10**12)
        t2 = time.perf_counter()
        out["product_peak_to_convergence"] = {"maxeval": 10**12, "seconds": t2 - t1,
                                              "evals_per_s": r["neval"] / (t2 - t1), **{k: r[k] for k in (
                                                  "integral", "error", "nregions", "neval", "fail",
                                                  "iterations", "peak_regions")}}
        # the same certified epsrel with the shared-allotment finish policy of
        # the guided loop (per-region test err <= epsrel((1-b)|est| + b|I| vrel),
        # b = 0.5, no network): regions and wall time against the plain run
        g = cuhre.Guide(ctx, seed=1)
        g.integrate(p, axis_mode=0, finish_mode=3, budget=0.5, maxeval=10**12)
        t1 = time.perf_counter()
        a = g.integrate(p, axis_mode=0, finish_mode=3, budget=0.5, maxeval=10**12)
        t2 = time.perf_counter()
        conv = out["product_peak_to_convergence"]
        conv["shared_allotment"] = {"budget": 0.5, "seconds": t2 - t1, **{k: a[k] for k in (
            "rc", "integral", "error", "nregions", "neval", "iterations")},
            "regions_vs_plain": a["nregions"] / conv["nregions"], "time_vs_plain": (t2 - t1) / conv["seconds"]}
        g.close()
    if not args.no_batched:
        out["batched"] = bench_batched(ctx)
    if args.learned:
        out["learned"] = bench_learned(ctx)
    if not args.no_learned8:
        # the headline comparison: time to epsrel 1e-6 on the 8-D product
        # peak, learned guide (trained here, on other centres) vs plain
        pk = out["learned"]["product_peak8"] if args.learned else bench_learned_peak8(ctx)
        row = pk["held_out"][0]
        out["learned_vs_plain_8d"] = {
            "integrand": row["integrand"], "epsrel": 1e-6, "maxeval": pk["maxeval"],
            "training": {"samples": pk["samples"], "collect_and_train_s": pk["collect_and_train_s"]},
            **{name: {k: row[name][k] for k in ("rc", "nregions", "ms", "integral", "error", "fp64_err_bound")
                      if k in row[name]} | {k: row[name][k] for k in ("regions_vs_plain", "time_vs_plain")
                                            if k in row[name]}
               for name in PEAK8_POLICIES}}
    if not args.no_reference:
        out["reference_method"] = bench_reference(ctx)
    if not args.no_table:
        out["table6d"] = bench_table6(ctx)
    if args.lclt:
        out["lclt_driver"] = bench_lclt(ctx)
    if args.sweep:
        out["segment_sweep"] = bench_sweep(ctx)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(problems)
        # the CPU leg also holds the closed forms: report each GPU result's
        # distance from the exact integral there
        out["cpu_baseline"]["gpu_vs_exact"] = check_against_exact(problems, per,
                                                                  out.get("product_peak_to_convergence"))
        if "reference_method" in out:
            sys.path.insert(0, os.path.join(ROOT, "oracle"))
            import oracle
            for row, p in zip(out["reference_method"], (W.c2_oscillatory8(), W.c2_product_peak8(), W.c5_corner7())):
                ex, _ = oracle.exact(p.f, p.lows, p.highs)
                for k in ("reference_method", "two_phase_gm75", "plain_gm75", "plain_degree9"):
                    row[k]["true_rel_err"] = abs(row[k]["integral"] - ex) / abs(ex)
        if "table6d" in out:
            sys.path.insert(0, os.path.join(ROOT, "oracle"))
            import oracle
            p = W.c3_table6()
            ex, _ = oracle.exact(p.f, p.lows, p.highs)
            t = out["table6d"]
            for row in (t["table6d_1e-5"], t["table6d_1e-3"]["plain"], t["table6d_1e-3"]["normalised_allotment_b0.25"]):
                row["true_rel_err"] = abs(row["integral"] - ex) / abs(ex)
        if "batched" in out:
            sys.path.insert(0, os.path.join(ROOT, "oracle"))
            import oracle
            probs = W.c4_batch()
            for key in ("plain_policy", "allotment"):
                rs = out["batched"]["_results"][key]
                outside = sum(1 for p, r in zip(probs, rs)
                              if r["fail"] == 0 and abs(r["integral"] - oracle.exact(p.f, p.lows, p.highs)[0])
                              > r["error"])
                outside_tol = sum(1 for p, r in zip(probs, rs)
                                  if r["fail"] == 0 and abs(r["integral"] - oracle.exact(p.f, p.lows, p.highs)[0])
                                  > p.epsrel * abs(r["integral"]))
                certified = sum(1 for r in rs if r["fail"] == 0)
                tgt = out["batched"] if key == "plain_policy" else out["batched"]["allotment"]
                tgt["certified"] = certified
                tgt["certified_outside_error_of_closed_form"] = outside
                tgt["certified_outside_epsrel_of_closed_form"] = outside_tol
        if "learned_vs_plain_8d" in out:
            sys.path.insert(0, os.path.join(ROOT, "oracle"))
            import oracle
            p = W.c2_product_peak8()
            ex, _ = oracle.exact(p.f, p.lows, p.highs)
            for name in PEAK8_POLICIES:
                row = out["learned_vs_plain_8d"][name]
                row["true_rel_err"] = abs(row["integral"] - ex) / abs(ex)
        if args.lclt:
            # the oracle's restatement of the same runs on the host, timed
            sys.path.insert(0, os.path.join(ROOT, "oracle"))
            import oracle
            t = lclt_tables()
            for row in out["lclt_driver"]:
                p = W.lclt_driver(t, row["epsrel"])
                t0 = time.perf_counter()
                o = oracle.integrate_two_phase(p, key=9, nthreads=16)
                t1 = time.perf_counter()
                b = oracle.integrate_bfs(p, nthreads=16)
                t2 = time.perf_counter()
                row["cpu_oracle"] = {"reference_method": {"integral": o["integral"], "nregions": o["nregions"],
                                                          "ms": (t1 - t0) * 1e3},
                                     "plain_gm75": {"integral": b["integral"], "nregions": b["nregions"],
                                                    "ms": (t2 - t1) * 1e3}, "threads": 16}
    ctx.close()
    if "batched" in out:
        del out["batched"]["_results"]
    print(json.dumps(out))


if __name__ == "__main__":
    main()
