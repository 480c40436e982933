"""Batched form: every instance of a batch must reproduce the plain method
run on that instance alone (same region counts and iterations, estimates
within 1e-12 relative), whatever the launch shape or graph use."""
import numpy as np
import pytest

import cuhre
import oracle as O
import workloads as W

pytestmark = pytest.mark.gpu


def _same(b, g, tag):
    assert b["fail"] == g["fail"], (tag, b, g)
    assert b["nregions"] == g["nregions"] and b["iterations"] == g["iterations"], (tag, b, g)
    assert b["neval"] == g["neval"]
    assert abs(b["integral"] - g["integral"]) <= 1e-12 * abs(g["integral"]) + 1e-300, (tag, b, g)
    assert abs(b["error"] - g["error"]) <= 1e-6 * g["error"] + 1e-300, (tag, b, g)


def test_batch_matches_plain_and_oracle(ctx):
    probs = W.c4_batch(count=48)
    res, info, rc = ctx.integrate_batch(probs)
    assert info["graph_used"] == 1
    assert rc == (0 if all(r["fail"] == 0 for r in res) else 1)
    for i, (p, b) in enumerate(zip(probs, res)):
        g = ctx.integrate(p)
        _same(b, g, ("plain", i))
        o = O.integrate_bfs(p, nthreads=8)
        assert b["nregions"] == o["nregions"] and b["iterations"] == o["iterations"], (i, b, o)
        assert abs(b["integral"] - o["integral"]) <= 1e-12 * abs(o["integral"])
        if b["fail"] == 0:
            ex, _ = O.exact(p.f, p.lows, p.highs)
            assert abs(b["integral"] - ex) <= max(b["error"], 1e-3 * p.epsrel * abs(ex)), (i, b, ex)


@pytest.mark.parametrize("segments,grid,graph", [(1, 256, True), (2, 1024, True), (4, 4096, False), (1, 2048, False)])
def test_batch_shapes_and_graph_agree(ctx, segments, grid, graph):
    probs = W.c4_batch(seed=9, count=64)
    ref, _, _ = ctx.integrate_batch(probs, point_segments=1, grid_blocks=2048, use_graph=True)
    res, info, _ = ctx.integrate_batch(probs, point_segments=segments, grid_blocks=grid, use_graph=graph)
    assert info["graph_used"] == int(graph)
    for i, (a, b) in enumerate(zip(res, ref)):
        assert a["nregions"] == b["nregions"] and a["iterations"] == b["iterations"]
        if segments == 1:  # same per-region summation order: bit-identical
            assert a["integral"] == b["integral"] and a["error"] == b["error"], i
        else:  # segments change the order of the per-region sums
            assert abs(a["integral"] - b["integral"]) <= 1e-12 * abs(b["integral"]), i
            assert abs(a["error"] - b["error"]) <= 1e-6 * b["error"], i


def test_batch_mixed_families_tolerances_budget_and_boxes(ctx):
    rng = W.SplitMix64(31)
    probs = []
    for i in range(24):
        d = 3
        a = [rng.uniform_range(0.5, 6) for _ in range(d)]
        w = [rng.uniform() for _ in range(d)]
        f = [W.oscillatory(a, rng.uniform()), W.continuous(a, w), W.gaussian(a, w),
             W.product_peak(a, w), W.corner_peak(a, 4)][i % 5]
        lo = [rng.uniform_range(-0.5, 0.2) for _ in range(d)]
        if f.family == W.CORNER_PEAK:
            lo = [abs(l) for l in lo]  # keep 1 + a.x away from zero
        hi = [l + rng.uniform_range(0.3, 1.5) for l in lo]
        eps = [1e-4, 1e-6, 1e-8][i % 3]
        maxeval = [10**9, 20_000, 10**7][i % 3]
        probs.append(W.Problem(f"m{i}", f, lo, hi, eps, 1e-14, maxeval))
    res, info, rc = ctx.integrate_batch(probs)
    for i, (p, b) in enumerate(zip(probs, res)):
        g = ctx.integrate(p)
        _same(b, g, ("mixed", i))
    assert info["total_regions"] == sum(r["nregions"] for r in res)


def test_batch_autotune_returns_a_valid_shape(ctx):
    probs = W.c4_batch(count=256)
    cfg, table = ctx.batch_autotune(probs)
    assert cfg["point_segments"] in (1, 2, 4) and cfg["grid_blocks"] in (256, 512, 1024, 2048, 4096)
    assert np.all(table > 0)
    res, info, _ = ctx.integrate_batch(probs)  # uses the tuned shape
    assert info["point_segments"] == cfg["point_segments"] and info["grid_blocks"] == cfg["grid_blocks"]
    ref, _, _ = ctx.integrate_batch(probs, point_segments=1, grid_blocks=1024)
    for a, b in zip(res, ref):
        assert a["nregions"] == b["nregions"]
        assert abs(a["integral"] - b["integral"]) <= 1e-12 * abs(b["integral"])


def test_batch_capacity_cut_and_errors(ctx):
    probs = W.c4_batch(seed=12, count=32)
    with cuhre.Context(0, capacity=4096) as small:
        res, info, rc = small.integrate_batch(probs)
    assert rc == 1 and info["not_converged"] > 0
    full, _, _ = ctx.integrate_batch(probs)
    for a, b in zip(res, full):
        if a["fail"] == 0:
            assert a["nregions"] == b["nregions"]
            assert abs(a["integral"] - b["integral"]) <= 1e-12 * abs(b["integral"])
    bad = probs[:3] + [W.Problem("bad", W.corner_peak([-1.0] * 5, 700), [0] * 5, [1] * 5, 1e-5, 0, 10**8)]
    with pytest.raises(cuhre.CuhreError) as e:
        ctx.integrate_batch(bad)
    assert e.value.code == cuhre.ERR_NONFINITE
    # the non-finite instance is marked in its own result; the others are
    # complete and equal their plain runs
    res, info, rc = ctx.integrate_batch(bad, raise_nonfinite=False)
    assert rc == cuhre.ERR_NONFINITE
    assert res[3]["fail"] == cuhre.ERR_NONFINITE
    assert info["not_converged"] >= 1
    for i in range(3):
        pl = ctx.integrate(bad[i])
        assert res[i]["fail"] == pl["fail"]
        assert res[i]["nregions"] == pl["nregions"]
        assert abs(res[i]["integral"] - pl["integral"]) <= 1e-12 * abs(pl["integral"])
    with pytest.raises(cuhre.CuhreError) as e:
        ctx.integrate_batch([W.c1_gaussian3(), W.c2_oscillatory8()])
    assert e.value.code == cuhre.ERR_ARG
    # the context is still usable
    assert ctx.integrate(W.c1_gaussian3())["rc"] == 0


@pytest.mark.parametrize("graph", [True, False])
def test_batch_list_growth_keeps_results(ctx, graph):
    # start from a 4096-region list so the run has to grow it several times
    probs = W.c4_batch(seed=21, count=40)
    ref, rinfo, _ = ctx.integrate_batch(probs, point_segments=1, grid_blocks=1024)
    res, info, _ = ctx.integrate_batch(probs, point_segments=1, grid_blocks=1024, use_graph=graph,
                                       initial_capacity=4096)
    assert info["grows"] >= 2 and info["not_converged"] == rinfo["not_converged"]
    for a, b in zip(res, ref):
        assert a["nregions"] == b["nregions"] and a["iterations"] == b["iterations"]
        assert a["integral"] == b["integral"] and a["error"] == b["error"]


def test_batch_allotment_matches_oracle(ctx):
    """The shared-allotment finish policy per instance (finish_budget 0.25)
    against the oracle's restatement of the same policy
    (or_integrate_bfs_allot): same region counts and iterations, estimates
    within 1e-12; every instance certified by its own global test and within
    the requested tolerance of the closed form (or within its reported
    error)."""
    probs = W.c4_batch(count=48)
    res, info, rc = ctx.integrate_batch(probs, finish_budget=0.25)
    assert rc == 0 and info["not_converged"] == 0
    for i, (p, b) in enumerate(zip(probs, res)):
        o = O.integrate_bfs_allot(p, 0.25, nthreads=8)
        assert b["fail"] == 0 and o["rc"] == 0, (i, b, o)
        assert b["nregions"] == o["nregions"] and b["iterations"] == o["iterations"], (i, b, o)
        assert abs(b["integral"] - o["integral"]) <= 1e-12 * abs(o["integral"])
        assert b["error"] <= p.epsrel * abs(b["integral"])
        ex, _ = O.exact(p.f, p.lows, p.highs)
        assert abs(b["integral"] - ex) <= max(b["error"], p.epsrel * abs(ex)), (i, b, ex)


def test_batch_allotment_on_the_4096_integrals(ctx):
    """The 4096 5-D integrals of the batched workload.  The Genz-Malik error
    is an estimate, not a bound: with the plain per-region test a few sharp
    product peaks and Gaussians (a_i up to 10) end certified but outside the
    requested tolerance of the closed form.  With the allotment every
    instance certifies at maxeval 1e10, on at most half the plain policy's
    regions, and such results are about as rare as with the plain policy."""
    probs = W.c4_batch()
    plain, pinfo, _ = ctx.integrate_batch(probs)
    res, info, rc = ctx.integrate_batch(probs, finish_budget=0.25)
    exact = [O.exact(p.f, p.lows, p.highs)[0] for p in probs]

    def misses(rs):
        out_err = sum(1 for r, ex in zip(rs, exact) if r["fail"] == 0 and abs(r["integral"] - ex) > r["error"])
        out_tol = sum(1 for p, r, ex in zip(probs, rs, exact)
                      if r["fail"] == 0 and abs(r["integral"] - ex) > p.epsrel * abs(ex))
        return out_err, out_tol

    pm, am = misses(plain), misses(res)
    print("regions plain policy", pinfo["total_regions"], "allotment", info["total_regions"],
          "not converged", pinfo["not_converged"], info["not_converged"],
          "outside error / outside tolerance: plain", pm, "allotment", am)
    assert rc == 0 and info["not_converged"] == 0
    assert info["total_regions"] <= 0.5 * pinfo["total_regions"]
    # measured: 52 (plain) and 56 (allotment) of 4096 certified results lie
    # beyond epsrel of the closed form
    assert am[1] <= pm[1] + max(8, pm[1] // 4)


# ------------------------------------------------------------ persistent form
# One persistent launch in which each workgroup takes instances from a queue
# and runs each one's loop in its own region queue; an instance that
# outgrows its queue is handed over (its list and sums) to the replayed
# iterations.  Per instance the result equals the replayed form's: same
# region counts, iterations and flags, sums within rounding.


@pytest.mark.parametrize("budget", [0.0, 0.25])
def test_persistent_matches_replayed_iterations(ctx, budget):
    probs = W.c4_batch(seed=31, count=96)
    ref, rinfo, rrc = ctx.integrate_batch(probs, finish_budget=budget)
    res, info, rc = ctx.integrate_batch(probs, finish_budget=budget, persistent=True)
    assert rinfo["persistent_finished"] == -1
    # most instances finish inside the persistent launch, the largest are
    # handed over (their lists outgrow the 16384-region queues)
    print("persistent_finished", info["persistent_finished"], "of", len(probs))
    assert 0 < info["persistent_finished"] < len(probs)
    assert rc == rrc and info["not_converged"] == rinfo["not_converged"]
    assert info["total_regions"] == rinfo["total_regions"]
    for i, (a, b) in enumerate(zip(res, ref)):
        _same(a, b, ("persistent", budget, i))


def test_persistent_only_small_instances_and_oracle(ctx):
    """Instances small enough to finish inside the persistent launch (no
    hand-over): the plain method per instance and the oracle."""
    pool = [W.Problem(p.name, p.f, p.lows, p.highs, 1e-4, 0.0, p.maxeval) for p in W.c4_batch(seed=32, count=64)]
    # lists that stay well inside a 16384-region queue
    probs = [p for p in pool if ctx.integrate(p)["peak_regions"] <= 4096][:24]
    assert len(probs) >= 12
    res, info, rc = ctx.integrate_batch(probs, persistent=True)
    assert info["persistent_finished"] == len(probs)
    for i, (p, b) in enumerate(zip(probs, res)):
        g = ctx.integrate(p)
        _same(b, g, ("plain", i))
        o = O.integrate_bfs(p, nthreads=8)
        assert b["nregions"] == o["nregions"] and b["iterations"] == o["iterations"], (i, b, o)
        assert abs(b["integral"] - o["integral"]) <= 1e-12 * abs(o["integral"])


def test_persistent_nonfinite_and_small_lists(ctx):
    bad = W.c4_batch(seed=33, count=3) + [
        W.Problem("bad", W.corner_peak([-1.0] * 5, 700), [0] * 5, [1] * 5, 1e-5, 0, 10**8)]
    res, info, rc = ctx.integrate_batch(bad, raise_nonfinite=False, persistent=True)
    assert rc == cuhre.ERR_NONFINITE and res[3]["fail"] == cuhre.ERR_NONFINITE
    for i in range(3):
        pl = ctx.integrate(bad[i])
        _same(res[i], pl, ("nonfinite batch", i))
    # lists too short for useful queues: the replayed iterations alone
    probs = W.c4_batch(seed=34, count=16)
    with cuhre.Context(0, capacity=8192) as small:
        a, ia, _ = small.integrate_batch(probs, persistent=True)
        b, ib, _ = small.integrate_batch(probs)
    assert ia["persistent_finished"] == -1
    for x, y in zip(a, b):
        assert x["nregions"] == y["nregions"] and x["integral"] == y["integral"]
