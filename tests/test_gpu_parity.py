"""GPU parity of the HIP path against the CPU oracle (through the C-ABI).

Bars:
  * per-region est / err / I5: |gpu - oracle| <= 1e-13 * volume*sum|w7 f|
    (FP64 rule evaluation; the bound is relative to the absolute weighted
    sum because GM weights have mixed signs);
  * split axis identical unless the two candidate fourth differences are a
    rounding-level tie;
  * compaction + bisection: bit-exact;
  * whole integrations: same region counts as the oracle's loop, estimates
    within 1e-12 relative of each other and within epsrel of the closed form.
"""
import json
import math
import os

import numpy as np
import pytest

import cuhre
import oracle as O
import workloads as W

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def random_regions(rng, n, d, lo=0.0, hi=1.0):
    L = np.empty((n, d)); Wd = np.empty((n, d))
    for i in range(n):
        for m in range(d):
            w = rng.uniform_range(0.02, 0.5) * (hi - lo)
            L[i, m] = lo + rng.uniform() * (hi - lo - w)
            Wd[i, m] = w
    return L, Wd


def family_instance(fam, d, rng):
    a = [rng.uniform_range(0.5, 5.0) for _ in range(d)]
    w = [rng.uniform() for _ in range(d)]
    if fam == W.GAUSSIAN:
        return W.gaussian(a, w)
    if fam == W.PRODUCT_PEAK:
        return W.product_peak(a, w)
    if fam == W.OSCILLATORY:
        return W.oscillatory(a, rng.uniform())
    if fam == W.CORNER_PEAK:
        return W.corner_peak(a, d + 1)
    if fam == W.CONTINUOUS:
        return W.continuous(a, w)
    if fam == W.POLYNOMIAL:
        terms = []
        for t in range(4):
            pw = [int(4 * rng.uniform()) for _ in range(d)]
            terms.append((rng.uniform_range(-1, 1), pw))
        return W.polynomial(d, terms)
    if fam == W.COSINE_SUM:
        return W.cosine_sum(d)
    raise ValueError(fam)


def assert_region_parity(g, o, d, tag, vol):
    est, err, ax, ab, i5, dv = g
    oest, oerr, oax, oab, oi5, odv = o
    scale = np.maximum(oab, 1e-300)
    fscale = scale / vol  # sum |w7 f|: the size of the f values behind D_i
    assert np.all(np.abs(est - oest) <= 1e-13 * scale), tag
    assert np.all(np.abs(err - oerr) <= 1e-13 * scale), tag
    assert np.all(np.abs(i5 - oi5) <= 1e-13 * scale), tag
    assert np.all(np.abs(ab - oab) <= 1e-12 * scale), tag
    dscale = np.maximum(np.max(odv, axis=1, keepdims=True), 1e-300)
    assert np.all(np.abs(dv - odv) <= 1e-10 * dscale + 1e-14 * fscale[:, None]), tag
    for r in np.nonzero(ax != oax)[0]:
        # only a rounding-level tie may flip the chosen axis
        gap = abs(odv[r, ax[r]] - odv[r, oax[r]])
        assert gap <= 1e-10 * dscale[r, 0] + 1e-14 * fscale[r], (tag, r, ax[r], oax[r], odv[r])


FAMILIES = [W.GAUSSIAN, W.PRODUCT_PEAK, W.OSCILLATORY, W.CORNER_PEAK, W.CONTINUOUS, W.POLYNOMIAL,
            W.COSINE_SUM]


@pytest.mark.parametrize("lanes", [1, 4, 16, 32])
@pytest.mark.parametrize("d", [2, 3, 5, 8])
@pytest.mark.parametrize("fam", FAMILIES)
def test_region_values_match_oracle(ctx, fam, d, lanes):
    # lanes = point segments per group of 64 regions
    rng = W.SplitMix64(1000 * fam + 10 * d + lanes)
    f = family_instance(fam, d, rng)
    n = 300 if d < 8 else 120
    L, Wd = random_regions(rng, n, d)
    ctx.set(False, lanes)
    try:
        g = ctx.eval_regions(f, L, Wd, want_aux=True)
    finally:
        ctx.set(False, 0)
    o = O.region_eval(f, L, Wd)
    assert_region_parity(g, o, d, (fam, d, lanes), np.prod(Wd, axis=1))


@pytest.mark.parametrize("lanes", [1, 2, 8, 32])
def test_region_values_dim10_and_table(ctx, lanes):
    rng = W.SplitMix64(77 + lanes)
    f = family_instance(W.PRODUCT_PEAK, 10, rng)
    L, Wd = random_regions(rng, 70, 10)
    ctx.set(False, lanes)
    try:
        g = ctx.eval_regions(f, L, Wd, want_aux=True)
        o = O.region_eval(f, L, Wd)
        assert_region_parity(g, o, 10, ("pp10", lanes), np.prod(Wd, axis=1))
        p = W.c3_table6()
        L, Wd = random_regions(rng, 200, 6)
        L[:, 2] = 0.1 + 0.9 * L[:, 2]
        Wd[:, 2] *= 0.9
        g = ctx.eval_regions(p.f, L, Wd, want_aux=True)
        o = O.region_eval(p.f, L, Wd)
        assert_region_parity(g, o, 6, ("table6", lanes), np.prod(Wd, axis=1))
    finally:
        ctx.set(False, 0)


def test_region_values_match_scipy_golden(ctx):
    g = json.load(open(os.path.join(GOLDEN, "gm_scipy_regions.json")))
    for c in g["cases"]:
        f = W.Integrand(c["family"], c["dim"], c["a"], c["w"], u=c["u"], exponent=c["exponent"])
        L = np.array([r["lower"] for r in c["regions"]])
        Wd = np.array([r["width"] for r in c["regions"]])
        est, err, ax, ab = ctx.eval_regions(f, L, Wd)
        for k, r in enumerate(c["regions"]):
            assert abs(est[k] - r["est"]) <= 1e-13 * ab[k]
            assert abs(err[k] - r["err"]) <= 1e-13 * ab[k]


@pytest.mark.parametrize("n", [1, 63, 256, 1000, 70001])
def test_split_bit_exact(ctx, n):
    d = 5
    rng = W.SplitMix64(n)
    L, Wd = random_regions(rng, n, d)
    est = np.array([rng.uniform_range(-1, 1) for _ in range(n)])
    err = np.array([rng.uniform() * 2e-6 for _ in range(n)])
    ax = np.array([int(d * rng.uniform()) for _ in range(n)], dtype=np.int32)
    eps = 1e-6
    A, cL, cW = ctx.split_regions(L, Wd, est, err, ax, eps)
    act = (err > np.maximum(eps * np.abs(est), 0.0)).astype(np.int32)
    oL, oW = O.split(L, Wd, act, ax)
    assert A == int(act.sum())
    assert np.array_equal(cL, oL) and np.array_equal(cW, oW)


def test_split_none_and_all_active(ctx):
    d = 3
    rng = W.SplitMix64(5)
    L, Wd = random_regions(rng, 500, d)
    est = np.ones(500)
    ax = np.zeros(500, dtype=np.int32)
    A, cL, cW = ctx.split_regions(L, Wd, est, np.zeros(500), ax, 1e-6)
    assert A == 0 and cL.shape[0] == 0
    A, cL, cW = ctx.split_regions(L, Wd, est, np.ones(500), ax, 1e-6)
    assert A == 500
    assert np.allclose(cW[0::2, 0] * 2, Wd[:, 0]) and np.array_equal(cL[0::2], L)


def test_sum_matches_fsum(ctx):
    rng = np.random.default_rng(3)
    for n in (1, 255, 4097, 1_000_003):
        x = rng.standard_normal(n) * np.exp(rng.uniform(-20, 20, n))
        s = ctx.sum(x)
        ex = math.fsum(x)
        assert abs(s - ex) <= 4 * np.finfo(float).eps * abs(ex) + 1e-300


def _compare_integrations(g, o, p, exact=None, exact_tol=None):
    assert g["rc"] == o["rc"], (g, o)
    assert g["nregions"] == o["nregions"] and g["iterations"] == o["iterations"], (g, o)
    assert g["neval"] == o["neval"]
    assert abs(g["integral"] - o["integral"]) <= 1e-12 * abs(o["integral"]) + 1e-15 * g["error"]
    assert abs(g["error"] - o["error"]) <= 1e-6 * o["error"]
    if exact is not None:
        assert abs(g["integral"] - exact) <= exact_tol


def test_integrate_gaussian3d(ctx):
    p = W.c1_gaussian3()
    g = ctx.integrate(p)
    o = O.integrate_bfs(p, nthreads=8)
    ex, _ = O.exact(p.f, p.lows, p.highs)
    _compare_integrations(g, o, p, ex, p.epsrel * abs(ex))
    assert g["rc"] == 0 and g["error"] <= p.epsrel * abs(g["integral"])
    assert g["fp64_err_bound"] < 1e-3 * g["error"]


def test_integrate_oscillatory8d(ctx):
    p = W.c2_oscillatory8()
    g = ctx.integrate(p)
    o = O.integrate_bfs(p, nthreads=8)
    ex, _ = O.exact(p.f, p.lows, p.highs)
    _compare_integrations(g, o, p, ex, p.epsrel * abs(ex))


@pytest.mark.parametrize("segments", [0, 1, 16])
def test_integrate_product_peak_budget_and_growth(ctx, segments):
    # 2e8 evaluations: the list outgrows the first allocation (growth path)
    p = W.c2_product_peak8()
    with cuhre.Context(0, segments=segments) as fresh:
        g = fresh.integrate(p, maxeval=200_000_000)
    o = O.integrate_bfs(p, nthreads=8, maxeval=200_000_000)
    assert g["rc"] == 1 and g["fail"] == 1 and g["neval"] <= 200_000_000
    _compare_integrations(g, o, p)
    ex, _ = O.exact(p.f, p.lows, p.highs)
    assert abs(g["integral"] - ex) <= g["error"]


def test_integrate_large_lists_two_level_finalisation(ctx):
    """maxeval 2e10 takes the 8-D product peak to lists of ~12 M regions:
    evaluation launches of ~47k blocks, whose sums go through more than 256
    groups of 64 block partials (the block-wide top level of the two-level
    finalisation) and whose split adds group bases to the block offsets."""
    p = W.c2_product_peak8()
    g = ctx.integrate(p, maxeval=20_000_000_000)
    assert g["peak_regions"] > 256 * 64 * 256  # more than 256 groups of 64 blocks of 256 regions
    o = O.integrate_bfs(p, nthreads=16, maxeval=20_000_000_000)
    _compare_integrations(g, o, p)


def test_integrate_reference_driver_pattern(ctx):
    # cudaCuhre.cu:43-83: construct once, integrate at two tolerances on the
    # box [1,2]x[.1,.3]x[0,1]; here with the driver's cosine-sum functor
    f = W.cosine_sum(3)
    lows, highs = [1.0, 0.1, 0.0], [2.0, 0.3, 1.0]
    ex, _ = O.exact(f, lows, highs)
    for eps in (1e-7, 1e-10):
        p = W.Problem("drv", f, lows, highs, eps, 1e-12, 10**9)
        g = ctx.integrate(p)
        o = O.integrate_bfs(p, nthreads=8)
        _compare_integrations(g, o, p, ex, max(eps * abs(ex), 1e-12))


def test_integrate_table6_and_corner7_short(ctx):
    p = W.c3_table6()
    g = ctx.integrate(p, epsrel=1e-3)
    o = O.integrate_bfs(p, nthreads=8, epsrel=1e-3)
    ex, _ = O.exact(p.f, p.lows, p.highs)
    _compare_integrations(g, o, p, ex, 1e-3 * abs(ex))
    p = W.c5_corner7()
    g = ctx.integrate(p, maxeval=50_000_000)
    o = O.integrate_bfs(p, nthreads=8, maxeval=50_000_000)
    _compare_integrations(g, o, p)


def test_deterministic_reruns(ctx):
    p = W.c2_product_peak8()
    a = ctx.integrate(p, maxeval=20_000_000)
    b = ctx.integrate(p, maxeval=20_000_000)
    for k in ("integral", "error", "fp64_err_bound", "nregions", "neval", "iterations"):
        assert a[k] == b[k]


def test_errors_are_codes(ctx):
    p = W.c1_gaussian3()
    bad = W.Problem("bad", W.corner_peak([-1.0, -1.0], 700), [0, 0], [1, 1], 1e-6, 0, 10**6)
    with pytest.raises(cuhre.CuhreError) as e:
        ctx.integrate(bad)
    assert e.value.code == cuhre.ERR_NONFINITE
    with pytest.raises(cuhre.CuhreError) as e:
        ctx.integrate(p, key=5)
    assert e.value.code == cuhre.ERR_UNSUPPORTED
    with pytest.raises(cuhre.CuhreError) as e:
        ctx.integrate(W.Problem("box", p.f, [0, 0, 1], [1, 1, 1], 1e-6, 0, 10**6))
    assert e.value.code == cuhre.ERR_ARG
    f11 = W.gaussian([1.0] * 11, [0.5] * 11)
    with pytest.raises(cuhre.CuhreError) as e:
        ctx.integrate(W.Problem("d11", f11, [0] * 11, [1] * 11, 1e-3, 0, 10**6))
    assert e.value.code == cuhre.ERR_UNSUPPORTED
    # the context stays usable after errors
    assert ctx.integrate(p)["rc"] == 0


def test_capacity_stop(ctx):
    import cuhre as cu
    p = W.c2_product_peak8()
    with cu.Context(0, capacity=4096) as small:
        g = small.integrate(p)
    o = O.integrate_bfs(p, nthreads=8, capacity=4096)
    assert g["rc"] == 1
    _compare_integrations(g, o, p)


@pytest.mark.parametrize("lanes", [1, 4, 16])
def test_table6_clustered_small_regions_match_oracle(ctx, lanes):
    """Short clustered regions of the table-driven integrand (the late
    iterations' shape, where rule points fall close to the table's cell
    edges): per-region values, fourth differences included, against the
    oracle.  The cell weights x (n - 1) - floor(...) are computed without
    FMA contraction, as the oracle computes them; with contraction they
    differed by up to an ulp of x (n - 1) and moved D_i by ~1e-14."""
    rng = np.random.default_rng(31 + lanes)
    p = W.c3_table6()
    n = 4096
    L = np.empty((n, 6)); Wd = np.empty((n, 6))
    # 16 clusters of 256 regions, each cluster inside a 0.1 x 0.1 patch of (x1, x2)
    for c in range(16):
        x0, y0 = rng.uniform(0, 0.9, 2)
        sl = slice(256 * c, 256 * (c + 1))
        Wd[sl, 0] = rng.uniform(0.001, 0.02, 256)
        Wd[sl, 1] = rng.uniform(0.001, 0.02, 256)
        L[sl, 0] = x0 + rng.uniform(0, 0.08, 256)
        L[sl, 1] = y0 + rng.uniform(0, 0.08, 256)
    Wd[:, 2:] = rng.uniform(0.05, 0.2, (n, 4))
    L[:, 2] = 0.1 + rng.uniform(0, 0.7, n)
    L[:, 3:] = rng.uniform(0, 0.8, (n, 3))
    # and a few wide ones
    Wd[-8:, 0] = 0.9
    L[-8:, 0] = 0.05
    ctx.set(False, lanes)
    try:
        g = ctx.eval_regions(p.f, L, Wd, want_aux=True)
    finally:
        ctx.set(False, 0)
    o = O.region_eval(p.f, L, Wd)
    assert_region_parity(g, o, 6, ("table6 clustered", lanes), np.prod(Wd, axis=1))
