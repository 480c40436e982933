"""The FP64 rounding bound reported with every result, checked against a
__float128 shadow of the same run (CPU, the oracle).

What the bound claims (or_fp64_bound, the library's fp64_rounding_bound):
|I - sum_r V_r sum_k w_k f_k| <= bound, where the sum runs over the regions
the result adds up and f_k are the FP64 integrand values the loop obtained.
The shadow evaluates that sum beside the FP64 loop in __float128 (w_k * f_k
of two doubles is exact in 113 bits; the quad sums are accurate to ~1e-30
relative), so |I - istar| is the loop's true FP64 arithmetic error.

Pins:
* the bound holds on every workload below, including a cancelling
  oscillatory integral whose value is ~0 against sum |w f| ~ 1 and the 8-D
  product peak at maxeval 2e8 (131k contributing regions);
* on the cancelling integral the error is many orders of magnitude above
  what the bound's other terms (2u|I| + 2n^2u^2 E) allow, so the rule-sum
  term is necessary: dropping it would fail there;
* the per-region form (est vs the region's quad shadow) holds on random
  regions of every family with a quad restatement;
* with the nodes and the integrand in __float128 too (irule), the total
  FP64 error, integrand rounding included, also falls inside the bound on
  these workloads (empirical: the integrand's rounding is not part of the
  bound's derivation).

The reference sums with thrust::reduce and reports no rounding error
(quad/GPUquad/GPUKernelquad.cu:565-577).
"""
from fractions import Fraction as Fr

import numpy as np
import pytest

import oracle
import workloads as W

U = 2.0 ** -53


def _cases():
    return [
        ("c1_gaussian3", W.c1_gaussian3(), dict()),
        ("c1_gaussian3_key9", W.c1_gaussian3(), dict(key=9)),
        ("c2_oscillatory8", W.c2_oscillatory8(), dict()),
        ("cancelling_oscillatory5", W.cancelling_oscillatory(5), dict(epsrel=1e-7, epsabs=0.0, maxeval=2 * 10**7)),
        ("cancelling_oscillatory8", W.cancelling_oscillatory(8, epsabs=1e-6), dict()),
        ("c5_corner7_1e-3", W.c5_corner7(), dict(epsrel=1e-3)),
        ("c3_table6_1e-3", W.c3_table6(), dict(epsrel=1e-3)),
        ("c2_product_peak8_2e8", W.c2_product_peak8(), dict(maxeval=2 * 10**8, want_rule=False)),
    ]


CASES = _cases()


def _err(I, hi, lo):
    return abs(Fr(I) - (Fr(hi) + Fr(lo)))


@pytest.fixture(scope="module")
def runs():
    out = {}
    for name, p, kw in CASES:
        out[name] = (p, kw, oracle.integrate_bfs_check(p, nthreads=8, **kw))
    return out


@pytest.mark.parametrize("name", [c[0] for c in CASES])
def test_bound_holds_against_quad_shadow(runs, name):
    p, kw, r = runs[name]
    assert r["rc"] in (0, 1), r
    e = _err(r["integral"], r["istar_hi"], r["istar_lo"])
    assert e <= Fr(r["bound"]), (name, float(e), r["bound"])
    # the bound is the formula over the run's own A, E and n
    key = kw.get("key", 7)
    pts = oracle.npoints(p.f.dim, key)
    assert r["bound"] == oracle.fp64_bound(pts, p.f.dim, r["ncontrib"], r["abs_total"], r["aest_total"],
                                           r["integral"])
    if r["have_rule"]:
        er = _err(r["integral"], r["irule_hi"], r["irule_lo"])
        assert er <= Fr(r["bound"]), (name, float(er), r["bound"])


def test_rule_sum_term_is_needed_on_cancellation(runs):
    """On the cancelling integral the FP64 error exceeds everything except
    the bound's rule-sum term: a bound without it fails there."""
    for name in ("cancelling_oscillatory5", "cancelling_oscillatory8"):
        p, kw, r = runs[name]
        e = float(_err(r["integral"], r["istar_hi"], r["istar_lo"]))
        rest = 2 * U * abs(r["integral"]) + 2 * r["ncontrib"] ** 2 * U * U * r["aest_total"]
        assert e > 1e3 * rest, (name, e, rest)
        assert abs(r["integral"]) < 1e-12 * r["abs_total"], r


def test_cancelling_integral_is_zero():
    p = W.cancelling_oscillatory(5)
    hi, lo = oracle.exact(p.f, p.lows, p.highs)
    assert abs(hi) < 1e-15


@pytest.mark.parametrize("fam", ["gaussian", "product_peak", "oscillatory", "corner_peak", "continuous"])
@pytest.mark.parametrize("key", [7, 9])
def test_per_region_bound(fam, key):
    rng = np.random.default_rng(11)
    d = 6
    if fam == "gaussian":
        f = W.gaussian(rng.uniform(1, 8, d), rng.uniform(0, 1, d))
    elif fam == "product_peak":
        f = W.product_peak(rng.uniform(1, 8, d), rng.uniform(0, 1, d))
    elif fam == "oscillatory":
        f = W.oscillatory(rng.uniform(0, 2, d), 0.3)
    elif fam == "corner_peak":
        f = W.corner_peak(rng.uniform(0.5, 5, d), d + 1)
    else:
        f = W.continuous(rng.uniform(1, 8, d), rng.uniform(0, 1, d))
    pts = oracle.npoints(d, key)
    K = pts + d + 2
    gK = K * U / (1 - K * U)
    worst = 0.0
    for _ in range(40):
        lo = rng.uniform(0, 0.7, d)
        wd = rng.uniform(0.01, 0.3, d)
        est, ab, sh, sl, rh, rl = oracle.region_check(f, lo, wd, key)
        e = _err(est, sh, sl)
        assert e <= Fr(gK * ab), (fam, key, float(e), gK * ab)
        worst = max(worst, float(e) / (U * ab))
        if key == 7:
            # nodes and integrand in quad: within the same bound (empirical)
            assert _err(est, rh, rl) <= Fr(gK * ab)
    assert worst < K


def test_two_phase_bound_equals_loop_when_phase1_converges():
    """Converged in Phase 1, the two-phase method returns the loop's own
    result, and the same bound over the same regions."""
    p = W.c1_gaussian3()
    a = oracle.integrate_bfs(p, nthreads=4)
    b = oracle.integrate_two_phase(p, key=7, phase1_max=10**9, pool_max=16, nthreads=4)
    assert b["rc"] == 0 and a["rc"] == 0
    assert b["integral"] == a["integral"] and b["nregions"] == a["nregions"]
    # (the two-phase restatement sums A and E in plain FP64, the loop with
    # compensation: they agree to rounding)
    assert b["bound"] == pytest.approx(a["bound"], rel=1e-12)


def test_two_phase_bound_covers_phase2():
    """With Phase 2 the bound also carries each region's RES update chain
    (three roundings per bisection).  Only its size is checked here: the
    Phase-2 part has no quad shadow."""
    p = W.c2_product_peak8()
    r = oracle.integrate_two_phase(p, key=9, phase1_max=2048, pool_max=64, nthreads=8)
    assert r["rc"] in (0, 1)
    assert r["bound"] > 0
    assert r["bound"] < 1e-8 * abs(r["integral"])
