"""The GPU's FP64 rounding bound against the oracle's __float128 shadow.

The GPU run and the oracle run evaluate the same regions (equal region
counts and iterations are asserted); istar is the exact-arithmetic sum of
the oracle's FP64 integrand values over them.  The GPU's own integrand
values differ from the oracle's by the device math (exp, cos and the
reciprocal are each within an ulp or two of the correctly rounded value),
so the GPU's exact-arithmetic sum is within 4 u A of istar (A = sum of
V sum |w f|): the check is |I_gpu - istar| <= bound_gpu + 4 u A.  The
bound is the same formula as the oracle's over the run's own A, E and n,
so the two bounds agree to rounding.

The reference's accumulation (thrust::reduce, quad/GPUquad/
GPUKernelquad.cu:565-577) reports no rounding error.
"""
from fractions import Fraction as Fr

import numpy as np
import pytest

import cuhre
import oracle as O
import workloads as W

pytestmark = pytest.mark.gpu
U = 2.0 ** -53

CASES = [
    ("c1_gaussian3", lambda: W.c1_gaussian3(), dict(), 0),
    ("c1_gaussian3_key9", lambda: W.c1_gaussian3(), dict(), 9),
    ("c2_oscillatory8", lambda: W.c2_oscillatory8(), dict(), 0),
    ("cancelling_oscillatory5", lambda: W.cancelling_oscillatory(5), dict(epsrel=1e-7, maxeval=2 * 10**7), 0),
    ("cancelling_oscillatory8", lambda: W.cancelling_oscillatory(8, epsabs=1e-6), dict(), 0),
    ("c2_product_peak8_2e8", lambda: W.c2_product_peak8(), dict(maxeval=2 * 10**8), 0),
]


@pytest.mark.parametrize("name,mk,kw,key", CASES, ids=[c[0] for c in CASES])
def test_gpu_bound_holds(ctx, name, mk, kw, key):
    p = mk()
    eps = kw.get("epsrel", p.epsrel)
    if "epsrel" in kw:
        p.epsrel = eps
    me = kw.get("maxeval", p.maxeval)
    print(f"[bound] {name} key={key} maxeval={me}", flush=True)
    g = ctx.integrate(p, maxeval=me, key=key)
    o = O.integrate_bfs_check(p, nthreads=8, maxeval=me, key=7 if key == 0 else key, want_rule=False)
    assert g["rc"] == o["rc"]
    assert g["nregions"] == o["nregions"] and g["iterations"] == o["iterations"]
    e = abs(Fr(g["integral"]) - (Fr(o["istar_hi"]) + Fr(o["istar_lo"])))
    slack = 4 * U * o["abs_total"]
    assert e <= Fr(g["fp64_err_bound"]) + Fr(slack), (name, float(e), g["fp64_err_bound"], slack)
    # same formula over the same regions
    assert g["fp64_err_bound"] == pytest.approx(o["bound"], rel=1e-9)


@pytest.mark.parametrize("key", [0, 9])
def test_gpu_per_region_bound(ctx, key):
    """est_gpu within gamma_K * abssum (+ 4 u abssum for the device math) of
    the oracle's quad shadow of the same region."""
    rng = np.random.default_rng(5)
    d = 8
    f = W.product_peak(rng.uniform(1, 8, d), rng.uniform(0, 1, d))
    n = 512
    lo = rng.uniform(0, 0.7, (n, d))
    wd = rng.uniform(0.01, 0.3, (n, d))
    est, err, ax, ab = ctx.eval_regions(f, lo, wd, key=key)
    pts = O.npoints(d, 9 if key == 9 else 7)
    K = pts + d + 2
    gK = K * U / (1 - K * U)
    for r in range(0, n, 7):
        c = O.region_check(f, lo[r], wd[r], 9 if key == 9 else 7)
        e = abs(Fr(est[r]) - (Fr(c[2]) + Fr(c[3])))
        assert e <= Fr((gK + 4 * U) * ab[r]), (r, float(e), gK * ab[r])


def test_gpu_two_phase_bound(ctx):
    """Two-phase: the GPU's bound carries the same Phase-2 budget (every
    evaluated region's rule sum, the RES update chains) as the oracle's
    restatement; the region sets agree to a fraction of a percent, so the
    bounds agree to a few percent."""
    p = W.c1_gaussian3()
    g = ctx.integrate_two_phase(p, key=9, phase1_regions=256, pool_max=64)
    o = O.integrate_two_phase(p, key=9, phase1_max=256, pool_max=64, nthreads=8)
    assert g["rc"] == o["rc"] and g["nzeroed"] == 0
    assert g["fp64_err_bound"] > 0
    assert g["fp64_err_bound"] == pytest.approx(o["bound"], rel=0.05)
