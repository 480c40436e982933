"""The plain loop with each iteration queued before the previous state is
read, against the loop that queues one iteration at a time, through the
C-ABI: every result field must be identical (the segment count of a launch
follows from the bound 2 n of the previous list and the reductions from
the real list length in both forms).  The cases stop on every condition:
convergence, the evaluation budget, list growth on the way (the lists
start small and grow when the children do not fit) and a non-finite value,
also while the first iterations run as fused half-wavefront launches (each
finding its regions from the previous list instead of a split launch).
"""
import json
import os

import pytest

import cuhre
import workloads as W

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(__file__), "golden")

FIELDS = ("integral", "error", "fp64_err_bound", "nregions", "neval", "fail", "iterations", "peak_regions",
          "nfinished", "rc")


def _both(ctx, p, **kw):
    ctx.queue_ahead(True)
    a = ctx.integrate(p, **kw)
    ctx.queue_ahead(False)
    try:
        b = ctx.integrate(p, **kw)
    finally:
        ctx.queue_ahead(True)
    return a, b


def _cases():
    with open(os.path.join(GOLDEN, "lc_lt_t_tables.json")) as fh:
        lclt = W.lclt_driver(json.load(fh), epsrel=1e-7)
    return [
        ("gaussian3d", W.c1_gaussian3(), {}),
        ("gaussian3d_key9", W.c1_gaussian3(), {"key": 9}),
        ("oscillatory8d", W.c2_oscillatory8(), {}),
        ("product_peak8d", W.c2_product_peak8(), {}),          # budget stop, lists grow to ~2.5M
        ("product_peak8d_1e7", W.c2_product_peak8(), {"maxeval": 10**7}),
        ("table6d", W.c3_table6(), {}),
        ("corner7d", W.c5_corner7(), {"maxeval": 10**9}),
        ("lc_lt_t3", lclt, {}),
        # the first iterations run as a chain of fused half-wavefront launches:
        # the smallest and largest initial grids, and a budget that stops the
        # loop inside the chain
        ("gaussian2d", W.Problem("g2", W.gaussian([6.0, 9.0], [0.3, 0.6]), [0.0] * 2, [1.0] * 2, 1e-9, 0.0,
                                 10**9), {}),
        ("product_peak10d", W.Problem("pp10", W.product_peak([3.0] * 10, [0.2 + 0.05 * i for i in range(10)]),
                                      [0.0] * 10, [1.0] * 10, 1e-6, 0.0, 10**8), {}),
        ("product_peak8d_budget_in_chain", W.c2_product_peak8(), {"maxeval": 401 * 1500}),
    ]


@pytest.mark.parametrize("name,p,kw", _cases(), ids=[c[0] for c in _cases()])
def test_queue_ahead_identical(name, p, kw):
    # a fresh context: its lists start small, so the long runs grow them
    with cuhre.Context(0) as ctx:
        print(f"[queue_ahead] {name} dim={p.f.dim}", flush=True)
        a, b = _both(ctx, p, **kw)
        for k in FIELDS:
            assert a[k] == b[k], (name, k, a[k], b[k])
        # and again on the grown lists (no growth on the way)
        a2, b2 = _both(ctx, p, **kw)
        for k in FIELDS:
            assert a2[k] == a[k] and b2[k] == a[k], (name, k)


def test_queue_ahead_nonfinite(ctx):
    p = W.Problem("bad", W.corner_peak([-1.0] * 5, 700), [0] * 5, [1] * 5, 1e-5, 0, 10**8)
    for on in (True, False, True):
        ctx.queue_ahead(on)
        with pytest.raises(cuhre.CuhreError) as e:
            ctx.integrate(p)
        assert e.value.code == cuhre.ERR_NONFINITE
    # the context still works afterwards, in both forms
    a, b = _both(ctx, W.c1_gaussian3())
    for k in FIELDS:
        assert a[k] == b[k], k
