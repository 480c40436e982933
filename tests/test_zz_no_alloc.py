"""cuhre_integrate_noalloc against cuhre_integrate, through the C-ABI.

The no-allocation entry point runs the same kernels in the same order on
region lists carved from the library's static device array, so every field
of its result must equal the allocating entry point's bit for bit.  Calls
come in a row on one context with inputs of decreasing size (the array
keeps the previous call's contents), then a run that outgrows the array
(a clear CUHRE_ERR_CAPACITY), after which both entry points still work.
"""
import json
import os

import pytest

import cuhre
import workloads as W

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(__file__), "golden")

FIELDS = ("integral", "error", "fp64_err_bound", "nregions", "neval", "fail", "iterations", "peak_regions",
          "nfinished")


def _same(a, b):
    for k in FIELDS:
        assert a[k] == b[k], (k, a[k], b[k])
    assert a["rc"] == b["rc"]


def _lclt():
    with open(os.path.join(GOLDEN, "lc_lt_t_tables.json")) as fh:
        return W.lclt_driver(json.load(fh), epsrel=1e-7)


def test_noalloc_equals_integrate_in_a_row():
    cases = [
        ("oscillatory8d", W.c2_oscillatory8(), {}),
        ("gaussian3d", W.c1_gaussian3(), {}),
        ("gaussian3d_key9", W.c1_gaussian3(), {"key": 9}),
        ("lc_lt_t3", _lclt(), {}),
        ("table6d_1e-3", W.c3_table6(), {"epsrel": 1e-3}),
        ("gaussian3d_again", W.c1_gaussian3(), {}),
    ]
    with cuhre.Context(0) as ctx:
        for name, p, kw in cases:
            print(f"[noalloc] {name} dim={p.f.dim} capacity={cuhre.noalloc_capacity(p.f.dim)}", flush=True)
            a = ctx.integrate_noalloc(p, **kw)
            b = ctx.integrate(p, **kw)
            _same(a, b)
            assert a["peak_regions"] <= cuhre.noalloc_capacity(p.f.dim)


def test_noalloc_rejects_what_outgrows_the_array():
    p = W.c2_product_peak8()  # needs ~2.5e6 regions at maxeval 1e9
    cap = cuhre.noalloc_capacity(8)
    assert 100_000 < cap < 1_000_000
    with cuhre.Context(0) as ctx:
        print("[noalloc] product_peak8d over capacity", flush=True)
        r = ctx.integrate_noalloc(p, check=False)
        assert r["rc"] == cuhre.ERR_CAPACITY, r
        assert "static workspace" in r["message"]
        assert r["fail"] == 1 and 0 < r["peak_regions"] <= cap
        # within the limit (a budget that stops below it) both agree again
        me = 300 * 401
        _same(ctx.integrate_noalloc(p, maxeval=me), ctx.integrate(p, maxeval=me))
        small = W.c1_gaussian3()
        _same(ctx.integrate_noalloc(small), ctx.integrate(small))


def test_noalloc_rejects_bad_arguments():
    with cuhre.Context(0) as ctx:
        p = W.c1_gaussian3()
        p.f.dim = 11
        p.lows = [0.0] * 11
        p.highs = [1.0] * 11
        r = ctx.integrate_noalloc(p, check=False)
        assert r["rc"] == cuhre.ERR_UNSUPPORTED
    assert cuhre.noalloc_capacity(1) == -1 and cuhre.noalloc_capacity(11) == -1
