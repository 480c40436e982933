"""GPU parity of the reference's LC_LT_t integrand (lc_lt_t.h:20-50, with
the float-table Interp2D of GPUQuadInterp2D.cu:49-100) against the oracle's
restatement, on the tables of the reference's header
(tests/golden/lc_lt_t_tables.json) and the reference driver's run
(cudaCuhre.cu:43-79: Volume lows (1, .1, 0), highs (2, .3, 1) in float,
epsabs 1e-12, epsrel 1e-7 and 1e-12).

Bars as for the other families: per region est / err within 1e-13 of
volume * sum |w f| (the GPU's exp/erfc are ROCm's, the oracle's glibc's),
the axis identical unless two fourth differences tie to rounding; whole
integrations with the oracle's region counts and iterations.
"""
import json
import os

import numpy as np
import pytest

import cuhre
import oracle as O
import workloads as W
from test_gpu_parity import assert_region_parity
from test_gpu_rule9 import assert_err_parity

pytestmark = pytest.mark.gpu

with open(os.path.join(os.path.dirname(__file__), "golden", "lc_lt_t_tables.json")) as fh:
    TABLES = json.load(fh)


def regions(rng, n):
    # inside the tables' range and beyond both ends of the lt nodes
    lo = np.stack([rng.uniform(0.5, 150.0, n), rng.uniform(0.08, 0.3, n), rng.uniform(-1.0, 2.5, n)], axis=1)
    wd = np.stack([rng.uniform(0.05, 20.0, n), rng.uniform(0.005, 0.05, n), rng.uniform(0.05, 1.0, n)], axis=1)
    return lo, wd


@pytest.mark.parametrize("lanes", [1, 32])
def test_lclt_region_values_match_oracle(ctx, lanes):
    f = W.lc_lt_t(TABLES)
    L, Wd = regions(np.random.default_rng(11 + lanes), 300)
    ctx.set(False, lanes)
    try:
        g = ctx.eval_regions(f, L, Wd, want_aux=True)
    finally:
        ctx.set(False, 0)
    o = O.region_eval(f, L, Wd)
    assert_region_parity(g, o, 3, ("lclt", lanes), np.prod(Wd, axis=1))


@pytest.mark.parametrize("lanes", [1, 32])
def test_lclt_rule9_region_values_match_oracle(ctx, lanes):
    f = W.lc_lt_t(TABLES)
    L, Wd = regions(np.random.default_rng(21 + lanes), 200)
    ctx.set(False, lanes)
    try:
        est, err, ax, ab = ctx.eval_regions(f, L, Wd, key=9)
    finally:
        ctx.set(False, 0)
    oest, oerr, oax, oab, odv, osums = O.region_eval9(f, L, Wd)
    scale = np.maximum(oab, 1e-300)
    assert np.all(np.abs(est - oest) <= 1e-13 * scale)
    assert_err_parity(err, oerr, osums, np.prod(Wd, axis=1), scale, 3, ("lclt9", lanes))


@pytest.mark.parametrize("key", [7, 9])
@pytest.mark.parametrize("epsrel", [1e-7, 1e-12])
def test_lclt_driver_run_matches_oracle(ctx, key, epsrel):
    p = W.lclt_driver(TABLES, epsrel)
    g = ctx.integrate(p, key=key)
    o = O.integrate_bfs(p, nthreads=8, key=key)
    assert g["rc"] == o["rc"], (g, o)
    assert g["nregions"] == o["nregions"] and g["iterations"] == o["iterations"], (g, o)
    assert abs(g["integral"] - o["integral"]) <= 1e-12 * abs(o["integral"])
    assert abs(g["error"] - o["error"]) <= 1e-6 * o["error"]


@pytest.mark.parametrize("epsrel", [1e-7, 1e-12])
def test_lclt_driver_two_phase_matches_oracle(ctx, epsrel):
    """The reference's own method on its own driver run."""
    p = W.lclt_driver(TABLES, epsrel)
    g = ctx.integrate_two_phase(p, key=9)
    o = O.integrate_two_phase(p, key=9)
    assert g["rc"] == o["rc"] and (g["fail"] == 0) == (o["fail"] == 0), (g, o)
    assert g["iterations"] == o["iterations"]
    assert abs(g["nregions"] - o["nregions"]) <= 5e-3 * o["nregions"]
    assert abs(g["integral"] - o["integral"]) <= max(g["error"], o["error"]) + 1e-12 * abs(o["integral"])


def test_lclt_table_validation(ctx):
    f = W.lc_lt_t(dict(TABLES, lt=list(reversed(TABLES["lt"]))))
    with pytest.raises(cuhre.CuhreError, match="increase"):
        ctx.integrate(W.Problem("bad", f, [1, .1, 0], [2, .3, 1], 1e-7, 0.0, 10**6))
    f = W.lc_lt_t(TABLES)
    f.dim = 4
    with pytest.raises(cuhre.CuhreError):
        ctx.integrate(W.Problem("bad", f, [1, .1, 0, 0], [2, .3, 1, 1], 1e-7, 0.0, 10**6))
