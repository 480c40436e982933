"""Pins the CPU oracle (the checker) before anything is compared against it.

Each check is chosen so that a plausible slip in the oracle -- a wrong
generator radius, weight, sign, index or scaling -- fails it:
  * rule: point counts, weight sums, exactness on every monomial up to the
    rule's degree and failure one degree above;
  * per-region values: SciPy's independent Genz-Malik implementation (golden
    fixture), exact rational integrals of polynomials;
  * fourth differences: the closed form D_i = |c_i| h_i^4 * 972/4900 for
    sum c_i x_i^4 and the widest-axis fallback for cubics
    (quad/GPUquad/GPUQuadSample.cu:113-125, 150-172);
  * closed forms of every family: adaptive quadrature (SciPy) in 1-3 D;
  * loop: convergence on the 3-D Gaussian workload against its erf form,
    budget stop, grid/bisection tiling.
"""
import itertools
import json
import math
import os
import zlib
from fractions import Fraction

import numpy as np
import pytest
from scipy import integrate

import oracle as O
import workloads as W

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def test_splitmix64_known_answers():
    # published first outputs of SplitMix64 seeded with 0
    r = W.SplitMix64(0)
    assert [r.next_u64() for _ in range(3)] == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4,
                                                0x06C45D188009454F]
    u = W.SplitMix64(7)
    vals = [u.uniform() for _ in range(1000)]
    assert min(vals) >= 0.0 and max(vals) < 1.0


@pytest.mark.parametrize("d", range(2, 11))
def test_rule_point_count_and_weights(d):
    g, w7, w5 = O.rule(d)
    assert len(w7) == (1 << d) + 2 * d * d + 2 * d + 1
    assert {3: 33, 5: 93, 6: 149, 7: 241, 8: 401}.get(d, len(w7)) == len(w7)
    assert abs(w7.sum() - 1.0) < 1e-14
    assert abs(w5.sum() - 1.0) < 1e-14
    assert np.all(w5[-(1 << d):] == 0.0)
    assert np.all(np.abs(g) <= 1.0)


def _monomial_exact(alpha):
    # mean of x^alpha over [-1,1]^d
    v = Fraction(1)
    for a in alpha:
        if a % 2:
            return Fraction(0)
        v *= Fraction(1, a + 1)
    return v


@pytest.mark.parametrize("d", [2, 3, 4, 5])
def test_rule_polynomial_degree(d):
    g, w7, w5 = O.rule(d)
    for deg in range(0, 9):
        for alpha in itertools.product(range(deg + 1), repeat=d):
            if sum(alpha) != deg:
                continue
            val7 = float(np.sum(w7 * np.prod(g ** np.array(alpha), axis=1)))
            val5 = float(np.sum(w5 * np.prod(g ** np.array(alpha), axis=1)))
            ex = float(_monomial_exact(alpha))
            if deg <= 7:
                assert abs(val7 - ex) < 1e-14, (alpha, val7, ex)
            if deg <= 5:
                assert abs(val5 - ex) < 1e-14, (alpha, val5, ex)
    # one degree higher is not integrated exactly (the rule is not over-fitted)
    a8 = (8,) + (0,) * (d - 1)
    assert abs(np.sum(w7 * g[:, 0] ** 8) - 1.0 / 9.0) > 1e-4
    a6 = (6,) + (0,) * (d - 1)
    assert abs(np.sum(w5 * g[:, 0] ** 6) - 1.0 / 7.0) > 1e-4
    del a8, a6


def test_oracle_matches_scipy_golden():
    g = json.load(open(os.path.join(GOLDEN, "gm_scipy_regions.json")))
    n = 0
    for c in g["cases"]:
        f = W.Integrand(c["family"], c["dim"], c["a"], c["w"], u=c["u"], exponent=c["exponent"])
        L = np.array([r["lower"] for r in c["regions"]])
        Wd = np.array([r["width"] for r in c["regions"]])
        est, err, ax, ab, i5, dv = O.region_eval(f, L, Wd)
        for k, r in enumerate(c["regions"]):
            assert abs(est[k] - r["est"]) <= 1e-13 * ab[k]
            assert abs(err[k] - r["err"]) <= 1e-13 * ab[k]
            n += 1
    assert n == 120


def _poly_exact_fraction(terms, lo, hi):
    tot = Fraction(0)
    for c, pw in terms:
        v = Fraction(c)
        for e, a, b in zip(pw, lo, hi):
            fa, fb = Fraction(a), Fraction(b)
            v *= (fb ** (e + 1) - fa ** (e + 1)) / (e + 1)
        tot += v
    return tot


def test_region_exact_on_polynomials():
    rng = W.SplitMix64(11)
    for d in (2, 3, 4, 6):
        for _ in range(4):
            terms = []
            for _t in range(5):
                pw = [0] * d
                deg = int(6 * rng.uniform())  # total degree <= 5
                for _k in range(deg):
                    pw[int(d * rng.uniform())] += 1
                terms.append((round(rng.uniform_range(-2, 2), 3), pw))
            f = W.polynomial(d, terms)
            lo = np.array([[round(rng.uniform(), 4) for _ in range(d)]])
            wd = np.array([[round(rng.uniform_range(0.05, 0.5), 4) for _ in range(d)]])
            est, err, ax, ab, i5, dv = O.region_eval(f, lo, wd)
            ex = float(_poly_exact_fraction(terms, lo[0], lo[0] + wd[0]))
            assert abs(est[0] - ex) <= 1e-14 * max(ab[0], 1e-300)
            assert err[0] <= 1e-14 * max(ab[0], 1e-300)
            assert abs(i5[0] - ex) <= 1e-14 * max(ab[0], 1e-300)


def test_fourth_difference_closed_form_and_axis():
    d = 5
    c = [0.3, -2.0, 1.1, 0.7, -0.2]
    terms = [(ci, [4 if k == i else 0 for k in range(d)]) for i, ci in enumerate(c)]
    f = W.polynomial(d, terms)
    lo = np.array([[0.1, 0.2, 0.3, 0.0, 0.5]])
    wd = np.array([[0.4, 0.3, 0.5, 0.2, 0.3]])
    est, err, ax, ab, i5, dv = O.region_eval(f, lo, wd)
    h = wd[0] / 2
    expect = np.abs(np.array(c)) * h ** 4 * 972.0 / 4900.0
    assert np.allclose(dv[0], expect, rtol=1e-9, atol=1e-15)
    assert ax[0] == int(np.argmax(expect))


def test_fourth_difference_zero_picks_widest_axis():
    d = 4
    f = W.polynomial(d, [(1.0, [1, 2, 0, 0]), (0.5, [0, 0, 3, 0]), (2.0, [0, 0, 0, 1])])
    lo = np.array([[0.0, 0.0, 0.0, 0.0]])
    wd = np.array([[0.1, 0.3, 0.2, 0.3]])  # first widest is axis 1
    est, err, ax, ab, i5, dv = O.region_eval(f, lo, wd)
    assert np.all(dv[0] <= 1e-15)
    assert ax[0] == 1


def _nquad(fn, lows, highs):
    v, e = integrate.nquad(lambda *x: fn(np.array(x)), list(zip(lows, highs)),
                           opts={"epsabs": 1e-13, "epsrel": 1e-12, "limit": 200})
    return v


@pytest.mark.parametrize("family", ["gaussian", "product_peak", "oscillatory", "corner", "continuous", "cosine"])
def test_closed_forms_against_quadrature(family):
    rng = W.SplitMix64(zlib.crc32(family.encode()))  # fixed per family (str hash is salted per process)
    d = 2
    lows = [rng.uniform_range(-0.2, 0.2) for _ in range(d)]
    highs = [lows[i] + rng.uniform_range(0.5, 1.2) for i in range(d)]
    a = [rng.uniform_range(0.5, 4) for _ in range(d)]
    w = [rng.uniform_range(0.1, 0.9) for _ in range(d)]
    f = {
        "gaussian": W.gaussian(a, w),
        "product_peak": W.product_peak(a, w),
        "oscillatory": W.oscillatory(a, rng.uniform()),
        "corner": W.corner_peak(a, 6),
        "continuous": W.continuous(a, w),
        "cosine": W.cosine_sum(d),
    }[family]
    if family == "corner":
        lows = [0.0, 0.1]
    hi, lo = O.exact(f, lows, highs)
    num = _nquad(lambda x: O.f_values(f, x[None, :])[0], lows, highs)
    assert abs(hi - num) <= 1e-10 * max(1.0, abs(num))


def test_corner_peak_closed_form_3d():
    f = W.corner_peak([0.7, 2.5, 4.0], 8)
    hi, lo = O.exact(f, [0, 0, 0], [1, 1, 1])
    num = integrate.tplquad(lambda z, y, x: (1 + 0.7 * x + 2.5 * y + 4.0 * z) ** -8, 0, 1, 0, 1, 0, 1,
                            epsabs=1e-15, epsrel=1e-12)[0]
    assert abs(hi - num) <= 1e-10 * abs(num)


def test_table6_closed_form_small_grid():
    # 3x3 table: bilinear interpolant integrated by quadrature vs trapezoid form
    T = np.array([[0.2, -0.5, 1.0], [0.3, 0.9, -0.1], [1.5, 0.0, 0.4]])
    f = W.table6(T)
    lows = [0, 0, 0.1, 0, 0, 0]
    highs = [1, 1, 1, 1, 1, 1]
    hi, lo = O.exact(f, lows, highs)

    def B(x1, x2):
        x = np.zeros(6)
        x[0], x[1], x[2], x[3], x[4], x[5] = x1, x2, 1.0, 0.5, 0.0, 0.0
        return O.f_values(f, x[None, :])[0]

    bint = integrate.dblquad(lambda y, x: B(x, y), 0, 1, 0, 1, epsabs=1e-13, epsrel=1e-12)[0]
    i3 = 2 * (1 / math.sqrt(0.1) - 1)
    i4 = integrate.quad(lambda t: math.exp(-((t - 0.5) ** 2) / 0.02), 0, 1, epsabs=1e-15)[0]
    assert abs(hi - bint * i3 * i4 * 1.25) <= 1e-9 * abs(hi)


def test_initial_grid_and_split_tile_the_box():
    for d in (3, 5):
        lows = [0.0] * d
        highs = [1.0] * (d - 1) + [2.0]
        L, Wd = O.initial_grid(d, lows, highs)
        k = 4 if d <= 4 else 2
        assert L.shape == (k ** d, d)
        assert abs(np.prod(Wd, axis=1).sum() - 2.0) < 1e-14
        # distinct cells
        assert len({tuple(r) for r in np.round(L, 12)}) == k ** d
        act = np.array([i % 3 != 0 for i in range(len(L))], dtype=np.int32)
        ax = np.array([i % d for i in range(len(L))], dtype=np.int32)
        cL, cW = O.split(L, Wd, act, ax)
        idx = np.nonzero(act)[0]
        assert cL.shape[0] == 2 * len(idx)
        for j, r in enumerate(idx):
            a = ax[r]
            for c in (0, 1):
                assert np.all(cW[2 * j + c, np.arange(d) != a] == Wd[r, np.arange(d) != a])
                assert cW[2 * j + c, a] == Wd[r, a] / 2
            assert cL[2 * j, a] == L[r, a]
            assert cL[2 * j + 1, a] == L[r, a] + Wd[r, a] / 2


def test_bfs_gaussian3d_converges_to_erf_form():
    p = W.c1_gaussian3()
    r = O.integrate_bfs(p)
    ex, _ = O.exact(p.f, p.lows, p.highs)
    assert r["rc"] == 0 and r["fail"] == 0
    assert r["error"] <= p.epsrel * abs(r["integral"])
    assert abs(r["integral"] - ex) <= p.epsrel * abs(ex)
    assert r["neval"] == r["nregions"] * 33
    assert r["neval"] <= p.maxeval


def test_bfs_budget_stop_and_heap():
    p = W.c2_product_peak8()
    r = O.integrate_bfs(p, nthreads=4, maxeval=2_000_000)
    assert r["rc"] == 1 and r["fail"] == 1 and r["neval"] <= 2_000_000
    p1 = W.c1_gaussian3()
    h = O.integrate_heap(p1)
    ex, _ = O.exact(p1.f, p1.lows, p1.highs)
    assert h["rc"] == 0 and abs(h["integral"] - ex) <= p1.epsrel * abs(ex)


def test_bfs_nonfinite_is_an_error():
    # x^-1.5 on a box touching 0 gives inf at a corner point? No: GM nodes are
    # interior.  Use a power that overflows instead: (1 + a.x)^-k with a < 0.
    f = W.corner_peak([-1.0, -1.0], 700)
    p = W.Problem("bad", f, [0.0, 0.0], [1.0, 1.0], 1e-6, 0.0, 10**6)
    r = O.integrate_bfs(p)
    assert r["rc"] == -2


# ---------------------------------------------------------------- degree 9
# The reference's hard-wired rule (KEY 9, quad/GPUquad/GPUQuadRule.cu:30-226):
# pinned by its point count, normalisation and polynomial degrees, which a
# wrong radius or weight breaks.

@pytest.mark.parametrize("d", range(2, 11))
def test_rule9_counts_and_normalisation(d):
    R = O.rule9(d)
    assert R["npts"] == R["count"].sum()
    assert {3: 77, 5: 273, 7: 717, 8: 1105}.get(d, R["npts"]) == R["npts"]
    assert abs((R["count"] * R["w"][:, 0]).sum() - 1.0) < 1e-13
    for q in range(1, 5):  # null rules integrate constants to zero
        assert abs((R["count"] * R["w"][:, q]).sum()) < 1e-13


def _r9_monomial_dev(d, deg):
    """Largest |rule - exact| over the monomials of total degree deg on
    [-1,1]^d, for rule 0 and each null rule (whose exact value is 0)."""
    dev = np.zeros(5)
    for alpha in itertools.product(range(deg + 1), repeat=d):
        if sum(alpha) != deg:
            continue
        f = W.polynomial(d, [(1.0, list(alpha))])
        est, err, ax, ab, dv, sums = O.region_eval9(f, np.full((1, d), -1.0), np.full((1, d), 2.0))
        ex = float(_monomial_exact(alpha)) * 2 ** d
        dev[0] = max(dev[0], abs(est[0] - ex))
        for q in range(1, 5):
            dev[q] = max(dev[q], abs(sums[0][q]))
    return dev


@pytest.mark.parametrize("d", [2, 3, 4])
def test_rule9_polynomial_degrees(d):
    # rule 0 has degree 9; the null rules 1..4 vanish up to degrees 7, 7, 5, 3
    exact_to = [9, 7, 7, 5, 3]
    devs = {deg: _r9_monomial_dev(d, deg) for deg in range(0, 11)}
    for q in range(5):
        for deg in range(0, exact_to[q] + 1):
            assert devs[deg][q] < 1e-13, (q, deg, devs[deg][q])
        # and not one even degree above (odd monomials vanish by symmetry)
        above = exact_to[q] + 1 if (exact_to[q] + 1) % 2 == 0 else exact_to[q] + 2
        assert devs[above][q] > 1e-4, (q, above, devs[above][q])


def test_rule9_fourth_difference_closed_form():
    # f = sum c_i x_i^4: the quadratic parts cancel by the choice of ratio,
    # leaving D_i = 2 |c_i| h_i^4 t2^2 (t1^2 - t2^2) (t in half-width units)
    d = 5
    c = [0.3, -2.0, 1.1, 0.7, -0.2]
    f = W.polynomial(d, [(ci, [4 if k == i else 0 for k in range(d)]) for i, ci in enumerate(c)])
    lo = np.array([[0.1, 0.2, 0.3, 0.0, 0.5]])
    wd = np.array([[0.4, 0.3, 0.5, 0.2, 0.3]])
    est, err, ax, ab, dv, sums = O.region_eval9(f, lo, wd)
    t = O.rule9(d)["t"]
    h = wd[0] / 2
    expect = 2 * np.abs(c) * h ** 4 * t[1] ** 2 * (t[0] ** 2 - t[1] ** 2)
    assert np.allclose(dv[0], expect, rtol=1e-9, atol=1e-15)
    assert ax[0] == int(np.argmax(expect))


def test_rule9_error_vanishes_for_cubics_and_axis_falls_back():
    d = 4
    f = W.polynomial(d, [(1.0, [1, 2, 0, 0]), (0.5, [0, 0, 3, 0]), (2.0, [0, 0, 0, 1])])
    lo = np.array([[0.0, 0.0, 0.0, 0.0]])
    wd = np.array([[0.1, 0.3, 0.2, 0.3]])
    est, err, ax, ab, dv, sums = O.region_eval9(f, lo, wd)
    ex = float(_poly_exact_fraction([(1.0, [1, 2, 0, 0]), (0.5, [0, 0, 3, 0]), (2.0, [0, 0, 0, 1])],
                                    lo[0], lo[0] + wd[0]))
    assert abs(est[0] - ex) <= 1e-14 * ab[0]
    assert err[0] <= 1e-15 * ab[0]
    assert np.all(dv[0] <= 1e-15 * ab[0] / np.prod(wd[0]))  # zero up to rounding
    # every D_i exactly 0 (the zero function): the first widest axis
    z = W.polynomial(d, [(0.0, [0, 0, 0, 0])])
    est, err, ax, ab, dv, sums = O.region_eval9(z, lo, wd)
    assert np.all(dv[0] == 0.0) and ax[0] == 1


def test_rule9_loop_gaussian():
    p = W.c1_gaussian3()
    r9 = O.integrate_bfs(p, key=9)
    r7 = O.integrate_bfs(p, key=7)
    ex, _ = O.exact(p.f, p.lows, p.highs)
    assert r9["rc"] == 0
    assert r9["error"] <= p.epsrel * abs(r9["integral"])
    assert abs(r9["integral"] - ex) <= r9["error"]
    assert r9["neval"] == r9["nregions"] * 77
    assert r7["rc"] == 0 and r7["neval"] == r7["nregions"] * 33


# ------------------------------------------------------- two-phase method
# The reference's Phase 1 + per-region Phase 2 (GPUquad.h:611-727,
# GPUQuadPhases.cu:422-536), restated in the oracle.

def test_two_phase_equals_breadth_first_when_phase1_converges():
    p = W.c1_gaussian3()
    bfs = O.integrate_bfs(p, key=9)
    tp = O.integrate_two_phase(p, key=9, phase1_max=1 << 20)
    assert bfs["rc"] == tp["rc"] == 0
    assert tp["nregions"] == bfs["nregions"] and tp["iterations"] == bfs["iterations"]
    assert tp["integral"] == bfs["integral"]


def test_two_phase_region_accounting():
    # phase1_max = 1: the initial 64 regions are evaluated, their active
    # ones split, and each child gets exactly one forced bisection when
    # pool_max = 1: it counts 2 regions (its pool: the two halves, the
    # reference's nregions) and 1 + 2 evaluations (GPUKernelquad.cu:883-888)
    p = W.c1_gaussian3()
    tp = O.integrate_two_phase(p, key=7, phase1_max=1, pool_max=1)
    assert tp["iterations"] == 1
    n2 = tp["nregions"] - 64
    assert n2 % 2 == 0 and n2 > 0
    assert tp["neval"] == 64 * 33 + (n2 // 2) * 3 * 33


@pytest.mark.parametrize("key", [7, 9])
def test_two_phase_phase2_converges(key):
    p = W.c1_gaussian3()
    ex, _ = O.exact(p.f, p.lows, p.highs)
    tp = O.integrate_two_phase(p, key=key, phase1_max=256)
    assert tp["rc"] == 0 and tp["fail"] == 0
    assert tp["error"] <= p.epsrel * abs(tp["integral"])
    assert abs(tp["integral"] - ex) <= tp["error"]


# ------------------------------------------------------------------ LC_LT_t
# The reference's only production integrand (lc_lt_t.h:20-50 with the
# Interp2D of GPUQuadInterp2D.cu:49-100).  Its data tables come from the
# reference's header through tests/golden/make_lclt_tables.py, which checks
# them against the FP64 twin tables of integrands/lc_lt_t.cc.

LCLT_PATH = os.path.join(os.path.dirname(__file__), "golden", "lc_lt_t_tables.json")


def lclt_tables():
    with open(LCLT_PATH) as fh:
        return json.load(fh)


def test_lclt_fixture_tables():
    t = lclt_tables()
    nx, ny = len(t["lt"]), len(t["zt"])
    assert (nx, ny) == (22, 5)
    assert np.all(np.diff(t["lt"]) > 0) and np.all(np.diff(t["zt"]) > 0)
    for k in ("tau", "mu", "sigma", "fmsk", "fprj"):
        a = np.asarray(t[k])
        assert a.size == nx * ny
        assert np.array_equal(a, a.astype(np.float32).astype(np.float64)), k  # float values
    assert min(t["sigma"]) > 0 and min(t["tau"]) > 0
    assert 0 <= min(t["fmsk"]) and max(t["fmsk"]) <= 1 and 0 <= min(t["fprj"]) and max(t["fprj"]) <= 1


def _bilinear_tables(t, mu_c, sg_c):
    """fmsk = fprj = 0 and mu, sigma bilinear in (lt, zt): the integrand is
    the normal density of lc, and interpolation of a bilinear function is
    exact (up to the float rounding of nodes, values and result)."""
    lt = np.asarray(t["lt"], dtype=np.float32).astype(np.float64)
    zt = np.asarray(t["zt"], dtype=np.float32).astype(np.float64)
    L, Z = np.meshgrid(lt, zt)  # [j, i]: row j = zt_j, column i = lt_i -> field[j*nx + i]
    bil = lambda c: c[0] + c[1] * L + c[2] * Z + c[3] * L * Z  # noqa: E731
    f32 = lambda a: a.astype(np.float32).astype(np.float64).ravel().tolist()  # noqa: E731
    nx, ny = lt.size, zt.size
    return dict(t, mu=f32(bil(mu_c)), sigma=f32(bil(sg_c)), fmsk=[0.0] * (nx * ny), fprj=[0.0] * (nx * ny))


def test_lclt_interpolation_reduces_to_normal_density():
    from scipy.stats import norm
    t = lclt_tables()
    mu_c, sg_c = (0.3, 0.01, -0.8, 0.05), (0.5, 0.004, 0.6, 0.02)
    f = W.lc_lt_t(_bilinear_tables(t, mu_c, sg_c))
    rng = np.random.default_rng(5)
    # inside the grid, below its first lt node and above its last
    lt = np.concatenate([rng.uniform(1, 160, 40), rng.uniform(0.2, 1, 5), rng.uniform(160, 200, 5)])
    zt = np.concatenate([rng.uniform(0.1, 0.3, 40), rng.uniform(0.1, 0.3, 10)])
    lc = rng.uniform(-1, 3, lt.size)
    got = O.f_values(f, np.stack([lt, zt, lc], axis=1))
    mu = mu_c[0] + mu_c[1] * lt + mu_c[2] * zt + mu_c[3] * lt * zt
    sg = sg_c[0] + sg_c[1] * lt + sg_c[2] * zt + sg_c[3] * lt * zt
    want = norm.pdf(lc, mu, sg)
    # mu and sigma pass through float: relative error a few 1e-7, times |dlog f|
    tol = 2e-6 * (1 + np.abs((lc - mu) / sg) ** 2) * want + 1e-12
    assert np.all(np.abs(got - want) <= tol), np.max(np.abs(got - want) / tol)


def test_lclt_interpolation_axes_are_not_transposed():
    # a field that depends on lt only, then on zt only: swapping the roles
    # of the two axes (or of q12 and q21) changes the value
    from scipy.stats import norm
    t = lclt_tables()
    for mu_c in ((0.0, 0.02, 0.0, 0.0), (0.0, 0.0, 3.0, 0.0)):
        f = W.lc_lt_t(_bilinear_tables(t, mu_c, (1.0, 0.0, 0.0, 0.0)))
        X = np.array([[50.0, 0.12, 0.4], [7.5, 0.27, 1.1]])
        mu = mu_c[1] * X[:, 0] + mu_c[2] * X[:, 1]
        assert np.allclose(O.f_values(f, X), norm.pdf(X[:, 2], mu, 1.0), rtol=1e-6)


@pytest.mark.parametrize("lt,zt", [(1.3, 0.13), (1.9, 0.27), (4.0, 0.2), (35.0, 0.11), (120.0, 0.29)])
def test_lclt_density_integrates_to_one(lt, zt):
    """For fixed (lt, zt) the integrand is a density in lc: the Gaussian term
    carries (1-fmsk)(1-fprj), the exponentially modified Gaussian
    (1-fmsk) fprj, and the fmsk terms (a uniform of width lt convolved
    with the Gaussian, minus the lt-shifted EMG over tau lt plus the EMG
    over tau lt) carry fmsk: the total is 1.  A dropped term or a wrong
    sign moves it by fmsk, fprj or fmsk*fprj/(tau*lt) (each at least 2e-4
    at these points, with the tables' interpolated values).  The reference forms mu + tau sigma^2 in float, which
    shifts the EMG terms' argument by a float rounding: the total is 1
    within ~2e-7, not to double precision.  (Far below the data, lc < -200,
    exp overflows against an erfc underflow and the formula gives NaN, in
    the reference as here; the integral starts at -60, where the density
    is below 1e-40.)"""
    t = lclt_tables()
    f = W.lc_lt_t(t)
    g = lambda c: O.f_values(f, np.array([[lt, zt, c]]))[0]  # noqa: E731
    edges = np.concatenate([np.linspace(-60.0, lt + 100.0, 200), np.linspace(lt + 100.0, 800.0, 50)[1:]])
    total = sum(integrate.quad(g, a, b, limit=200, epsabs=1e-15, epsrel=1e-12)[0]
                for a, b in zip(edges[:-1], edges[1:]))
    assert abs(total - 1.0) <= 1e-6, total
    # the terms a mistake would drop are far above that tolerance here
    fz = lambda k: O.f_values(W.lc_lt_t(dict(t, **{k: [0.0] * len(t[k])})),  # noqa: E731
                              np.array([[lt, zt, c] for c in (0.5, lt / 2)]))
    assert np.max(np.abs(fz("fmsk") - O.f_values(f, np.array([[lt, zt, 0.5], [lt, zt, lt / 2]])))) > 1e-5


# The degree-9 rule's error (GPUQuadSample.cu:195-216; constants
# GPUQuadRule.cu:209-225): pinned by what the construction means, not by
# restating it.

@pytest.mark.parametrize("d", range(2, 11))
def test_rule9_null_combinations_are_normalised_null_rules(d):
    """For each set t and null-rule pair (q, q+1) the combination
    N_{q+1} + scale[t][q] N_q removes set t (unless set t has no weight in
    N_q, where the reference uses scale 100), still integrates constants to
    zero, and norm[t][q] makes its weighted 1-norm exactly 1."""
    R = O.rule9(d)
    w, cnt, sc, nm = R["w"], R["count"], R["scale"], R["norm"]
    for q in range(1, 4):
        for t in range(9):
            comb = w[:, q + 1] + sc[t, q] * w[:, q]
            if w[t, q] != 0.0:
                assert abs(comb[t]) <= 1e-15 * (abs(w[t, q + 1]) + 1.0), (d, q, t)
            else:
                assert sc[t, q] == 100.0
            assert abs((cnt * comb).sum()) <= 1e-12 * (cnt * np.abs(comb)).sum()
            assert (cnt * np.abs(comb)).sum() * nm[t, q] == pytest.approx(1.0, rel=1e-14)


@pytest.mark.parametrize("d", [3, 8])
def test_rule9_error_branch_on_chosen_sums(d):
    """errcoeff {5, 1, 5}: the error is e1 when e1, e2, e3 fall off by 5x
    each, else 5 max(e1, e2, e3).  Sums that leave only the last null rule
    give e1 = e2 = 0 < e3 and an error of exactly 0; sums that leave only
    the first give e2 = e3 = 0 < e1 and an error of 5 e1 (a swapped
    coefficient or branch fails one of the two)."""
    R = O.rule9(d)
    err, e = O.r9_error(d, [1.0, 0.0, 0.0, 0.0, 1.0])
    assert e[0] == 0.0 and e[1] == 0.0 and e[2] == pytest.approx(R["norm"][:, 3].max())
    assert err == 0.0
    err, e = O.r9_error(d, [1.0, 1.0, 0.0, 0.0, 0.0])
    assert e[1] == 0.0 and e[2] == 0.0
    assert e[0] == pytest.approx((np.abs(R["scale"][:, 1]) * R["norm"][:, 1]).max())
    assert err == pytest.approx(5.0 * e[0])
    # geometric fall-off taken from real sums: the error is homogeneous of
    # degree 1 in the sums
    err1, _ = O.r9_error(d, [0.3, 1e-3, -2e-3, 5e-2, 0.7])
    err2, _ = O.r9_error(d, [0.6, 2e-3, -4e-3, 1e-1, 1.4])
    assert err2 == pytest.approx(2.0 * err1, rel=1e-14)


# Phase 2 (GPUQuadPhases.cu:422-536): the Cuba refinement of the children's
# errors from the parent difference (GPUQuadPhases.cu:498-515), checked by
# hand on one and two bisections.

def _refine(pe, l, r, el, er):
    diff = abs(0.25 * (l + r - pe))
    e = el + er
    if e > 0:
        c = 1 + 2 * diff / e
        el, er = el * c, er * c
    return el + diff, er + diff


def _eval(key, f, lo, wd):
    if key == 9:
        est, err, ax, ab, dv, sums = O.region_eval9(f, lo[None, :], wd[None, :])
    else:
        est, err, ax, ab, i5, dv = O.region_eval(f, lo[None, :], wd[None, :])
    return est[0], err[0], int(ax[0])


def _bisect(lo, wd, k):
    hw = 0.5 * wd[k]
    lo0, wd0 = lo.copy(), wd.copy()
    wd0[k] = hw
    lo1, wd1 = lo.copy(), wd.copy()
    wd1[k] = hw
    lo1[k] = lo[k] + hw
    return (lo0, wd0), (lo1, wd1)


@pytest.mark.parametrize("key", [7, 9])
def test_phase2_one_bisection_by_hand(key):
    d = 4
    f = W.product_peak([6.0, 3.0, 5.0, 2.0], [0.3, 0.6, 0.45, 0.2])
    lo = np.array([0.25, 0.5, 0.25, 0.0])
    wd = np.array([0.25, 0.25, 0.25, 0.5])
    pe, pr, k = _eval(key, f, lo, wd)
    (l0, w0), (l1, w1) = _bisect(lo, wd, k)
    a, ea, _ = _eval(key, f, l0, w0)
    b, eb, _ = _eval(key, f, l1, w1)
    ea2, eb2 = _refine(pe, a, b, ea, eb)
    res, err, cnt, act, zero = O.phase2_region(f, lo, wd, key=key, pool_max=1)
    assert cnt == 2 and zero == 0
    assert res == pe + (a + b - pe)
    assert err == pr + (ea2 + eb2 - pr)
    # the refinement only grows errors, by at least the parent difference
    assert ea2 >= ea + abs(0.25 * (a + b - pe)) and eb2 >= eb


@pytest.mark.parametrize("key", [7, 9])
def test_phase2_two_bisections_take_the_larger_error(key):
    d = 3
    f = W.gaussian([7.0, 4.0, 9.0], [0.31, 0.62, 0.17])
    lo = np.array([0.0, 0.5, 0.0])
    wd = np.array([0.5, 0.5, 0.5])
    pe, pr, k = _eval(key, f, lo, wd)
    (l0, w0), (l1, w1) = _bisect(lo, wd, k)
    a, ea, ka = _eval(key, f, l0, w0)
    b, eb, kb = _eval(key, f, l1, w1)
    ea, eb = _refine(pe, a, b, ea, eb)
    RES, ERR = pe + (a + b - pe), pr + (ea + eb - pr)
    # second bisection: the child with the larger refined error (the first
    # created on a tie)
    (lp, wp, ep, rp, kp) = (l0, w0, a, ea, ka) if ea >= eb else (l1, w1, b, eb, kb)
    (m0, v0), (m1, v1) = _bisect(lp, wp, kp)
    c0, ec0, _ = _eval(key, f, m0, v0)
    c1, ec1, _ = _eval(key, f, m1, v1)
    ec0, ec1 = _refine(ep, c0, c1, ec0, ec1)
    RES += c0 + c1 - ep
    ERR += ec0 + ec1 - rp
    res, err, cnt, act, zero = O.phase2_region(f, lo, wd, key=key, epsrel=1e-12, pool_max=2)
    assert cnt == 3
    assert res == RES and err == ERR
