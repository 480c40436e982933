"""ctypes loader for the CPU oracle (test infrastructure only).

Only tests/, ``__graft_entry__.smoke()`` and bench.py's CPU-baseline leg may
import this module; the product path never does.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "libcuhre_oracle.so")
sys.path.insert(0, os.path.join(os.path.dirname(HERE), "old-gpuintegration_amd"))
import workloads as W  # noqa: E402

MAXDIM = 16
MAXTERMS = 8


class OrIntegrand(C.Structure):
    _fields_ = [
        ("family", C.c_int32),
        ("dim", C.c_int32),
        ("a", C.c_double * MAXDIM),
        ("w", C.c_double * MAXDIM),
        ("u", C.c_double),
        ("exponent", C.c_int32),
        ("nterms", C.c_int32),
        ("coef", C.c_double * MAXTERMS),
        ("powers", (C.c_int32 * MAXDIM) * MAXTERMS),
        ("table", C.POINTER(C.c_double)),
        ("table_nx", C.c_int32),
        ("table_ny", C.c_int32),
    ]


class OrResult(C.Structure):
    _fields_ = [
        ("integral", C.c_double),
        ("error", C.c_double),
        ("bound", C.c_double),
        ("nregions", C.c_int64),
        ("neval", C.c_int64),
        ("fail", C.c_int32),
        ("iterations", C.c_int32),
        ("peak_regions", C.c_int64),
        ("nfinished", C.c_int64),
    ]

    def asdict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class OrCheck(C.Structure):
    _fields_ = [
        ("istar_hi", C.c_double),
        ("istar_lo", C.c_double),
        ("irule_hi", C.c_double),
        ("irule_lo", C.c_double),
        ("abs_total", C.c_double),
        ("aest_total", C.c_double),
        ("ncontrib", C.c_int64),
        ("have_rule", C.c_int32),
        ("pad", C.c_int32),
    ]

    def asdict(self):
        return {k: getattr(self, k) for k, _ in self._fields_ if k != "pad"}


class OrRule(C.Structure):
    _fields_ = [
        ("dim", C.c_int),
        ("npts", C.c_int),
        ("g", C.POINTER(C.c_double)),
        ("w7", C.POINTER(C.c_double)),
        ("w5", C.POINTER(C.c_double)),
    ]


class OrRule9(C.Structure):
    _fields_ = [
        ("dim", C.c_int),
        ("npts", C.c_int),
        ("t", C.c_double * 5),
        ("w", (C.c_double * 5) * 9),
        ("scale", (C.c_double * 5) * 9),
        ("norm", (C.c_double * 5) * 9),
        ("count", C.c_int * 9),
    ]


_lib = None


def build():
    subprocess.check_call(["make", "-s", "-C", HERE])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = C.CDLL(LIB)
        dp = C.POINTER(C.c_double)
        ip = C.POINTER(C.c_int32)
        L.or_rule_init.argtypes = [C.POINTER(OrRule), C.c_int]
        L.or_rule_free.argtypes = [C.POINTER(OrRule)]
        L.or_f.argtypes = [C.POINTER(OrIntegrand), dp]
        L.or_f.restype = C.c_double
        L.or_region_eval.argtypes = [C.POINTER(OrRule), C.POINTER(OrIntegrand), dp, dp,
                                     dp, dp, C.POINTER(C.c_int), dp, dp, dp]
        L.or_integrate_bfs.argtypes = [C.POINTER(OrIntegrand), dp, dp, C.c_double, C.c_double,
                                       C.c_int64, C.c_int64, C.c_int, C.POINTER(OrResult)]
        L.or_rule9_init.argtypes = [C.POINTER(OrRule9), C.c_int]
        L.or_r9_npoints.argtypes = [C.c_int]
        L.or_region_eval9.argtypes = [C.POINTER(OrRule9), C.POINTER(OrIntegrand), dp, dp,
                                      dp, dp, C.POINTER(C.c_int), dp, dp, dp]
        L.or_integrate_bfs_key.argtypes = [C.POINTER(OrIntegrand), dp, dp, C.c_double, C.c_double,
                                           C.c_int64, C.c_int64, C.c_int, C.c_int, C.POINTER(OrResult)]
        L.or_integrate_bfs_check.argtypes = [C.POINTER(OrIntegrand), dp, dp, C.c_double, C.c_double,
                                             C.c_int64, C.c_int64, C.c_int, C.c_int, C.c_int,
                                             C.POINTER(OrResult), C.POINTER(OrCheck)]
        L.or_integrate_bfs_allot.argtypes = [C.POINTER(OrIntegrand), dp, dp, C.c_double, C.c_double,
                                             C.c_int64, C.c_int64, C.c_int, C.c_double, C.POINTER(OrResult)]
        L.or_integrate_bfs_normalized.argtypes = [C.POINTER(OrIntegrand), dp, dp, C.c_double, C.c_double,
                                                  C.c_int64, C.c_int64, C.c_int, C.c_double, C.POINTER(OrResult)]
        L.or_r9_error.argtypes = [C.POINTER(OrRule9), dp, dp]
        L.or_r9_error.restype = C.c_double
        L.or_phase2_region.argtypes = [C.c_int, C.POINTER(OrIntegrand), dp, dp, C.c_double, C.c_double, C.c_int, dp]
        L.or_fp64_bound.argtypes = [C.c_int, C.c_int, C.c_int64, C.c_double, C.c_double, C.c_double]
        L.or_fp64_bound.restype = C.c_double
        L.or_region_check.argtypes = [C.c_int, C.POINTER(OrIntegrand), dp, dp, dp]
        L.or_fq_residual.argtypes = [C.POINTER(OrIntegrand), dp, dp]
        L.or_fq_residual.restype = C.c_double
        L.or_integrate_two_phase.argtypes = [C.POINTER(OrIntegrand), dp, dp, C.c_double, C.c_double, C.c_int,
                                             C.c_int64, C.c_int, C.c_int, C.POINTER(OrResult)]
        L.or_integrate_heap.argtypes = [C.POINTER(OrIntegrand), dp, dp, C.c_double, C.c_double,
                                        C.c_int64, C.POINTER(OrResult)]
        L.or_exact.argtypes = [C.POINTER(OrIntegrand), dp, dp, dp, dp]
        L.or_split.argtypes = [C.c_int, C.c_int64, dp, dp, ip, ip, dp, dp]
        L.or_split.restype = C.c_int64
        L.or_initial_grid.argtypes = [C.c_int, dp, dp, dp, dp]
        L.or_initial_grid.restype = C.c_int64
        _lib = L
    return _lib


def _dp(arr):
    return arr.ctypes.data_as(C.POINTER(C.c_double))


def to_struct(f: W.Integrand):
    s = OrIntegrand()
    s.family = f.family
    s.dim = f.dim
    for i, v in enumerate(f.a):
        s.a[i] = v
    for i, v in enumerate(f.w):
        s.w[i] = v
    s.u = f.u
    s.exponent = f.exponent
    s.nterms = len(f.coef)
    for t, (c, pw) in enumerate(zip(f.coef, f.powers)):
        s.coef[t] = c
        for i, e in enumerate(pw):
            s.powers[t][i] = e
    keep = None
    if f.table is not None:
        keep = np.ascontiguousarray(f.table, dtype=np.float64)
        s.table = _dp(keep)
        s.table_nx, s.table_ny = f.table_shape if f.table_shape is not None else keep.shape
    return s, keep


def rule(d):
    """Return (g[p,d], w7[p], w5[p]) of the oracle's GM7/5 rule."""
    r = OrRule()
    rc = lib().or_rule_init(C.byref(r), d)
    assert rc == 0
    p = r.npts
    g = np.ctypeslib.as_array(r.g, shape=(p * d,)).reshape(p, d).copy()
    w7 = np.ctypeslib.as_array(r.w7, shape=(p,)).copy()
    w5 = np.ctypeslib.as_array(r.w5, shape=(p,)).copy()
    lib().or_rule_free(C.byref(r))
    return g, w7, w5


def f_values(f: W.Integrand, X):
    X = np.ascontiguousarray(X, dtype=np.float64)
    s, keep = to_struct(f)
    out = np.empty(X.shape[0])
    for i in range(X.shape[0]):
        out[i] = lib().or_f(C.byref(s), _dp(X[i]))
    return out


def region_eval(f: W.Integrand, lower, width):
    """Per-region (est, err, axis, abssum, i5, dvec) for arrays lower/width [n, d]."""
    lower = np.ascontiguousarray(lower, dtype=np.float64)
    width = np.ascontiguousarray(width, dtype=np.float64)
    n, d = lower.shape
    r = OrRule()
    assert lib().or_rule_init(C.byref(r), d) == 0
    s, keep = to_struct(f)
    est = np.empty(n); err = np.empty(n); ab = np.empty(n); i5 = np.empty(n)
    ax = np.empty(n, dtype=np.int32)
    dv = np.empty((n, d))
    e = C.c_double(); er = C.c_double(); a = C.c_int(); abv = C.c_double(); i5v = C.c_double()
    for i in range(n):
        lib().or_region_eval(C.byref(r), C.byref(s), _dp(lower[i]), _dp(width[i]), C.byref(e),
                             C.byref(er), C.byref(a), C.byref(abv), C.byref(i5v), _dp(dv[i]))
        est[i], err[i], ax[i], ab[i], i5[i] = e.value, er.value, a.value, abv.value, i5v.value
    lib().or_rule_free(C.byref(r))
    return est, err, ax, ab, i5, dv


def rule9(d):
    """The degree-9 rule's tables: dict of t[5], w[9][5], scale, norm, count[9], npts."""
    r = OrRule9()
    assert lib().or_rule9_init(C.byref(r), d) == 0
    return {"t": np.array(r.t[:]), "w": np.array([r.w[i][:] for i in range(9)]),
            "scale": np.array([r.scale[i][:] for i in range(9)]),
            "norm": np.array([r.norm[i][:] for i in range(9)]),
            "count": np.array(r.count[:]), "npts": r.npts}


def region_eval9(f: W.Integrand, lower, width):
    """Per-region (est, err, axis, abssum, dvec, sums[5]) under the degree-9 rule."""
    lower = np.ascontiguousarray(lower, dtype=np.float64)
    width = np.ascontiguousarray(width, dtype=np.float64)
    n, d = lower.shape
    r = OrRule9()
    assert lib().or_rule9_init(C.byref(r), d) == 0
    s, keep = to_struct(f)
    est = np.empty(n); err = np.empty(n); ab = np.empty(n)
    ax = np.empty(n, dtype=np.int32)
    dv = np.empty((n, d)); sums = np.empty((n, 5))
    e = C.c_double(); er = C.c_double(); a = C.c_int(); abv = C.c_double()
    for i in range(n):
        lib().or_region_eval9(C.byref(r), C.byref(s), _dp(lower[i]), _dp(width[i]), C.byref(e),
                              C.byref(er), C.byref(a), C.byref(abv), _dp(dv[i]), _dp(sums[i]))
        est[i], err[i], ax[i], ab[i] = e.value, er.value, a.value, abv.value
    return est, err, ax, ab, dv, sums


def npoints(d, key=7):
    return lib().or_r9_npoints(d) if key == 9 else W.gm_npoints(d)


def _default_cap(maxeval, d, key):
    """List capacity of the oracle's loop: what maxeval allows, at most 2^25
    regions (the oracle allocates its lists up front; no workload of the
    tests or the bench needs more in one iteration)."""
    return min(max(1, maxeval // npoints(d, key)) + 64, 1 << 25)


def integrate_bfs(p: W.Problem, nthreads=1, capacity=None, maxeval=None, epsrel=None, key=7):
    s, keep = to_struct(p.f)
    lo = np.array(p.lows, dtype=np.float64)
    hi = np.array(p.highs, dtype=np.float64)
    res = OrResult()
    me = p.maxeval if maxeval is None else maxeval
    cap = capacity if capacity is not None else _default_cap(me, p.f.dim, key)
    rc = lib().or_integrate_bfs_key(C.byref(s), _dp(lo), _dp(hi), p.epsrel if epsrel is None else epsrel,
                                    p.epsabs, me, cap, nthreads, key, C.byref(res))
    d = res.asdict()
    d["rc"] = rc
    return d


def integrate_bfs_check(p: W.Problem, nthreads=8, maxeval=None, epsrel=None, epsabs=None, key=7,
                        want_rule=True):
    """The breadth-first loop with its __float128 shadow: the FP64 result
    plus istar (exact-arithmetic sum of the loop's own FP64 integrand values
    over the contributing regions) and irule (nodes and integrand in
    __float128), each as a (hi, lo) pair, and the bound's A, E, n."""
    s, keep = to_struct(p.f)
    lo = np.array(p.lows, dtype=np.float64)
    hi = np.array(p.highs, dtype=np.float64)
    res = OrResult()
    chk = OrCheck()
    me = p.maxeval if maxeval is None else maxeval
    cap = _default_cap(me, p.f.dim, key)
    rc = lib().or_integrate_bfs_check(C.byref(s), _dp(lo), _dp(hi), p.epsrel if epsrel is None else epsrel,
                                      p.epsabs if epsabs is None else epsabs, me, cap, nthreads, key,
                                      1 if want_rule else 0, C.byref(res), C.byref(chk))
    d = res.asdict()
    d["rc"] = rc
    d.update(chk.asdict())
    return d


def integrate_bfs_allot(p: W.Problem, allot, nthreads=8, maxeval=None, epsrel=None):
    """The breadth-first GM7/5 loop with the shared-allotment finish policy
    (allot = b; 0 is the plain loop)."""
    s, keep = to_struct(p.f)
    lo = np.array(p.lows, dtype=np.float64)
    hi = np.array(p.highs, dtype=np.float64)
    res = OrResult()
    me = p.maxeval if maxeval is None else maxeval
    cap = _default_cap(me, p.f.dim, 7)
    rc = lib().or_integrate_bfs_allot(C.byref(s), _dp(lo), _dp(hi), p.epsrel if epsrel is None else epsrel,
                                      p.epsabs, me, cap, nthreads, allot, C.byref(res))
    d = res.asdict()
    d["rc"] = rc
    return d


def integrate_bfs_normalized(p: W.Problem, allot, nthreads=8, maxeval=None, epsrel=None):
    """The breadth-first GM7/5 loop with the normalized allotment (per-region
    tolerances adding up to epsrel |I|)."""
    s, keep = to_struct(p.f)
    lo = np.array(p.lows, dtype=np.float64)
    hi = np.array(p.highs, dtype=np.float64)
    res = OrResult()
    me = p.maxeval if maxeval is None else maxeval
    cap = _default_cap(me, p.f.dim, 7)
    rc = lib().or_integrate_bfs_normalized(C.byref(s), _dp(lo), _dp(hi), p.epsrel if epsrel is None else epsrel,
                                           p.epsabs, me, cap, nthreads, allot, C.byref(res))
    d = res.asdict()
    d["rc"] = rc
    return d


def r9_error(d, sums):
    """Error of the degree-9 rule (before the volume) from five rule sums;
    returns (err, (e1, e2, e3))."""
    r = OrRule9()
    assert lib().or_rule9_init(C.byref(r), d) == 0
    s = np.ascontiguousarray(sums, dtype=np.float64)
    e = np.zeros(3)
    err = lib().or_r9_error(C.byref(r), _dp(s), _dp(e))
    return err, e


def phase2_region(f: W.Integrand, lower, width, key=9, epsrel=1e-6, epsabs=0.0, pool_max=2048):
    """Phase 2 of one region: (result, error, pool size, active, zeroed)."""
    s, keep = to_struct(f)
    lo = np.ascontiguousarray(lower, dtype=np.float64)
    wd = np.ascontiguousarray(width, dtype=np.float64)
    out = np.empty(5)
    lib().or_phase2_region(key, C.byref(s), _dp(lo), _dp(wd), epsrel, epsabs, pool_max, _dp(out))
    return out[0], out[1], int(out[2]), int(out[3]), int(out[4])


def fp64_bound(p, d, n, A, E, I):
    """The rounding bound formula the library reports (or_fp64_bound)."""
    return lib().or_fp64_bound(p, d, n, A, E, I)


def region_check(f: W.Integrand, lower, width, key=7):
    """est, abssum, istar (hi, lo), irule (hi, lo) of one region."""
    s, keep = to_struct(f)
    lo = np.ascontiguousarray(lower, dtype=np.float64)
    wd = np.ascontiguousarray(width, dtype=np.float64)
    out = np.empty(6)
    lib().or_region_check(key, C.byref(s), _dp(lo), _dp(wd), _dp(out))
    return out


def integrate_two_phase(p: W.Problem, key=9, phase1_max=16384, pool_max=2048, nthreads=8):
    """The reference's two-phase method (breadth-first Phase 1, per-region Phase 2)."""
    s, keep = to_struct(p.f)
    lo = np.array(p.lows, dtype=np.float64)
    hi = np.array(p.highs, dtype=np.float64)
    res = OrResult()
    rc = lib().or_integrate_two_phase(C.byref(s), _dp(lo), _dp(hi), p.epsrel, p.epsabs, key, phase1_max,
                                      pool_max, nthreads, C.byref(res))
    d = res.asdict()
    d["rc"] = rc
    return d


def integrate_heap(p: W.Problem, maxeval=None):
    s, keep = to_struct(p.f)
    lo = np.array(p.lows, dtype=np.float64)
    hi = np.array(p.highs, dtype=np.float64)
    res = OrResult()
    rc = lib().or_integrate_heap(C.byref(s), _dp(lo), _dp(hi), p.epsrel, p.epsabs,
                                 p.maxeval if maxeval is None else maxeval, C.byref(res))
    d = res.asdict()
    d["rc"] = rc
    return d


def exact(f: W.Integrand, lows, highs):
    s, keep = to_struct(f)
    lo = np.array(lows, dtype=np.float64)
    hi = np.array(highs, dtype=np.float64)
    h = C.c_double(); l = C.c_double()
    rc = lib().or_exact(C.byref(s), _dp(lo), _dp(hi), C.byref(h), C.byref(l))
    if rc != 0:
        raise ValueError("no closed form for this integrand/box")
    return h.value, l.value


def initial_grid(d, lows, highs):
    lo = np.array(lows, dtype=np.float64)
    hi = np.array(highs, dtype=np.float64)
    n = lib().or_initial_grid(d, _dp(lo), _dp(hi), None, None)
    L = np.empty(n * d); Wd = np.empty(n * d)
    lib().or_initial_grid(d, _dp(lo), _dp(hi), _dp(L), _dp(Wd))
    return L.reshape(d, n).T.copy(), Wd.reshape(d, n).T.copy()


def split(lower, width, active, axis):
    """lower/width [n,d]; returns children [2A, d] (child 0 at 2j, child 1 at 2j+1)."""
    n, d = lower.shape
    L = np.ascontiguousarray(lower.T).ravel()
    Wd = np.ascontiguousarray(width.T).ravel()
    act = np.ascontiguousarray(active, dtype=np.int32)
    ax = np.ascontiguousarray(axis, dtype=np.int32)
    A = int(act.astype(bool).sum())
    oL = np.empty(max(1, 2 * A * d)); oW = np.empty(max(1, 2 * A * d))
    lib().or_split(d, n, _dp(L), _dp(Wd), act.ctypes.data_as(C.POINTER(C.c_int32)),
                   ax.ctypes.data_as(C.POINTER(C.c_int32)), _dp(oL), _dp(oW))
    return (oL[: 2 * A * d].reshape(d, 2 * A).T.copy(), oW[: 2 * A * d].reshape(d, 2 * A).T.copy())
