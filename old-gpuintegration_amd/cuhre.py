"""ctypes binding of libcuhre_amd.so (the C-ABI declared in include/cuhre.h).

This is the Python view of the drop-in boundary used by tests and bench.py.
It never falls back to anything: if the HIP library is missing or no GPU is
present, the calls raise.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Optional

import numpy as np

import workloads as W

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.environ.get("CUHRE_AMD_LIB") or os.path.join(HERE, "libcuhre_amd.so")
MAXDIM = 16
MAXTERMS = 8

OK = 0
NOT_CONVERGED = 1
ERR_ARG = -1
ERR_NONFINITE = -2
ERR_ALLOC = -3
ERR_DEVICE = -4
ERR_UNSUPPORTED = -5
ERR_CAPACITY = -6

# Symbols every build exports (kept in step with include/cuhre.h).
EXPORTS = [
    "cuhre_context_create", "cuhre_context_destroy", "cuhre_context_set", "cuhre_context_queue_ahead",
    "cuhre_integrate", "cuhre_integrate_two_phase", "cuhre_eval_regions", "cuhre_split_regions", "cuhre_sum",
    "cuhre_last_error", "cuhre_points_per_region", "cuhre_integrate_batch", "cuhre_batch_autotune",
    "cuhre_guide_create", "cuhre_guide_destroy", "cuhre_guide_set_params", "cuhre_guide_get_params",
    "cuhre_guide_forward", "cuhre_guide_gradient", "cuhre_guide_collect", "cuhre_guide_samples",
    "cuhre_guide_get_samples", "cuhre_guide_train", "cuhre_integrate_guided",
]
GUIDE_PARAMS = 11472


class Integrand(C.Structure):
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


class Result(C.Structure):
    _fields_ = [
        ("integral", C.c_double),
        ("error", C.c_double),
        ("fp64_err_bound", C.c_double),
        ("nregions", C.c_int64),
        ("neval", C.c_int64),
        ("fail", C.c_int32),
        ("iterations", C.c_int32),
        ("peak_regions", C.c_int64),
        ("nfinished", C.c_int64),
        ("ms_total", C.c_double),
        ("ms_eval", C.c_double),
        ("ms_split", C.c_double),
        ("eval_launches", C.c_int32),
        ("split_launches", C.c_int32),
        ("eval_flops", C.c_double),
        ("ms_eval_largest", C.c_double),
        ("nzeroed", C.c_int32),
        ("pad1", C.c_int32),
        ("largest_regions", C.c_int64),
    ]

    def asdict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class BatchConfig(C.Structure):
    _fields_ = [
        ("point_segments", C.c_int32),
        ("grid_blocks", C.c_int32),
        ("iterations_per_check", C.c_int32),
        ("use_graph", C.c_int32),
        ("persistent", C.c_int32),
        ("initial_capacity", C.c_int64),
        ("finish_budget", C.c_double),
    ]

    def asdict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class BatchInfo(C.Structure):
    _fields_ = [
        ("iterations", C.c_int32),
        ("launched_iterations", C.c_int32),
        ("total_regions", C.c_int64),
        ("total_neval", C.c_int64),
        ("not_converged", C.c_int32),
        ("graph_used", C.c_int32),
        ("point_segments", C.c_int32),
        ("grid_blocks", C.c_int32),
        ("ms_total", C.c_double),
        ("peak_list", C.c_int64),
        ("capacity", C.c_int64),
        ("grows", C.c_int32),
        ("persistent_finished", C.c_int32),
    ]

    def asdict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class GuidePolicy(C.Structure):
    _fields_ = [
        ("axis_mode", C.c_int32),
        ("finish_mode", C.c_int32),
        ("axis_confidence", C.c_float),
        ("finish_threshold", C.c_float),
        ("budget", C.c_double),
    ]


class GuideStats(C.Structure):
    _fields_ = [
        ("retired", C.c_int64),
        ("samples", C.c_int64),
        ("overridden", C.c_int64),
        ("guide_iterations", C.c_int32),
        ("pad", C.c_int32),
    ]

    def asdict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class TrainStats(C.Structure):
    _fields_ = [
        ("loss_first", C.c_double),
        ("loss_last", C.c_double),
        ("axis_accuracy", C.c_double),
        ("finish_accuracy", C.c_double),
        ("epochs", C.c_int32),
        ("steps", C.c_int32),
        ("samples", C.c_int64),
        ("ms", C.c_double),
    ]

    def asdict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class CuhreError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"cuhre error {code}: {msg}")
        self.code = code


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIBPATH):
            raise RuntimeError(f"{LIBPATH} is not built (run __graft_entry__.build())")
        L = C.CDLL(LIBPATH)
        dp = C.POINTER(C.c_double)
        ctxp = C.c_void_p
        L.cuhre_context_create.argtypes = [C.POINTER(C.c_void_p), C.c_int, C.c_int64]
        L.cuhre_context_destroy.argtypes = [ctxp]
        L.cuhre_context_set.argtypes = [ctxp, C.c_int, C.c_int]
        L.cuhre_context_queue_ahead.argtypes = [ctxp, C.c_int]
        L.cuhre_integrate.argtypes = [ctxp, C.POINTER(Integrand), C.c_int, C.c_double, C.c_double,
                                      C.c_int64, C.c_int, dp, dp, C.POINTER(Result)]
        L.cuhre_integrate_noalloc.argtypes = [ctxp, C.POINTER(Integrand), C.c_int, C.c_double, C.c_double,
                                              C.c_int64, C.c_int, dp, dp, C.POINTER(Result)]
        L.cuhre_noalloc_capacity.argtypes = [C.c_int]
        L.cuhre_noalloc_capacity.restype = C.c_int64
        L.cuhre_integrate_two_phase.argtypes = [ctxp, C.POINTER(Integrand), C.c_int, C.c_double, C.c_double,
                                                C.c_int, C.c_int64, C.c_int32, dp, dp, C.POINTER(Result)]
        L.cuhre_phase2_regions.argtypes = [ctxp, C.POINTER(Integrand), C.c_int, C.c_int, C.c_int64, dp, dp,
                                           C.c_double, C.c_double, C.c_int32, dp, dp, C.POINTER(C.c_int32)]
        L.cuhre_eval_regions.argtypes = [ctxp, C.POINTER(Integrand), C.c_int, C.c_int, C.c_int64,
                                         dp, dp, dp, dp, C.POINTER(C.c_int32), dp, dp, dp]
        L.cuhre_split_regions.argtypes = [ctxp, C.c_int, C.c_int64, dp, dp, dp, dp,
                                          C.POINTER(C.c_int32), C.c_double, C.c_double, dp, dp,
                                          C.POINTER(C.c_int64)]
        L.cuhre_sum.argtypes = [ctxp, C.c_int64, dp, dp]
        i64p = C.POINTER(C.c_int64)
        L.cuhre_integrate_batch.argtypes = [ctxp, C.c_int32, C.POINTER(Integrand), C.c_int, dp, dp, i64p,
                                            dp, dp, C.POINTER(BatchConfig), C.POINTER(Result),
                                            C.POINTER(BatchInfo)]
        L.cuhre_batch_autotune.argtypes = [ctxp, C.c_int32, C.POINTER(Integrand), C.c_int, dp, dp,
                                           C.POINTER(BatchConfig), dp]
        fp = C.POINTER(C.c_float)
        i32p = C.POINTER(C.c_int32)
        L.cuhre_guide_create.argtypes = [ctxp, C.c_uint64, C.POINTER(C.c_void_p)]
        L.cuhre_guide_destroy.argtypes = [C.c_void_p]
        L.cuhre_guide_set_params.argtypes = [C.c_void_p, fp]
        L.cuhre_guide_get_params.argtypes = [C.c_void_p, fp]
        L.cuhre_guide_forward.argtypes = [C.c_void_p, C.c_int64, fp, fp]
        L.cuhre_guide_gradient.argtypes = [C.c_void_p, C.c_int64, fp, i32p, fp, C.c_int, fp,
                                           C.POINTER(C.c_double)]
        L.cuhre_guide_collect.argtypes = [C.c_void_p, C.POINTER(Integrand), C.c_int, C.c_double, C.c_double,
                                          C.c_int64, dp, dp, C.c_int64, C.POINTER(GuidePolicy), C.POINTER(Result)]
        L.cuhre_guide_samples.argtypes = [C.c_void_p]
        L.cuhre_guide_samples.restype = C.c_int64
        L.cuhre_guide_get_samples.argtypes = [C.c_void_p, C.c_int64, C.c_int64, fp, i32p, fp]
        L.cuhre_guide_train.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_float, C.c_uint64,
                                        C.POINTER(TrainStats)]
        L.cuhre_integrate_guided.argtypes = [ctxp, C.c_void_p, C.POINTER(Integrand), C.c_int, C.c_double,
                                             C.c_double, C.c_int64, dp, dp, C.POINTER(GuidePolicy),
                                             C.POINTER(Result), C.POINTER(GuideStats)]
        L.cuhre_last_error.restype = C.c_char_p
        L.cuhre_points_per_region.argtypes = [C.c_int, C.c_int]
        L.cuhre_points_per_region.restype = C.c_int64
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _box(v):
    """A box bound as a ctypes double array (cheaper per call than a numpy
    array's pointer)."""
    return (C.c_double * len(v))(*v)


def _check(rc):
    if rc < 0:
        raise CuhreError(rc, lib().cuhre_last_error().decode())
    return rc


def to_struct(f: W.Integrand):
    s = Integrand()
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


class Context:
    """One engine context on one HIP device (cuhre_context_create)."""

    def __init__(self, device: int = 0, capacity: int = 0, profile: bool = False, segments: int = 0):
        self.h = C.c_void_p()
        _check(lib().cuhre_context_create(C.byref(self.h), device, capacity))
        self.set(profile, segments)

    def set(self, profile: bool = False, segments: int = 0):
        """profile: HIP-event timing of every launch; segments: 0 automatic or
        1/2/4/8/16 wavefronts per group of 64 regions."""
        _check(lib().cuhre_context_set(self.h, int(profile), segments))

    def queue_ahead(self, on: bool = True):
        """Plain loop: queue each iteration before the previous state is read
        (the default) or one at a time; results are identical."""
        _check(lib().cuhre_context_queue_ahead(self.h, int(on)))

    def close(self):
        if self.h:
            lib().cuhre_context_destroy(self.h)
            self.h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def integrate(self, p: W.Problem, epsrel: Optional[float] = None, maxeval: Optional[int] = None,
                  key: int = 0):
        s, keep = to_struct(p.f)
        res = Result()
        rc = lib().cuhre_integrate(self.h, C.byref(s), p.f.dim,
                                   p.epsrel if epsrel is None else epsrel, p.epsabs,
                                   p.maxeval if maxeval is None else maxeval, key,
                                   _box(p.lows), _box(p.highs), C.byref(res))
        _check(rc)
        d = res.asdict()
        d["rc"] = rc
        return d

    def integrate_noalloc(self, p: W.Problem, epsrel: Optional[float] = None, maxeval: Optional[int] = None,
                          key: int = 0, check: bool = True):
        """cuhre_integrate_noalloc: the plain method with its region lists in
        the library's static device array (no allocation in the call).
        With check=False the return code is returned instead of raised."""
        s, keep = to_struct(p.f)
        res = Result()
        rc = lib().cuhre_integrate_noalloc(self.h, C.byref(s), p.f.dim,
                                           p.epsrel if epsrel is None else epsrel, p.epsabs,
                                           p.maxeval if maxeval is None else maxeval, key,
                                           _box(p.lows), _box(p.highs), C.byref(res))
        if check:
            _check(rc)
        d = res.asdict()
        d["rc"] = rc
        if rc < 0:
            d["message"] = lib().cuhre_last_error().decode()
        return d

    def integrate_two_phase(self, p: W.Problem, key: int = 9, phase1_regions: int = 0, pool_max: int = 0):
        """The reference's two-phase method: breadth-first Phase 1 (at most 100
        iterations, until phase1_regions regions; 0: 2^14), then per-region
        Phase 2 (at most pool_max bisections each; 0: 2048)."""
        s, keep = to_struct(p.f)
        lo = np.ascontiguousarray(p.lows, dtype=np.float64)
        hi = np.ascontiguousarray(p.highs, dtype=np.float64)
        res = Result()
        rc = lib().cuhre_integrate_two_phase(self.h, C.byref(s), p.f.dim, p.epsrel, p.epsabs, key,
                                             phase1_regions, pool_max, _dp(lo), _dp(hi), C.byref(res))
        _check(rc)
        d = res.asdict()
        d["rc"] = rc
        return d

    def eval_regions(self, f: W.Integrand, lower, width, want_aux=False, key=0):
        """lower/width: [n, d] arrays.  Returns est, err, axis, abssum (+ i5, dvec[n,d])."""
        lower = np.asarray(lower, dtype=np.float64)
        width = np.asarray(width, dtype=np.float64)
        n, d = lower.shape
        L = np.ascontiguousarray(lower.T)
        Wd = np.ascontiguousarray(width.T)
        est = np.empty(n); err = np.empty(n); ab = np.empty(n)
        ax = np.empty(n, dtype=np.int32)
        i5 = np.empty(n) if want_aux else None
        dv = np.empty((d, n)) if want_aux else None
        s, keep = to_struct(f)
        _check(lib().cuhre_eval_regions(self.h, C.byref(s), d, key, n, _dp(L), _dp(Wd), _dp(est),
                                        _dp(err), ax.ctypes.data_as(C.POINTER(C.c_int32)), _dp(ab),
                                        _dp(i5) if want_aux else None,
                                        _dp(dv) if want_aux else None))
        if want_aux:
            return est, err, ax, ab, i5, dv.T.copy()
        return est, err, ax, ab

    def phase2_regions(self, f: W.Integrand, lower, width, key=9, epsrel=1e-6, epsabs=0.0, pool_max=0):
        """Phase 2 of the two-phase method on each region: (res, err, count)."""
        lower = np.asarray(lower, dtype=np.float64)
        width = np.asarray(width, dtype=np.float64)
        n, d = lower.shape
        L = np.ascontiguousarray(lower.T)
        Wd = np.ascontiguousarray(width.T)
        res = np.empty(n); err = np.empty(n)
        cnt = np.empty(n, dtype=np.int32)
        s, keep = to_struct(f)
        _check(lib().cuhre_phase2_regions(self.h, C.byref(s), d, key, n, _dp(L), _dp(Wd), epsrel, epsabs, pool_max,
                                          _dp(res), _dp(err), cnt.ctypes.data_as(C.POINTER(C.c_int32))))
        return res, err, cnt

    def split_regions(self, lower, width, est, err, axis, epsrel, epsabs=0.0):
        lower = np.asarray(lower, dtype=np.float64)
        width = np.asarray(width, dtype=np.float64)
        n, d = lower.shape
        L = np.ascontiguousarray(lower.T)
        Wd = np.ascontiguousarray(width.T)
        est = np.ascontiguousarray(est, dtype=np.float64)
        err = np.ascontiguousarray(err, dtype=np.float64)
        ax = np.ascontiguousarray(axis, dtype=np.int32)
        oL = np.empty(max(1, 2 * n * d)); oW = np.empty(max(1, 2 * n * d))
        na = C.c_int64()
        _check(lib().cuhre_split_regions(self.h, d, n, _dp(L), _dp(Wd), _dp(est), _dp(err),
                                         ax.ctypes.data_as(C.POINTER(C.c_int32)), epsrel, epsabs,
                                         _dp(oL), _dp(oW), C.byref(na)))
        A = na.value
        m = 2 * A
        return A, oL[: m * d].reshape(d, m).T.copy(), oW[: m * d].reshape(d, m).T.copy()

    def _batch_inputs(self, problems):
        count = len(problems)
        d = problems[0].f.dim
        arr = (Integrand * count)()
        keep = []
        for i, p in enumerate(problems):
            st, k = to_struct(p.f)
            arr[i] = st
            keep.append(k)
        lows = np.ascontiguousarray([x for p in problems for x in p.lows], dtype=np.float64)
        highs = np.ascontiguousarray([x for p in problems for x in p.highs], dtype=np.float64)
        return count, d, arr, lows, highs, keep

    def integrate_batch(self, problems, point_segments=0, grid_blocks=0, iterations_per_check=0,
                        use_graph=True, initial_capacity=0, raise_nonfinite=True, finish_budget=0.0,
                        persistent=False):
        """Integrates every problem (same dimension) in one batched call.
        Returns (results list in input order, info dict, return code).  With
        raise_nonfinite=False a non-finite integrand value returns rc -2 with
        the per-instance results (that instance's fail is -2).  persistent:
        one persistent launch runs the instances' loops in per-workgroup
        region queues first."""
        count, d, arr, lows, highs, keep = self._batch_inputs(problems)
        er = np.ascontiguousarray([p.epsrel for p in problems], dtype=np.float64)
        ea = np.ascontiguousarray([p.epsabs for p in problems], dtype=np.float64)
        me = np.ascontiguousarray([p.maxeval for p in problems], dtype=np.int64)
        cfg = BatchConfig(point_segments, grid_blocks, iterations_per_check, int(bool(use_graph)),
                          int(bool(persistent)), initial_capacity, finish_budget)
        out = (Result * count)()
        info = BatchInfo()
        rc = lib().cuhre_integrate_batch(self.h, count, arr, d, _dp(er), _dp(ea),
                                         me.ctypes.data_as(C.POINTER(C.c_int64)), _dp(lows), _dp(highs),
                                         C.byref(cfg), out, C.byref(info))
        if raise_nonfinite or rc != -2:
            _check(rc)
        return [out[i].asdict() for i in range(count)], info.asdict(), rc

    def batch_autotune(self, problems):
        count, d, arr, lows, highs, keep = self._batch_inputs(problems)
        best = BatchConfig()
        table = np.zeros(15)
        _check(lib().cuhre_batch_autotune(self.h, count, arr, d, _dp(lows), _dp(highs), C.byref(best),
                                          _dp(table)))
        return best.asdict(), table

    def sum(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        out = C.c_double()
        _check(lib().cuhre_sum(self.h, x.size, _dp(x), C.byref(out)))
        return out.value


def _fp(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


class Guide:
    """Learned split guide (BF16 MLP on MFMA) attached to a Context."""

    def __init__(self, ctx: Context, seed: int = 1):
        self.ctx = ctx
        self.h = C.c_void_p()
        _check(lib().cuhre_guide_create(ctx.h, seed, C.byref(self.h)))

    def close(self):
        if self.h:
            lib().cuhre_guide_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def params(self):
        p = np.empty(GUIDE_PARAMS, dtype=np.float32)
        _check(lib().cuhre_guide_get_params(self.h, _fp(p)))
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        assert p.size == GUIDE_PARAMS
        _check(lib().cuhre_guide_set_params(self.h, _fp(p)))

    def forward(self, X):
        X = np.ascontiguousarray(X, dtype=np.float32)
        n = X.shape[0]
        Y = np.empty((n, 16), dtype=np.float32)
        _check(lib().cuhre_guide_forward(self.h, n, _fp(X), _fp(Y)))
        return Y

    def gradient(self, X, axis, finish, dim):
        X = np.ascontiguousarray(X, dtype=np.float32)
        a = np.ascontiguousarray(axis, dtype=np.int32)
        f = np.ascontiguousarray(finish, dtype=np.float32)
        g = np.empty(GUIDE_PARAMS, dtype=np.float32)
        loss = C.c_double()
        _check(lib().cuhre_guide_gradient(self.h, X.shape[0], _fp(X), a.ctypes.data_as(C.POINTER(C.c_int32)),
                                          _fp(f), dim, _fp(g), C.byref(loss)))
        return g, loss.value

    def collect(self, p: W.Problem, max_samples: int, maxeval=None, axis_mode=0, axis_confidence=0.6):
        """Run that logs up to max_samples; axis_mode 1 lets the network pick
        the axes it is confident about (so the log covers its own regions)."""
        s, keep = to_struct(p.f)
        lo = np.ascontiguousarray(p.lows, dtype=np.float64)
        hi = np.ascontiguousarray(p.highs, dtype=np.float64)
        res = Result()
        pol = GuidePolicy(axis_mode, 0, axis_confidence, 1.0, 0.0)
        rc = lib().cuhre_guide_collect(self.h, C.byref(s), p.f.dim, p.epsrel, p.epsabs,
                                       p.maxeval if maxeval is None else maxeval, _dp(lo), _dp(hi),
                                       max_samples, C.byref(pol), C.byref(res))
        _check(rc)
        d = res.asdict()
        d["rc"] = rc
        return d

    def samples(self):
        return lib().cuhre_guide_samples(self.h)

    def get_samples(self, first=0, count=None):
        n = self.samples()
        count = n - first if count is None else count
        X = np.empty((count, 32), dtype=np.float32)
        a = np.empty(count, dtype=np.int32)
        f = np.empty(count, dtype=np.float32)
        _check(lib().cuhre_guide_get_samples(self.h, first, count, _fp(X), a.ctypes.data_as(C.POINTER(C.c_int32)),
                                             _fp(f)))
        return X, a, f

    def train(self, dim, epochs=4, batch=4096, lr=1e-3, seed=7):
        st = TrainStats()
        _check(lib().cuhre_guide_train(self.h, dim, epochs, batch, lr, seed, C.byref(st)))
        return st.asdict()

    def integrate(self, p: W.Problem, axis_mode=1, finish_mode=1, axis_confidence=0.6, finish_threshold=0.5,
                  budget=0.25, maxeval=None, epsrel=None):
        s, keep = to_struct(p.f)
        lo = np.ascontiguousarray(p.lows, dtype=np.float64)
        hi = np.ascontiguousarray(p.highs, dtype=np.float64)
        pol = GuidePolicy(axis_mode, finish_mode, axis_confidence, finish_threshold, budget)
        res = Result()
        gst = GuideStats()
        rc = lib().cuhre_integrate_guided(self.ctx.h, self.h, C.byref(s), p.f.dim,
                                          p.epsrel if epsrel is None else epsrel, p.epsabs,
                                          p.maxeval if maxeval is None else maxeval, _dp(lo), _dp(hi),
                                          C.byref(pol), C.byref(res), C.byref(gst))
        _check(rc)
        d = res.asdict()
        d["rc"] = rc
        d["guide"] = gst.asdict()
        return d


def noalloc_capacity(dim: int) -> int:
    """Regions per list the static device array of integrate_noalloc holds."""
    return lib().cuhre_noalloc_capacity(dim)


def points_per_region(dim: int, key: int = 0) -> int:
    return lib().cuhre_points_per_region(dim, key)
