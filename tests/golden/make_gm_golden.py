"""Writes tests/golden/gm_scipy_regions.json: per-region Genz-Malik degree-7
estimates and |I7 - I5| error estimates computed by SciPy's independent
implementation (scipy.integrate._rules.GenzMalikCubature, SciPy 1.15.3, a
third-party library of this image -- not the reference).

The fixture pins the oracle's rule (nodes, weights, scaling to a box) to a
second implementation.  Run from the repository root:

    python tests/golden/make_gm_golden.py
"""
import json
import os
import sys

import numpy as np
import scipy
from scipy.integrate._rules import GenzMalikCubature

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "..", "old-gpuintegration_amd"))
import workloads as W  # noqa: E402


def np_f(f: W.Integrand):
    a = np.array(f.a[: f.dim])
    w = np.array(f.w[: f.dim])
    if f.family == W.GAUSSIAN:
        return lambda x: np.exp(-np.sum((a * (x - w)) ** 2, axis=-1))
    if f.family == W.PRODUCT_PEAK:
        return lambda x: 1.0 / np.prod(1.0 / a**2 + (x - w) ** 2, axis=-1)
    if f.family == W.OSCILLATORY:
        return lambda x: np.cos(2 * np.pi * f.u + x @ a)
    if f.family == W.CORNER_PEAK:
        return lambda x: (1.0 + x @ a) ** (-f.exponent)
    if f.family == W.CONTINUOUS:
        return lambda x: np.exp(-np.sum(a * np.abs(x - w), axis=-1))
    raise ValueError(f.family)


def main():
    rng = W.SplitMix64(20261203)
    cases = []
    for d in (2, 3, 5, 8):
        rule = GenzMalikCubature(d)
        for fam in (W.GAUSSIAN, W.PRODUCT_PEAK, W.OSCILLATORY, W.CORNER_PEAK, W.CONTINUOUS):
            a = [rng.uniform_range(0.5, 5.0) for _ in range(d)]
            w = [rng.uniform() for _ in range(d)]
            if fam == W.GAUSSIAN:
                f = W.gaussian(a, w)
            elif fam == W.PRODUCT_PEAK:
                f = W.product_peak(a, w)
            elif fam == W.OSCILLATORY:
                f = W.oscillatory(a, rng.uniform())
            elif fam == W.CORNER_PEAK:
                f = W.corner_peak(a, d + 1)
            else:
                f = W.continuous(a, w)
            fn = np_f(f)
            regions = []
            for _ in range(6):
                lo = np.array([rng.uniform_range(0.0, 0.7) for _ in range(d)])
                wd = np.array([rng.uniform_range(0.01, 0.3) for _ in range(d)])
                est = float(rule.estimate(fn, lo, lo + wd))
                err = float(rule.estimate_error(fn, lo, lo + wd))
                regions.append({"lower": lo.tolist(), "width": wd.tolist(), "est": est, "err": err})
            cases.append({"family": fam, "dim": d, "a": f.a, "w": f.w, "u": f.u,
                          "exponent": f.exponent, "regions": regions})
    out = {"generator": "scipy.integrate._rules.GenzMalikCubature", "scipy": scipy.__version__,
           "cases": cases}
    with open(os.path.join(HERE, "gm_scipy_regions.json"), "w") as fh:
        json.dump(out, fh, indent=1)
    print("wrote", sum(len(c["regions"]) for c in cases), "regions")


if __name__ == "__main__":
    main()
