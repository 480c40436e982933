"""Synthetic integrands and the named workloads, generated from fixed seeds.

Every random parameter comes from SplitMix64 (Steele, Lea & Flood 2014)
mapped to U[0,1) through its top 53 bits, so the same seed gives the same
integrand in every process and on every machine.

Integrand families follow the reference's integrand concept,
``__device__ double operator()(double x[], int dim)`` (cudaCuhre.cu:11-22),
and the Genz test families listed in function.cu:111-156.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

GAUSSIAN = 1
PRODUCT_PEAK = 2
OSCILLATORY = 3
CORNER_PEAK = 4
CONTINUOUS = 5
POLYNOMIAL = 6
COSINE_SUM = 7
TABLE6 = 8
LC_LT_T = 9

FAMILY_NAMES = {
    GAUSSIAN: "gaussian",
    PRODUCT_PEAK: "product_peak",
    OSCILLATORY: "oscillatory",
    CORNER_PEAK: "corner_peak",
    CONTINUOUS: "continuous",
    POLYNOMIAL: "polynomial",
    COSINE_SUM: "cosine_sum",
    TABLE6: "table6",
    LC_LT_T: "lc_lt_t",
}

MAXDIM = 16
MAXTERMS = 8
_MASK = (1 << 64) - 1


class SplitMix64:
    """SplitMix64 generator; ``uniform()`` is U[0,1) with 53 random bits."""

    def __init__(self, seed: int):
        self.state = seed & _MASK

    def next_u64(self) -> int:
        self.state = (self.state + 0x9E3779B97F4A7C15) & _MASK
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & _MASK
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & _MASK
        return z ^ (z >> 31)

    def uniform(self) -> float:
        return (self.next_u64() >> 11) * (1.0 / 9007199254740992.0)

    def uniform_range(self, lo: float, hi: float) -> float:
        return lo + (hi - lo) * self.uniform()

    def normal(self) -> float:
        """Box-Muller, one value per call (the sine branch is discarded)."""
        u1 = 1.0 - self.uniform()  # (0, 1]
        u2 = self.uniform()
        return math.sqrt(-2.0 * math.log(u1)) * math.cos(2.0 * math.pi * u2)


def gm_npoints(d: int) -> int:
    """Points of the Genz-Malik degree-7 rule: 2^d + 2d^2 + 2d + 1."""
    return (1 << d) + 2 * d * d + 2 * d + 1


@dataclass
class Integrand:
    family: int
    dim: int
    a: List[float] = field(default_factory=list)
    w: List[float] = field(default_factory=list)
    u: float = 0.0
    exponent: int = 0
    coef: List[float] = field(default_factory=list)
    powers: List[List[int]] = field(default_factory=list)
    table: Optional[np.ndarray] = None  # (nx, ny) float64, table[i, j] = B(i/(nx-1), j/(ny-1))
    table_shape: Optional[tuple] = None  # (nx, ny) when `table` is a packed 1-D array (LC_LT_T)

    @property
    def name(self) -> str:
        return FAMILY_NAMES[self.family]


@dataclass
class Problem:
    name: str
    f: Integrand
    lows: List[float]
    highs: List[float]
    epsrel: float
    epsabs: float
    maxeval: int


def unit_box(d: int):
    return [0.0] * d, [1.0] * d


# ---------------------------------------------------------------- families

def gaussian(a, w) -> Integrand:
    return Integrand(GAUSSIAN, len(a), list(a), list(w))


def product_peak(a, w) -> Integrand:
    return Integrand(PRODUCT_PEAK, len(a), list(a), list(w))


def oscillatory(a, u) -> Integrand:
    return Integrand(OSCILLATORY, len(a), list(a), [0.0] * len(a), u=u)


def corner_peak(a, k) -> Integrand:
    return Integrand(CORNER_PEAK, len(a), list(a), [0.0] * len(a), exponent=int(k))


def continuous(a, w) -> Integrand:
    return Integrand(CONTINUOUS, len(a), list(a), list(w))


def polynomial(d, terms) -> Integrand:
    """terms: list of (coef, [powers...])."""
    return Integrand(POLYNOMIAL, d, [0.0] * d, [0.0] * d,
                     coef=[float(c) for c, _ in terms],
                     powers=[list(p) for _, p in terms])


def cosine_sum(d) -> Integrand:
    return Integrand(COSINE_SUM, d, [0.0] * d, [0.0] * d)


def table6(table: np.ndarray) -> Integrand:
    return Integrand(TABLE6, 6, [0.0] * 6, [0.0] * 6,
                     table=np.ascontiguousarray(table, dtype=np.float64))


def lc_lt_t(tables: dict) -> Integrand:
    """The reference's LC_LT_t integrand (lc_lt_t.h:20-50) in x = (lt, zt, lc):
    tables holds the lt and zt nodes and the fields tau, mu, sigma, fmsk, fprj
    (field[j*nx + i] = value at (lt_i, zt_j)), packed as the C-ABI expects."""
    lt = np.asarray(tables["lt"], dtype=np.float64)
    zt = np.asarray(tables["zt"], dtype=np.float64)
    parts = [lt, zt] + [np.asarray(tables[k], dtype=np.float64) for k in ("tau", "mu", "sigma", "fmsk", "fprj")]
    for k, a in zip(("tau", "mu", "sigma", "fmsk", "fprj"), parts[2:]):
        assert a.size == lt.size * zt.size, k
    return Integrand(LC_LT_T, 3, [0.0] * 3, [0.0] * 3, table=np.concatenate(parts),
                     table_shape=(lt.size, zt.size))


def lclt_driver(tables: dict, epsrel: float = 1e-7) -> Problem:
    """The reference driver's run (cudaCuhre.cu:43-79): LC_LT_t over the
    Volume lows (1, .1, 0), highs (2, .3, 1) given as floats, epsabs 1e-12."""
    f32 = lambda v: float(np.float32(v))  # noqa: E731
    return Problem("lc_lt_t3", lc_lt_t(tables), [f32(1.0), f32(0.1), f32(0.0)], [f32(2.0), f32(0.3), f32(1.0)],
                   epsrel, 1e-12, 10**9)


# ---------------------------------------------------------------- workloads

def c1_gaussian3(seed: int = 1) -> Problem:
    """3-D Gaussian peak, a_i = 5, w_i ~ U(0.2, 0.8); epsrel 1e-6, maxeval 1e6."""
    rng = SplitMix64(seed)
    w = [rng.uniform_range(0.2, 0.8) for _ in range(3)]
    lo, hi = unit_box(3)
    return Problem("gaussian3d", gaussian([5.0] * 3, w), lo, hi, 1e-6, 0.0, 10**6)


def c2_oscillatory8(seed: int = 2) -> Problem:
    """8-D oscillatory cos(2 pi u + sum a_i x_i), a_i = 9/8 U(0,1), u ~ U(0,1)."""
    rng = SplitMix64(seed)
    a = [1.125 * rng.uniform() for _ in range(8)]
    u = rng.uniform()
    lo, hi = unit_box(8)
    return Problem("oscillatory8d", oscillatory(a, u), lo, hi, 1e-6, 0.0, 10**9)


def c2_product_peak8(seed: int = 3) -> Problem:
    """8-D product peak prod 1/(a^-2 + (x_i - w_i)^2), a_i = 3, w_i ~ U(0,1)."""
    rng = SplitMix64(seed)
    w = [rng.uniform() for _ in range(8)]
    lo, hi = unit_box(8)
    return Problem("product_peak8d", product_peak([3.0] * 8, w), lo, hi, 1e-6, 0.0, 10**9)


def cancelling_oscillatory(d: int = 5, seed: int = 7, epsabs: float = 1e-8) -> Problem:
    """Oscillatory integrand whose integral over [0,1]^d is zero: a_i = 9/8
    U(0,1) and u chosen so that 2 pi u + sum a_i / 2 = pi / 2 (mod 2 pi), so
    the closed form cos(2 pi u + sum a_i / 2) prod 2 sin(a_i / 2) / a_i
    vanishes up to the rounding of u.  The positive and negative halves
    cancel, so |I| is tiny against sum |w f|; tolerance epsabs."""
    rng = SplitMix64(seed)
    a = [1.125 * rng.uniform() for _ in range(d)]
    u = ((0.5 * math.pi - 0.5 * sum(a)) / (2.0 * math.pi)) % 1.0
    lo, hi = unit_box(d)
    return Problem(f"cancelling_oscillatory{d}d", oscillatory(a, u), lo, hi, 0.0, epsabs, 10**9)


def c3_field(seed: int = 4, n: int = 256, ncos: int = 32) -> np.ndarray:
    """Smooth random field on an n x n grid: sum of ncos cosines
    c_k cos(2 pi (m_k x1 + n_k x2) + phi_k), c_k ~ N(0,1), m_k, n_k in {0..4},
    phi_k ~ U(0, 2 pi), sampled at nodes (i/(n-1), j/(n-1))."""
    rng = SplitMix64(seed)
    terms = []
    for _ in range(ncos):
        c = rng.normal()
        m = int(5 * rng.uniform())
        k = int(5 * rng.uniform())
        phi = 2.0 * math.pi * rng.uniform()
        terms.append((c, m, k, phi))
    x = np.arange(n, dtype=np.float64) / (n - 1)
    X1, X2 = np.meshgrid(x, x, indexing="ij")
    B = np.zeros((n, n), dtype=np.float64)
    for c, m, k, phi in terms:
        B += c * np.cos(2.0 * math.pi * (m * X1 + k * X2) + phi)
    return B


def c3_table6(seed: int = 4) -> Problem:
    """6-D table-driven integrand on [0,1]^2 x [0.1,1] x [0,1]^3, epsrel 1e-5."""
    f = table6(c3_field(seed))
    lo = [0.0, 0.0, 0.1, 0.0, 0.0, 0.0]
    hi = [1.0] * 6
    return Problem("table6d", f, lo, hi, 1e-5, 0.0, 10**9)


def c4_batch(seed: int = 5, count: int = 4096, d: int = 5) -> List[Problem]:
    """count independent 5-D integrals: family uniform over {Gaussian, product
    peak, corner peak (exponent 6)}, a_i ~ U(1,10), w_i ~ U(0,1); epsrel 1e-5,
    maxeval 1e10 per instance (the sharpest product peaks and Gaussians need
    up to ~1e8 evaluations with the plain per-region test; at this budget
    every instance certifies)."""
    rng = SplitMix64(seed)
    out = []
    lo, hi = unit_box(d)
    for i in range(count):
        fam = int(3 * rng.uniform())
        a = [rng.uniform_range(1.0, 10.0) for _ in range(d)]
        w = [rng.uniform() for _ in range(d)]
        if fam == 0:
            f = gaussian(a, w)
        elif fam == 1:
            f = product_peak(a, w)
        else:
            f = corner_peak(a, 6)
        out.append(Problem(f"batch{i}", f, lo, hi, 1e-5, 0.0, 10**10))
    return out


def c5_corner7(seed: int = 6) -> Problem:
    """7-D corner peak (1 + sum a_i x_i)^-8, a_i ~ U(0.5, 5); epsrel 1e-6."""
    rng = SplitMix64(seed)
    a = [rng.uniform_range(0.5, 5.0) for _ in range(7)]
    lo, hi = unit_box(7)
    return Problem("corner_peak7d", corner_peak(a, 8), lo, hi, 1e-6, 0.0, 10**9)


def c5_train_problems(first_seed: int = 100, count: int = 16) -> List[Problem]:
    """Training integrands for the learned split guide (held apart from c5)."""
    return [c5_corner7(s) for s in range(first_seed, first_seed + count)]
