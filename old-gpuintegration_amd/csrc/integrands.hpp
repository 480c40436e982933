// integrands.hpp -- FP64 device integrand families.
//
// The reference evaluates a user functor `double operator()(double x[], int
// dim)` at each cubature point (quad/GPUquad/GPUQuadSample.cu:84,
// cudaCuhre.cu:11-22).  Functors cannot cross a C ABI, so the families are
// compiled in and selected by DevIntegrand::family; each is templated on the
// dimension so the point vector stays in registers.
#pragma once

#include "common.hpp"

namespace cuhre {

// b^e, e >= 0: right-to-left binary powering.
__device__ __forceinline__ double ipow_dev(double b, int e)
{
  double r = 1.0;
  while (e > 0) {
    if (e & 1) r *= b;
    e >>= 1;
    if (e == 0) break;
    b *= b;
  }
  return r;
}

// b^E for a compile-time E, unrolled (E = 8: three squarings).
template <int E>
__device__ __forceinline__ double ipow_const(double b)
{
  if constexpr (E == 0) {
    return 1.0;
  } else if constexpr (E == 1) {
    return b;
  } else if constexpr (E % 2 == 0) {
    const double h = ipow_const<E / 2>(b);
    return h * h;
  } else {
    return b * ipow_const<E - 1>(b);
  }
}

// 1/p from the hardware estimate (v_rcp_f64) refined by two Newton-Raphson
// steps r <- r + r (1 - p r): within one ulp of the rounded quotient for
// normal finite p.  p = +-inf or 0 gives NaN (the steps meet inf * 0), which
// the evaluation reports as a non-finite integrand value.
__device__ __forceinline__ double recip_dev(double p)
{
  double r = __builtin_amdgcn_rcp(p);
  double e = fma(-p, r, 1.0);
  r = fma(r, e, r);
  e = fma(-p, r, 1.0);
  return fma(r, e, r);
}

// Each family states unrolls(d) (kUnroll at its D): whether the rule's
// point loops unroll around its evaluation (seg_eval).  True where the
// integrand accumulates per-coordinate terms that the compiler can share
// between points with equal coordinates (products and sums of
// one-coordinate factors, the table lookups in (x1, x2)); false where one
// expensive function of all coordinates dominates, so unrolled copies of it
// would only fill the instruction cache: cos of a sum, exp of a sum above
// six dimensions (up to six its shared sum still pays), the polynomial's
// runtime term loop, the cosine sum's d cosines per point and LC_LT_t's
// erfc and exponentials.
template <int FAM, int D>
struct Family;

template <int D>
struct Family<CUHRE_GAUSSIAN, D> {
  static constexpr bool unrolls(int d) { return d <= 6; }
  static constexpr bool kUnroll = unrolls(D);
  __device__ __forceinline__ static double eval(const DevIntegrand& P, const double (&x)[D])
  {
    double s = 0.0;
#pragma unroll
    for (int i = 0; i < D; ++i) {
      double t = P.a[i] * (x[i] - P.w[i]);
      s = fma(t, t, s);
    }
    return exp(-s);
  }
};

template <int D>
struct Family<CUHRE_PRODUCT_PEAK, D> {
  static constexpr bool unrolls(int) { return true; }
  static constexpr bool kUnroll = unrolls(D);
  __device__ __forceinline__ static double eval(const DevIntegrand& P, const double (&x)[D])
  {
    double p = 1.0;
#pragma unroll
    for (int i = 0; i < D; ++i) {
      double t = x[i] - P.w[i];
      p *= fma(t, t, P.b[i]);
    }
    return recip_dev(p);
  }
};

template <int D>
struct Family<CUHRE_OSCILLATORY, D> {
  static constexpr bool unrolls(int) { return false; }
  static constexpr bool kUnroll = unrolls(D);
  __device__ __forceinline__ static double eval(const DevIntegrand& P, const double (&x)[D])
  {
    double s = P.s0;
#pragma unroll
    for (int i = 0; i < D; ++i) s = fma(P.a[i], x[i], s);
    return cos(s);
  }
};

template <int D>
struct Family<CUHRE_CORNER_PEAK, D> {
  static constexpr bool unrolls(int) { return true; }
  static constexpr bool kUnroll = unrolls(D);
  __device__ __forceinline__ static double eval(const DevIntegrand& P, const double (&x)[D])
  {
    double s = 1.0;
#pragma unroll
    for (int i = 0; i < D; ++i) s = fma(P.a[i], x[i], s);
    const double r = recip_dev(s);
    // Genz's exponent d + 1 unrolled; any other exponent by the loop
    return P.exponent == D + 1 ? ipow_const<D + 1>(r) : ipow_dev(r, P.exponent);
  }
};

template <int D>
struct Family<CUHRE_CONTINUOUS, D> {
  static constexpr bool unrolls(int d) { return d <= 6; }
  static constexpr bool kUnroll = unrolls(D);
  __device__ __forceinline__ static double eval(const DevIntegrand& P, const double (&x)[D])
  {
    double s = 0.0;
#pragma unroll
    for (int i = 0; i < D; ++i) s = fma(P.a[i], fabs(x[i] - P.w[i]), s);
    return exp(-s);
  }
};

template <int D>
struct Family<CUHRE_POLYNOMIAL, D> {
  static constexpr bool unrolls(int) { return false; }
  static constexpr bool kUnroll = unrolls(D);
  __device__ __forceinline__ static double eval(const DevIntegrand& P, const double (&x)[D])
  {
    double s = 0.0;
    for (int t = 0; t < P.nterms; ++t) {
      double p = P.coef[t];
#pragma unroll
      for (int i = 0; i < D; ++i) p *= ipow_dev(x[i], P.powers[t][i]);
      s += p;
    }
    return s;
  }
};

// The reference driver's test functor (cudaCuhre.cu:11-22, function.cu FUN 5).
template <int D>
struct Family<CUHRE_COSINE_SUM, D> {
  static constexpr bool unrolls(int) { return false; }
  static constexpr bool kUnroll = unrolls(D);
  __device__ __forceinline__ static double eval(const DevIntegrand&, const double (&x)[D])
  {
    double s = 0.0;
#pragma unroll
    for (int i = 0; i < D; ++i) s = s - cos(10.0 * x[i]) / 0.054402111088937;
    return s / 2.0;
  }
};

// Table-driven 6-D integrand: B(x1,x2) x3^-1.5 exp(-(x4-0.5)^2/0.02) (1+x5 x6),
// B bilinear on an nx*ny FP64 grid over [0,1]^2 (table[i*ny + j]).  The grid
// (512 KiB at 256x256) exceeds one CU's LDS, so it is read through L2.
template <int D>
struct Family<CUHRE_TABLE6, D> {
  static constexpr bool unrolls(int) { return true; }
  static constexpr bool kUnroll = unrolls(D);
  // Cell of the table holding x (the same clamp the oracle applies).
  __device__ __forceinline__ static void cell(const DevIntegrand& P, double x0, double x1, int& ix, int& iy,
                                              double& tx, double& ty)
  {
    // no contraction: fusing fx - ix with the product into an FMA gives
    // other interpolation weights than the oracle's (by up to an ulp of fx)
#pragma clang fp contract(off)
    const int nx = P.nx, ny = P.ny;
    const double fx = x0 * (double)(nx - 1), fy = x1 * (double)(ny - 1);
    ix = (int)floor(fx);
    iy = (int)floor(fy);
    ix = ix < 0 ? 0 : (ix > nx - 2 ? nx - 2 : ix);
    iy = iy < 0 ? 0 : (iy > ny - 2 ? ny - 2 : iy);
    tx = fx - (double)ix;
    ty = fy - (double)iy;
  }
  __device__ __forceinline__ static double eval(const DevIntegrand& P, const double (&x)[D])
  {
    int ix, iy;
    double tx, ty;
    cell(P, x[0], x[1], ix, iy, tx, ty);
    const int ny = P.ny;
    const double* T = P.table + (size_t)ix * ny + iy;
    return eval_corners(x, tx, ty, T[0], T[1], T[ny], T[ny + 1]);
  }
  // Bilinear value from the cell's four corners, times the other factors.
  __device__ __forceinline__ static double eval_corners(const double (&x)[D], double tx, double ty, double t00,
                                                        double t01, double t10, double t11)
  {
    double b0 = fma(tx, t10 - t00, t00);
    double b1 = fma(tx, t11 - t01, t01);
    double B = fma(ty, b1 - b0, b0);
    double x3 = x[2];
    double p3 = 1.0 / (x3 * sqrt(x3));
    double t4 = x[3] - 0.5;
    double g4 = exp(-(t4 * t4) / 0.02);
    double q = fma(x[4], x[5], 1.0);
    return B * p3 * g4 * q;
  }
};

// The reference's LC_LT_t integrand (lc_lt_t.h:20-50) in x = (lt, zt, lc).
// Five fields tau, mu, sigma, fmsk, fprj are interpolated bilinearly in
// (lt, zt) on float tables in the arithmetic of the reference's Interp2D
// (GPUQuadInterp2D.cu:49-100): float nodes and values, double weights, the
// result rounded to float.  The node interval is the reference's
// SmallNearestN (the last node below the coordinate, the first interval
// below the grid), kept inside the grid above its last node, where the
// reference reads one texel past the end.  The closed form follows in
// double with the reference's float sub-expressions kept float.
// Contraction is off so the FP64 operations are the ones written.
__device__ __forceinline__ int lclt_interval(const float* g, int n, double t)
{
  int k = 0;
  for (int i = 1; i < n; ++i) k += (double)g[i] < t ? 1 : 0;
  return k > n - 2 ? n - 2 : k;
}

__device__ __forceinline__ float lclt_field(const float* F, int nx, int ix, int iy, double wy1, double wy2,
                                            double t1, double t2)
{
#pragma clang fp contract(off)
  const float q11 = F[ix * nx + iy], q12 = F[ix * nx + iy + 1];
  const float q21 = F[(ix + 1) * nx + iy], q22 = F[(ix + 1) * nx + iy + 1];
  return (float)((q11 * wy2 + q12 * wy1) * t1 + (q21 * wy2 + q22 * wy1) * t2);
}

__device__ __forceinline__ double lclt_gaussian(double x, double mu, double sigma)
{
#pragma clang fp contract(off)
  const double z = (x - mu) / sigma;
  return exp(-z * z / 2.) * 0.3989422804014327 / sigma;
}

template <int D>
struct Family<CUHRE_LC_LT_T, D> {
  static constexpr bool unrolls(int) { return false; }
  static constexpr bool kUnroll = unrolls(D);
  __device__ __forceinline__ static double eval(const DevIntegrand& P, const double (&x)[D])
  {
#pragma clang fp contract(off)
    const float* T = P.ftab;
    const int nx = P.nx, ny = P.ny;
    const double lt = x[0], zt = x[1], lc = x[2];
    const int iy = lclt_interval(T, nx, lt);
    const int ix = lclt_interval(T + nx, ny, zt);
    const float y1 = T[iy], y2 = T[iy + 1];
    const float x1 = T[nx + ix], x2 = T[nx + ix + 1];
    const float den = (x2 - x1) * (y2 - y1);
    const double t1 = (x2 - zt) / den, t2 = (zt - x1) / den;
    const double wy2 = y2 - lt, wy1 = lt - y1;
    const float* F = T + nx + ny;
    const int fs = nx * ny;
    const float tau = lclt_field(F, nx, ix, iy, wy1, wy2, t1, t2);
    const float mu = lclt_field(F + fs, nx, ix, iy, wy1, wy2, t1, t2);
    const float sigma = lclt_field(F + 2 * fs, nx, ix, iy, wy1, wy2, t1, t2);
    const float fmsk = lclt_field(F + 3 * fs, nx, ix, iy, wy1, wy2, t1, t2);
    const float fprj = lclt_field(F + 4 * fs, nx, ix, iy, wy1, wy2, t1, t2);
    const double exptau = exp(tau * (2.0 * mu + tau * sigma * sigma - 2.0 * lc) / 2.0);
    const double root_two_sigma = sqrt(2.0) * sigma;
    const double mu_tau_sig_sqr = mu + tau * sigma * sigma;
    return (1.0 - fmsk) * (1.0 - fprj) * lclt_gaussian(lc, mu, sigma) +
           0.5 * ((1.0 - fmsk) * fprj * tau + fmsk * fprj / lt) * exptau *
               erfc((mu_tau_sig_sqr - lc) / root_two_sigma) +
           0.5 * fmsk / lt * (erfc((lc - mu) / root_two_sigma) - erfc((lc + lt - mu) / root_two_sigma)) -
           0.5 * fmsk * fprj / lt * (exp(-tau * lt) * exptau * erfc((mu_tau_sig_sqr - (lc + lt)) / root_two_sigma));
  }
};

}  // namespace cuhre
