// eval_kernels.hpp -- Genz-Malik degree-7/5 rule evaluation on gfx950.
//
// Replaces the reference's per-region evaluation: INTEGRATE_GPU_PHASE1
// (quad/GPUquad/GPUQuadPhases.cu:37-79), which runs SampleRegionBlock
// (quad/GPUquad/GPUQuadSample.cu:99-220) with one 256-thread block per
// region, five LDS tree reductions and a serial fourth-difference loop on
// thread 0.  Two mappings are provided:
//
//  * k_eval_lane  -- one region per lane.  All 64 lanes of a wave walk the
//    same point sequence, so the generator offsets are wave-uniform values
//    computed on the scalar unit; the FP64 VALU only does the coordinate
//    FMAs, the integrand and the weighted sums.  No cross-lane reduction is
//    needed for a region.  Used when the list is large enough to fill the
//    chip (n >= ~32k regions).
//  * k_eval_wave  -- one region per wavefront.  The points of a region are
//    spread over the 64 lanes; packed generator codes are staged in LDS per
//    block, sums are folded with wave shuffles, and the fourth differences
//    are gathered from the lanes holding the axis points.  Used for short
//    lists, where per-region latency rather than throughput dominates.
//
// Both fuse the classification err > max(epsrel |est|, epsabs), the
// per-block partial sums and the last-block finalisation (reduce.hpp).
#pragma once

#include "integrands.hpp"
#include "reduce.hpp"

namespace cuhre {

template <int D>
constexpr int gm_npoints() { return (1 << D) + 2 * D * D + 2 * D + 1; }

// Packed generator of point k (order: centre; per axis i: +l2,-l2,+l3,-l3;
// per pair i<j: (+,+),(+,-),(-,+),(-,-) at l4; corners with coordinate m
// negative iff bit m of the corner index is set).
//   bits 0-2 type (0 centre, 1 axis l2, 2 axis l3, 3 pair, 4 corner)
//   bits 3-7 i, bits 8-12 j, bit 13 sign of i, bit 14 sign of j,
//   bits 16-31 corner index.
__device__ __forceinline__ uint32_t gm_code(int D, int k)
{
  if (k == 0) return 0u;
  k -= 1;
  if (k < 4 * D) {
    const int i = k >> 2, s = k & 3;
    return (uint32_t)((s < 2 ? 1 : 2) | (i << 3) | ((s & 1) << 13));
  }
  k -= 4 * D;
  const int npairs = D * (D - 1) / 2;
  if (k < 4 * npairs) {
    int q = k >> 2;
    const int s = k & 3;
    int i = 0;
    while (q >= D - 1 - i) {
      q -= D - 1 - i;
      ++i;
    }
    const int j = i + 1 + q;
    return (uint32_t)(3 | (i << 3) | (j << 8) | ((s >> 1) << 13) | ((s & 1) << 14));
  }
  k -= 4 * npairs;
  return (uint32_t)(4 | (k << 16));
}

template <int D, int FAM>
__global__ __launch_bounds__(kBlock) void k_eval_lane(EvalArgs A, DevIntegrand P, RuleConst R)
{
  using F = Family<FAM, D>;
  __shared__ Partial lds4[kBlock / kWave];
  const int64_t r = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  const bool valid = r < A.n;
  const int64_t rr = valid ? r : A.n - 1;

  double c[D], h[D], x[D];
  double vol = 1.0, maxw = 0.0;
  int widest = 0;
#pragma unroll
  for (int m = 0; m < D; ++m) {
    const double lo = A.lower[m * A.ld + rr];
    const double w = A.width[m * A.ld + rr];
    c[m] = fma(0.5, w, lo);
    h[m] = 0.5 * w;
    vol *= w;
    if (w > maxw) {
      maxw = w;
      widest = m;
    }
  }

  const double f0 = F::eval(P, c);
  double s7 = fma(R.w7c, f0, 0.0);
  double s5 = fma(R.w5c, f0, 0.0);
  double sa = fma(fabs(R.w7c), fabs(f0), 0.0);
  double maxd = 0.0;
  int ax = widest;
  double* const dvec = A.dvec;

  // axis points: four per axis, then the fourth difference of that axis
#pragma unroll 1
  for (int i = 0; i < D; ++i) {
    double fv[4];
#pragma unroll
    for (int s = 0; s < 4; ++s) {
      const double gv = s == 0 ? R.l2 : (s == 1 ? -R.l2 : (s == 2 ? R.l3 : -R.l3));
#pragma unroll
      for (int m = 0; m < D; ++m) x[m] = fma(h[m], m == i ? gv : 0.0, c[m]);
      fv[s] = F::eval(P, x);
      const double w7 = s < 2 ? R.w7a2 : R.w7a3;
      const double w5 = s < 2 ? R.w5a2 : R.w5a3;
      s7 = fma(w7, fv[s], s7);
      s5 = fma(w5, fv[s], s5);
      sa = fma(fabs(w7), fabs(fv[s]), sa);
    }
    const double t2 = fma(-2.0, f0, fv[0] + fv[1]);
    const double t3 = fma(-2.0, f0, fv[2] + fv[3]);
    const double di = fabs(fma(-R.ratio, t3, t2));
    if (dvec != nullptr && valid) dvec[i * A.ld + r] = di;
    if (di > maxd) {
      maxd = di;
      ax = i;
    }
  }

  // pair points at l4
#pragma unroll 1
  for (int i = 0; i < D - 1; ++i) {
#pragma unroll 1
    for (int j = i + 1; j < D; ++j) {
#pragma unroll
      for (int s = 0; s < 4; ++s) {
        const double gi = s < 2 ? R.l4 : -R.l4;
        const double gj = (s & 1) ? -R.l4 : R.l4;
#pragma unroll
        for (int m = 0; m < D; ++m) x[m] = fma(h[m], m == i ? gi : (m == j ? gj : 0.0), c[m]);
        const double v = F::eval(P, x);
        s7 = fma(R.w7p, v, s7);
        s5 = fma(R.w5p, v, s5);
        sa = fma(fabs(R.w7p), fabs(v), sa);
      }
    }
  }

  // 2^D corner points at l5 (degree-7 rule only)
#pragma unroll 1
  for (int k = 0; k < (1 << D); ++k) {
#pragma unroll
    for (int m = 0; m < D; ++m) x[m] = fma(h[m], ((k >> m) & 1) ? -R.l5 : R.l5, c[m]);
    const double v = F::eval(P, x);
    s7 = fma(R.w7v, v, s7);
    sa = fma(fabs(R.w7v), fabs(v), sa);
  }

  const double est = vol * s7;
  const double err = vol * fabs(s7 - s5);
  const double ab = vol * sa;
  if (valid) {
    A.est[r] = est;
    A.err[r] = err;
    A.axis[r] = (int8_t)ax;
    if (A.abssum) A.abssum[r] = ab;
    if (A.i5) A.i5[r] = vol * s5;
  }
  if (A.finalize) {
    Partial p;
    partial_from_region(p, est, err, ab, valid, A.epsrel, A.epsabs);
    block_reduce_partial(p, lds4);
    finish_iteration(A, p, lds4);
  }
}

template <int D, int FAM>
__global__ __launch_bounds__(kBlock) void k_eval_wave(EvalArgs A, DevIntegrand P, RuleConst R)
{
  using F = Family<FAM, D>;
  constexpr int NP = gm_npoints<D>();
  static_assert(4 * D + 1 <= kWave, "axis points must fit in one pass");
  __shared__ uint32_t code[NP];
  __shared__ Partial lds4[kBlock / kWave];
  for (int k = threadIdx.x; k < NP; k += kBlock) code[k] = gm_code(D, k);
  __syncthreads();

  const int lane = threadIdx.x & (kWave - 1);
  const int wave = threadIdx.x / kWave;
  const int64_t r = (int64_t)blockIdx.x * (kBlock / kWave) + wave;
  const bool valid = r < A.n;
  const int64_t rr = valid ? r : A.n - 1;

  double c[D], h[D], x[D];
  double vol = 1.0, maxw = 0.0;
  int widest = 0;
#pragma unroll
  for (int m = 0; m < D; ++m) {
    const double lo = A.lower[m * A.ld + rr];
    const double w = A.width[m * A.ld + rr];
    c[m] = fma(0.5, w, lo);
    h[m] = 0.5 * w;
    vol *= w;
    if (w > maxw) {
      maxw = w;
      widest = m;
    }
  }

  double s7 = 0.0, s5 = 0.0, sa = 0.0, fpass0 = 0.0;
  for (int k0 = 0; k0 < NP; k0 += kWave) {
    const int k = k0 + lane;
    if (k < NP) {
      const uint32_t cd = code[k];
      const int type = cd & 7, i = (cd >> 3) & 31, j = (cd >> 8) & 31;
      const bool si = (cd >> 13) & 1, sj = (cd >> 14) & 1;
      const uint32_t cidx = cd >> 16;
      const double lam = type == 1 ? R.l2 : (type == 2 ? R.l3 : R.l4);
      const double gi = si ? -lam : lam, gj = sj ? -lam : lam;
#pragma unroll
      for (int m = 0; m < D; ++m) {
        double g;
        if (type == 4) g = ((cidx >> m) & 1) ? -R.l5 : R.l5;
        else g = (type != 0 && m == i) ? gi : ((type == 3 && m == j) ? gj : 0.0);
        x[m] = fma(h[m], g, c[m]);
      }
      const double v = F::eval(P, x);
      const double w7 = type == 0 ? R.w7c : type == 1 ? R.w7a2 : type == 2 ? R.w7a3 : type == 3 ? R.w7p : R.w7v;
      const double w5 = type == 0 ? R.w5c : type == 1 ? R.w5a2 : type == 2 ? R.w5a3 : type == 3 ? R.w5p : 0.0;
      s7 = fma(w7, v, s7);
      s5 = fma(w5, v, s5);
      sa = fma(fabs(w7), fabs(v), sa);
      if (k0 == 0) fpass0 = v;
    }
  }
#pragma unroll
  for (int mask = 1; mask < kWave; mask <<= 1) {
    const double o7 = shfl_xor_d(s7, mask), o5 = shfl_xor_d(s5, mask), oa = shfl_xor_d(sa, mask);
    const bool lo_lane = (lane & mask) == 0;
    s7 = lo_lane ? s7 + o7 : o7 + s7;
    s5 = lo_lane ? s5 + o5 : o5 + s5;
    sa = lo_lane ? sa + oa : oa + sa;
  }

  // fourth differences: lane i < D gathers its axis points from pass 0
  const int ia = lane < D ? lane : 0;
  const double f0 = __shfl(fpass0, 0, kWave);
  const double fp2 = __shfl(fpass0, 1 + 4 * ia, kWave);
  const double fm2 = __shfl(fpass0, 2 + 4 * ia, kWave);
  const double fp3 = __shfl(fpass0, 3 + 4 * ia, kWave);
  const double fm3 = __shfl(fpass0, 4 + 4 * ia, kWave);
  const double t2 = fma(-2.0, f0, fp2 + fm2);
  const double t3 = fma(-2.0, f0, fp3 + fm3);
  const double di = lane < D ? fabs(fma(-R.ratio, t3, t2)) : -1.0;
  double mx = di;
#pragma unroll
  for (int mask = 1; mask < kWave; mask <<= 1) mx = fmax(mx, shfl_xor_d(mx, mask));
  const uint64_t hit = __ballot(lane < D && di == mx);
  int ax = widest;
  if (mx > 0.0 && hit != 0ull) ax = __ffsll((unsigned long long)hit) - 1;

  const double est = vol * s7;
  const double err = vol * fabs(s7 - s5);
  const double ab = vol * sa;
  if (valid) {
    if (lane == 0) {
      A.est[r] = est;
      A.err[r] = err;
      A.axis[r] = (int8_t)ax;
      if (A.abssum) A.abssum[r] = ab;
      if (A.i5) A.i5[r] = vol * s5;
    }
    if (A.dvec != nullptr && lane < D) A.dvec[lane * A.ld + r] = di;
  }
  if (A.finalize) {
    Partial p;
    partial_from_region(p, est, err, ab, valid && lane == 0, A.epsrel, A.epsabs);
    block_reduce_partial(p, lds4);
    finish_iteration(A, p, lds4);
  }
}

template <int D, int FAM, int L>
void launch_eval(const EvalArgs& A, const DevIntegrand& P, const RuleConst& R, hipStream_t st)
{
  constexpr int rpb = kBlock / L;
  const int64_t nb = (A.n + rpb - 1) / rpb;
  if constexpr (L == 1)
    hipLaunchKernelGGL((k_eval_lane<D, FAM>), dim3((unsigned)nb), dim3(kBlock), 0, st, A, P, R);
  else
    hipLaunchKernelGGL((k_eval_wave<D, FAM>), dim3((unsigned)nb), dim3(kBlock), 0, st, A, P, R);
}

template <int FAM, int D>
EvalLauncher pick_lanes(int lanes)
{
  return lanes == 1 ? &launch_eval<D, FAM, 1> : &launch_eval<D, FAM, kWave>;
}

// Launchers for one family over dimensions 2..10.
template <int FAM>
EvalLauncher family_launcher(int dim, int lanes)
{
  switch (dim) {
  case 2: return pick_lanes<FAM, 2>(lanes);
  case 3: return pick_lanes<FAM, 3>(lanes);
  case 4: return pick_lanes<FAM, 4>(lanes);
  case 5: return pick_lanes<FAM, 5>(lanes);
  case 6: return pick_lanes<FAM, 6>(lanes);
  case 7: return pick_lanes<FAM, 7>(lanes);
  case 8: return pick_lanes<FAM, 8>(lanes);
  case 9: return pick_lanes<FAM, 9>(lanes);
  case 10: return pick_lanes<FAM, 10>(lanes);
  default: return nullptr;
  }
}

}  // namespace cuhre
