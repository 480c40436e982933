// eval_kernels.hpp -- Genz-Malik degree-7/5 rule evaluation on gfx950.
//
// Replaces the reference's per-region evaluation: INTEGRATE_GPU_PHASE1
// (quad/GPUquad/GPUQuadPhases.cu:37-79), which runs SampleRegionBlock
// (quad/GPUquad/GPUQuadSample.cu:99-220) with one 256-thread block per
// region, five LDS tree reductions and a serial fourth-difference loop on
// thread 0.
//
// Mapping: one region per lane.  All 64 lanes of a wavefront walk the same
// point sequence for their own 64 regions, so every generator offset and
// weight is wave-uniform: the scalar unit selects it (s_cselect on the loop
// counters) and the FP64 VALU only does the coordinate FMAs, the integrand
// and the weighted sums.  A region needs no cross-lane reduction.
//
// Short lists: the point sequence is cut into S segments (S = 1..16) and a
// block runs S wavefronts over the same 64 regions, wave s taking the axes
// i = s (mod S), the pairs q = s (mod S) and the corners k = s (mod S).  The
// partial sums, fourth-difference maxima and argmax candidates of segments
// 1..S-1 meet segment 0 in LDS.  S is picked per iteration so that n*S lanes
// fill the chip; S = 1 once the list holds ~256k regions.  (A region per
// whole wavefront was measured at ~3 TFLOP/s against ~40 for this mapping,
// because every region then pays a cross-lane reduction and a per-lane
// generator decode.)
//
// The kernel also fuses the classification err > max(epsrel |est|, epsabs)
// and the per-block partial sums; k_finish (reduce.hpp) reduces them.
#pragma once

#include <hip/hip_ext.h>

#include "integrands.hpp"
#include "reduce.hpp"

namespace cuhre {

template <int D>
constexpr int gm_npoints() { return (1 << D) + 2 * D * D + 2 * D + 1; }

// LDS bytes a block needs to merge S segments (0 when S == 1).
inline size_t seg_lds_bytes(int S)
{
  if (S <= 1) return 0;
  const int W = S < 4 ? 4 : S;
  return (size_t)W * kWave * (4 * sizeof(double) + sizeof(int32_t));
}

inline int seg_waves(int S) { return S < 4 ? 4 : S; }

// Shortest lists: one region per half wavefront, its points in S = 32
// segments (lanes 0-31 / 32-63) merged by a xor butterfly, 8 regions per
// 256-thread block.  A few thousand regions then spread over every CU instead
// of stacking 16 segment waves of 64 regions on one CU.
constexpr int kHalfS = 32;

inline int seg_regions_per_block(int S)
{
  return S == kHalfS ? kBlock / kHalfS : kWave * (seg_waves(S) / S);
}

// The launch's list length: A.n, or the length on the device when the
// iteration was queued ahead of the host (A.n is then an upper bound and
// the blocks beyond the list store zero partials).
__device__ __forceinline__ int64_t eval_len(const EvalArgs& A) { return A.n_dev != nullptr ? *A.n_dev : A.n; }

// A block of a queued-ahead launch that lies wholly beyond the list (first:
// its first region) stores a zero partial and a zero active count and
// exits; when the list is empty (the loop stopped before this iteration)
// it writes nothing.
__device__ __forceinline__ bool block_beyond(const EvalArgs& A, int64_t n, int64_t first)
{
  if (A.n_dev == nullptr || first < n) return false;
  if (n > 0 && A.finalize && threadIdx.x == 0) {
    Partial z;
    partial_zero(z);
    A.partials[blockIdx.x] = z;
    A.offsets[blockIdx.x] = 0;
  }
  return true;
}

// Integrand of the plain kernel: one family, parameters in kernel arguments.
template <int D, int FAM>
struct PlainIntegrand {
  static constexpr bool kUnroll = Family<FAM, D>::kUnroll;
  const DevIntegrand& P;
  __device__ __forceinline__ double operator()(const double (&x)[D]) const
  {
    return Family<FAM, D>::eval(P, x);
  }
};

// Per-region outputs of seg_eval (valid in segment 0 of each group).
struct RegionOut {
  double est, err, ab, i5;
  int axis;
};

// Half-wave merges of the segment results (HALF = true below): segments
// s = lane mod 32 of one region sit in one half of a wavefront and meet by a
// xor butterfly, so every lane of the half ends with the same values (each
// level adds the same two operands in both partner lanes).
__device__ __forceinline__ double half_sum(double v)
{
#pragma unroll
  for (int o = 16; o > 0; o >>= 1) v += shfl_xor_d(v, o);
  return v;
}

// first strict argmax over increasing axis index; a lane whose largest
// difference is 0 holds axis -1, so partners always agree
__device__ __forceinline__ void half_argmax(double& d, int& a)
{
#pragma unroll
  for (int o = 16; o > 0; o >>= 1) {
    const double d2 = shfl_xor_d(d, o);
    const int a2 = __shfl_xor(a, o, kWave);
    if (d2 > d || (d2 == d && d2 > 0.0 && a2 < a)) {
      d = d2;
      a = a2;
    }
  }
}

// e[m] = h[m] at the coordinate i that a point set moves and 0 elsewhere.
// With i a per-lane value (the half-wavefront form), a point's coordinates
// x[m] = fma(e[m], g, c[m]) are c[m] + h[m] g where m == i and c[m] plus a
// signed zero, i.e. c[m] itself, elsewhere: the same values as
// fma(h[m], m == i ? g : 0, c[m]), with the index test paid once per set
// instead of at every coordinate of every point.  Nested for two or three
// moved coordinates (pairs, triples), the inner results are c[m] exactly
// wherever an outer coordinate moves.
template <int D>
__device__ __forceinline__ void moved_coord(double (&e)[D], const double (&h)[D], int i)
{
#pragma unroll
  for (int m = 0; m < D; ++m) e[m] = m == i ? h[m] : 0.0;
}

// The four axis points of axis i (a compile-time index once inlined in an
// unrolled loop: only x[i] leaves the centre) and that axis' fourth
// difference.
template <int D, class F>
__device__ __forceinline__ void gm_axis(const F& f, const double (&c)[D], const double (&h)[D], int i, double f0,
                                        const RuleConst& R, double& s7, double& s5, double& sa, double& maxd,
                                        int& ax, double* dvec, int64_t ld, int64_t r, bool valid)
{
  double x[D], fv[4];
#pragma unroll
  for (int m = 0; m < D; ++m) x[m] = c[m];
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    const double gv = t == 0 ? R.l2 : (t == 1 ? -R.l2 : (t == 2 ? R.l3 : -R.l3));
    x[i] = fma(h[i], gv, c[i]);
    fv[t] = f(x);
    const double w7 = t < 2 ? R.w7a2 : R.w7a3;
    const double w5 = t < 2 ? R.w5a2 : R.w5a3;
    s7 = fma(w7, fv[t], s7);
    s5 = fma(w5, fv[t], s5);
    sa = fma(fabs(w7), fabs(fv[t]), sa);
  }
  const double t2 = fma(-2.0, f0, fv[0] + fv[1]);
  const double t3 = fma(-2.0, f0, fv[2] + fv[3]);
  const double di = fabs(fma(-R.ratio, t3, t2));
  if (dvec != nullptr && valid) dvec[i * ld + r] = di;
  if (di > maxd) {
    maxd = di;
    ax = i;
  }
}

// Segmented Genz-Malik evaluation of the region held by this lane (see the
// file comment).  All waves of the block must call it together when S > 1
// (it contains a block barrier).  F(x) evaluates the lane's integrand.
// HALF: S = 32 segments of one region in the 32 lanes of a half wavefront
// (s = lane mod 32), merged by half_sum / half_argmax instead of LDS.
// MOVED (half form only): axis and pair points take their coordinates from
// per-set moved-coordinate vectors (moved_coord) instead of per-coordinate
// selects: fewer instructions, more registers.
template <int D, bool HALF = false, bool ONE = false, bool MOVED = HALF, class F>
__device__ __forceinline__ RegionOut seg_eval(const F& f, const double* __restrict__ lower,
                                              const double* __restrict__ width, int64_t ld,
                                              int64_t r, bool valid, int64_t rr, int S, int s,
                                              int wave, int lane, int W, double* seg_lds,
                                              const RuleConst& R, double* dvec)
{
  double c[D], h[D], x[D];
  double vol = 1.0, maxw = 0.0;
  int widest = 0;
#pragma unroll
  for (int m = 0; m < D; ++m) {
    const double lo = lower[m * ld + rr];
    const double w = width[m * ld + rr];
    c[m] = fma(0.5, w, lo);
    h[m] = 0.5 * w;
    vol *= w;
    if (w > maxw) {
      maxw = w;
      widest = m;
    }
  }

  // every segment needs f at the centre for its fourth differences; only
  // segment 0 adds its weight
  const double f0 = f(c);
  const double c0 = s == 0 ? 1.0 : 0.0;
  double s7 = fma(c0 * R.w7c, f0, 0.0);
  double s5 = fma(c0 * R.w5c, f0, 0.0);
  double sa = fma(c0 * fabs(R.w7c), fabs(f0), 0.0);
  double maxd = 0.0;
  int ax = -1;

  // Point coordinates: a coordinate that sits at the centre is c[m] itself
  // (equal to fma(h, 0, c), c never being -0), so each point computes only
  // the coordinates its generator moves.  With one segment (ONE) the axis
  // points are straight-line code whose moved coordinate is a compile-time
  // index, so an integrand that accumulates over the coordinates in order
  // shares its partial results with the centre's (the compiler's common
  // subexpressions); corners run in groups of 16 whose last four
  // coordinates are compile-time, so such an integrand's partial result over
  // the group's first coordinates is computed once per group.  (Pair points
  // unrolled the same way needed ~380 registers.)  Families without kUnroll
  // keep runtime point loops (x = c + h * g at every point), as does the
  // half-wavefront form for its axes and pairs.

  // axis points: four per axis, then the fourth difference of that axis
  constexpr bool COMPACT = HALF || !F::kUnroll;  // runtime point loops
  if constexpr (COMPACT) {
    // axes i = s, s + S, ... (in a half, s varies across its lanes)
#pragma unroll 1
    for (int i = s; i < D; i += S) {
      double fv[4], e[D];
      if constexpr (MOVED) moved_coord(e, h, i);
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        const double gv = t == 0 ? R.l2 : (t == 1 ? -R.l2 : (t == 2 ? R.l3 : -R.l3));
#pragma unroll
        for (int m = 0; m < D; ++m) {
          if constexpr (MOVED)
            x[m] = fma(e[m], gv, c[m]);
          else
            x[m] = fma(h[m], m == i ? gv : 0.0, c[m]);
        }
        fv[t] = f(x);
        const double w7 = t < 2 ? R.w7a2 : R.w7a3;
        const double w5 = t < 2 ? R.w5a2 : R.w5a3;
        s7 = fma(w7, fv[t], s7);
        s5 = fma(w5, fv[t], s5);
        sa = fma(fabs(w7), fabs(fv[t]), sa);
      }
      const double t2 = fma(-2.0, f0, fv[0] + fv[1]);
      const double t3 = fma(-2.0, f0, fv[2] + fv[3]);
      const double di = fabs(fma(-R.ratio, t3, t2));
      if (dvec != nullptr && valid) dvec[i * ld + r] = di;
      if (di > maxd) {
        maxd = di;
        ax = i;
      }
    }
  } else if (ONE) {
#pragma unroll
    for (int i = 0; i < D; ++i) gm_axis<D>(f, c, h, i, f0, R, s7, s5, sa, maxd, ax, dvec, ld, r, valid);
  } else {
    // wave-uniform segment: axes i = s, s + S, ... (i compile-time in the
    // unrolled loop)
    int next = s;
#pragma unroll
    for (int i = 0; i < D; ++i) {
      if (i != next) continue;
      next += S;
      gm_axis<D>(f, c, h, i, f0, R, s7, s5, sa, maxd, ax, dvec, ld, r, valid);
    }
  }

  // pair points at l4, pair q = (i, j) in the order (0,1),(0,2),...,(D-2,D-1),
  // each pair's four points in the order (+,+), (+,-), (-,+), (-,-) of (x_i, x_j)
  constexpr int NPAIR = D * (D - 1) / 2;
  if constexpr (COMPACT) {
#pragma unroll 2
    for (int q = s; q < NPAIR; q += S) {
      int i = 0, rem = q;
      while (rem >= D - 1 - i) {
        rem -= D - 1 - i;
        ++i;
      }
      const int j = i + 1 + rem;
      double ei[D], ej[D];
      if constexpr (MOVED) {
        moved_coord(ei, h, i);
        moved_coord(ej, h, j);
      }
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        const double gi = t < 2 ? R.l4 : -R.l4;
        const double gj = (t & 1) ? -R.l4 : R.l4;
#pragma unroll
        for (int m = 0; m < D; ++m) {
          if constexpr (MOVED)
            x[m] = fma(ei[m], gi, fma(ej[m], gj, c[m]));
          else
            x[m] = fma(h[m], m == i ? gi : (m == j ? gj : 0.0), c[m]);
        }
        const double v = f(x);
        s7 = fma(R.w7p, v, s7);
        s5 = fma(R.w5p, v, s5);
        sa = fma(fabs(R.w7p), fabs(v), sa);
      }
    }
  } else {
    // rows i (runtime, wave-uniform) hold the row vectors xp / xm (x_i at
    // +-l4, the rest at the centre); j is a compile-time index, so each point
    // computes only x_j
    int next = s;  // next pair index of this segment
    int q0 = 0;    // index of the pair (i, i + 1)
#pragma unroll 1
    for (int i = 0; i < D - 1; ++i) {
      const int row = D - 1 - i;
      if (next < q0 + row) {
        double xp[D], xm[D];
#pragma unroll
        for (int m = 0; m < D; ++m) {
          const double g = m == i ? R.l4 : 0.0;
          xp[m] = fma(h[m], g, c[m]);
          xm[m] = fma(h[m], -g, c[m]);
        }
#pragma unroll
        for (int j = 1; j < D; ++j) {
          if (j <= i || q0 + (j - i - 1) != next) continue;
          next += S;
          const double xjp = fma(h[j], R.l4, c[j]);
          const double xjm = fma(h[j], -R.l4, c[j]);
#pragma unroll
          for (int t = 0; t < 4; ++t) {
#pragma unroll
            for (int m = 0; m < D; ++m) x[m] = t < 2 ? xp[m] : xm[m];
            x[j] = (t & 1) ? xjm : xjp;
            const double v = f(x);
            s7 = fma(R.w7p, v, s7);
            s5 = fma(R.w5p, v, s5);
            sa = fma(fabs(R.w7p), fabs(v), sa);
          }
        }
      }
      q0 += row;
    }
  }

  // 2^D corner points at l5 (degree-7 rule only)
  if constexpr (!F::kUnroll) {
    // corner k has x_m at -l5 where bit m of k is set; k = s, s + S, ...
#pragma unroll 2
    for (int k = s; k < (1 << D); k += S) {
#pragma unroll
      for (int m = 0; m < D; ++m) x[m] = fma(h[m], ((k >> m) & 1) ? -R.l5 : R.l5, c[m]);
      const double v = f(x);
      s7 = fma(R.w7v, v, s7);
      sa = fma(fabs(R.w7v), fabs(v), sa);
    }
  } else if constexpr (HALF) {
    // corner k has x_m at -l5 where bit m of k is set; corners k = s, s + 32,
    // ...: bits 0..4 of k are the lane's s, the next (up to three) are
    // unrolled
    if constexpr (D <= 5) {
      if (s < (1 << D)) {
#pragma unroll
        for (int m = 0; m < D; ++m) x[m] = fma(h[m], ((s >> m) & 1) ? -R.l5 : R.l5, c[m]);
        const double v = f(x);
        s7 = fma(R.w7v, v, s7);
        sa = fma(fabs(R.w7v), fabs(v), sa);
      }
    } else {
      constexpr int LT = D - 5;              // bits of k above the lane's five
      constexpr int LU = LT < 3 ? LT : 3;    // of which unrolled
      double cpu[LU], cmu[LU];
#pragma unroll
      for (int m = 0; m < 5; ++m) x[m] = fma(h[m], ((s >> m) & 1) ? -R.l5 : R.l5, c[m]);
#pragma unroll
      for (int b = 0; b < LU; ++b) {
        cpu[b] = fma(h[5 + b], R.l5, c[5 + b]);
        cmu[b] = fma(h[5 + b], -R.l5, c[5 + b]);
      }
#pragma unroll 1
      for (int g = 0; g < (1 << (LT - LU)); ++g) {
#pragma unroll
        for (int m = 5 + LU; m < D; ++m)
          x[m] = fma(h[m], ((g >> (m - 5 - LU)) & 1) ? -R.l5 : R.l5, c[m]);
#pragma unroll
        for (int u = 0; u < (1 << LU); ++u) {
#pragma unroll
          for (int b = 0; b < LU; ++b) x[5 + b] = ((u >> b) & 1) ? cmu[b] : cpu[b];
          const double v = f(x);
          s7 = fma(R.w7v, v, s7);
          sa = fma(fabs(R.w7v), fabs(v), sa);
        }
      }
    }
  } else {
    // corners in lexicographic order of their signs, the last coordinate
    // fastest: corner k has x_m at -l5 where bit D-1-m of k is set.  Groups
    // of 2^LU corners (group g = s, s + S, ...) share the first D - LU
    // coordinates; the last LU come from precomputed c +- l5 h.
    constexpr int LU = D < 4 ? D : 4;
    constexpr int M0 = D - LU;  // first unrolled coordinate
    double cpu[LU], cmu[LU];
#pragma unroll
    for (int b = 0; b < LU; ++b) {
      cpu[b] = fma(h[M0 + b], R.l5, c[M0 + b]);
      cmu[b] = fma(h[M0 + b], -R.l5, c[M0 + b]);
    }
#pragma unroll 1
    for (int g = ONE ? 0 : s; g < (1 << M0); g += ONE ? 1 : S) {
#pragma unroll
      for (int m = 0; m < M0; ++m) x[m] = fma(h[m], ((g >> (M0 - 1 - m)) & 1) ? -R.l5 : R.l5, c[m]);
#pragma unroll
      for (int u = 0; u < (1 << LU); ++u) {
#pragma unroll
        for (int b = 0; b < LU; ++b) x[M0 + b] = ((u >> (LU - 1 - b)) & 1) ? cmu[b] : cpu[b];
        const double v = f(x);
        s7 = fma(R.w7v, v, s7);
        sa = fma(fabs(R.w7v), fabs(v), sa);
      }
    }
  }

  if constexpr (HALF) {
    s7 = half_sum(s7);
    s5 = half_sum(s5);
    sa = half_sum(sa);
    half_argmax(maxd, ax);
  } else if (S > 1) {
    // segments 1..S-1 hand their partial sums to segment 0 of their group
    double* const sh = seg_lds;  // [W][4][64] doubles, then [W][64] ints
    int32_t* const shax = reinterpret_cast<int32_t*>(seg_lds + (size_t)W * 4 * kWave);
    if (s != 0) {
      sh[(wave * 4 + 0) * kWave + lane] = s7;
      sh[(wave * 4 + 1) * kWave + lane] = s5;
      sh[(wave * 4 + 2) * kWave + lane] = sa;
      sh[(wave * 4 + 3) * kWave + lane] = maxd;
      shax[wave * kWave + lane] = ax;
    }
    __syncthreads();
    if (s == 0) {
      for (int t = 1; t < S; ++t) {
        const int w2 = wave + t;
        s7 += sh[(w2 * 4 + 0) * kWave + lane];
        s5 += sh[(w2 * 4 + 1) * kWave + lane];
        sa += sh[(w2 * 4 + 2) * kWave + lane];
        const double d2 = sh[(w2 * 4 + 3) * kWave + lane];
        const int a2 = shax[w2 * kWave + lane];
        // first strict argmax over increasing axis index
        if (d2 > maxd || (d2 == maxd && d2 > 0.0 && a2 < ax)) {
          maxd = d2;
          ax = a2;
        }
      }
    }
    __syncthreads();  // the LDS block may be reused by the caller
  }
  if (ax < 0) ax = widest;  // every fourth difference is zero

  RegionOut o;
  o.est = vol * s7;
  o.err = vol * fabs(s7 - s5);
  o.ab = vol * sa;
  o.i5 = vol * s5;
  o.axis = ax;
  return o;
}

template <int D, int FAM>
__global__ __launch_bounds__(1024) void k_eval_seg(EvalArgs A, DevIntegrand P, RuleConst R, int S)
{
  extern __shared__ double seg_lds[];
  const int W = S < 4 ? 4 : S;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int lane = threadIdx.x & (kWave - 1);
  const int g = wave / S, s = wave - g * S;
  const int64_t r = (int64_t)blockIdx.x * (kWave * (W / S)) + g * kWave + lane;
  const int64_t n = eval_len(A);
  if (block_beyond(A, n, (int64_t)blockIdx.x * (kWave * (W / S)))) return;
  const bool valid = r < n;
  const int64_t rr = valid ? r : (n > 0 ? n - 1 : 0);
  const PlainIntegrand<D, FAM> f{P};
  const RegionOut o = seg_eval<D>(f, A.lower, A.width, A.ld, r, valid, rr, S, s, wave, lane, W,
                                  seg_lds, R, A.dvec);
  const bool owner = s == 0;
  if (valid && owner) {
    A.est[r] = o.est;
    A.err[r] = o.err;
    A.axis[r] = (int8_t)o.axis;
    if (A.abssum) A.abssum[r] = o.ab;
    if (A.i5) A.i5[r] = o.i5;
  }
  if (A.finalize) {
    Partial p;
    block_reduce_regions(owner, g * kWave + lane, kWave * (W / S), o.est, o.err, o.ab, valid, A.loc_epsrel,
                         A.loc_epsabs, p);
    store_block_partial(A, p);
  }
}

// One segment (long lists): 256-thread blocks, a region per lane, the point
// loops straight-line code (seg_eval with ONE).  Its own launch bounds give
// the compiler the registers that code needs.
template <int D, int FAM>
__global__ __launch_bounds__(kBlock) void k_eval_one(EvalArgs A, DevIntegrand P, RuleConst R)
{
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int lane = threadIdx.x & (kWave - 1);
  const int64_t r = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  const int64_t n = eval_len(A);
  if (block_beyond(A, n, (int64_t)blockIdx.x * kBlock)) return;
  const bool valid = r < n;
  const int64_t rr = valid ? r : (n > 0 ? n - 1 : 0);
  const PlainIntegrand<D, FAM> f{P};
  const RegionOut o = seg_eval<D, false, true>(f, A.lower, A.width, A.ld, r, valid, rr, 1, 0, wave, lane,
                                               kBlock / kWave, nullptr, R, A.dvec);
  if (valid) {
    A.est[r] = o.est;
    A.err[r] = o.err;
    A.axis[r] = (int8_t)o.axis;
    if (A.abssum) A.abssum[r] = o.ab;
    if (A.i5) A.i5[r] = o.i5;
  }
  if (A.finalize) {
    Partial p;
    block_reduce_regions(true, threadIdx.x, kBlock, o.est, o.err, o.ab, valid, A.loc_epsrel, A.loc_epsabs, p);
    store_block_partial(A, p);
  }
}

// Half-wavefront form (S = kHalfS), defined after the degree-9 evaluator.
template <int D, int FAM, int KEY>
void launch_half(const EvalArgs& A, const DevIntegrand& P, const RuleConst& R, hipStream_t st);

template <int D, int FAM>
void launch_eval(const EvalArgs& A, const DevIntegrand& P, const RuleConst& R, int S,
                 hipStream_t st)
{
  if (S == kHalfS) return launch_half<D, FAM, 7>(A, P, R, st);
  const int rpb = seg_regions_per_block(S);
  const int64_t nb = (A.n + rpb - 1) / rpb;
  if (S == 1) {
    hipExtLaunchKernelGGL((k_eval_one<D, FAM>), dim3((unsigned)nb), dim3(kBlock), 0, st, A.prof_start,
                          A.prof_stop, 0, A, P, R);
    return;
  }
  hipExtLaunchKernelGGL((k_eval_seg<D, FAM>), dim3((unsigned)nb), dim3(kWave * seg_waves(S)),
                        (uint32_t)seg_lds_bytes(S), st, A.prof_start, A.prof_stop, 0, A, P, R, S);
}

// ---------------------------------------------------------------- degree 9
// The reference's degree-9 rule (KEY 9; quad/GPUquad/GPUQuadRule.cu:30-226,
// evaluated per region by quad/GPUquad/GPUQuadSample.cu:99-220), on the same
// lane-per-region mapping.  The point sets are walked set by set: the
// centre; per axis i the eight points +-t1..t4 along i (whose first two
// radii give the fourth difference); per pair i<j the four (+-t1, +-t1) and
// the eight (+-t1, +-t2) / (+-t2, +-t1) points; per triple i<j<k the eight
// (+-t1)^3 points; the 2^d corners at +-t5.  f is summed per set group and
// the group sums enter the five rule sums with the set's weights.  Segments
// take axes, pairs, triples and corners round robin as in seg_eval.

inline size_t seg9_lds_bytes(int S)
{
  if (S <= 1) return 0;
  const int W = S < 4 ? 4 : S;
  return (size_t)W * kWave * (7 * sizeof(double) + sizeof(int32_t));
}

template <int D, bool HALF = false, class F>
__device__ __forceinline__ RegionOut seg_eval9(const F& f, const double* __restrict__ lower,
                                               const double* __restrict__ width, int64_t ld, int64_t r,
                                               bool valid, int64_t rr, int S, int s, int wave, int lane, int W,
                                               double* seg_lds, const Rule9Const& R, double* dvec)
{
  double c[D], h[D], x[D];
  double vol = 1.0, maxw = 0.0;
  int widest = 0;
#pragma unroll
  for (int m = 0; m < D; ++m) {
    const double lo = lower[m * ld + rr];
    const double w = width[m * ld + rr];
    c[m] = fma(0.5, w, lo);
    h[m] = 0.5 * w;
    vol *= w;
    if (w > maxw) {
      maxw = w;
      widest = m;
    }
  }
  double sum[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
  double sa = 0.0;
  // adds one set group (sum of f, sum of |f|) to the rule sums
  auto group = [&](int set, double fs, double fa) {
#pragma unroll
    for (int q = 0; q < 5; ++q) sum[q] = fma(R.w[set][q], fs, sum[q]);
    sa = fma(fabs(R.w[set][0]), fa, sa);
  };
  const double f0 = f(c);
  if (s == 0) group(0, f0, fabs(f0));
  double maxd = 0.0;
  int ax = -1;

  // axis sets 1..4
#pragma unroll 1
  for (int i = s; i < D; i += S) {
    double fv[2][2], e[D];
    if constexpr (HALF) moved_coord(e, h, i);
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      double fs = 0.0, fa = 0.0;
#pragma unroll
      for (int sg = 0; sg < 2; ++sg) {
        const double gv = sg ? -R.t[k] : R.t[k];
#pragma unroll
        for (int m = 0; m < D; ++m) {
          if constexpr (HALF)
            x[m] = fma(e[m], gv, c[m]);
          else
            x[m] = fma(h[m], m == i ? gv : 0.0, c[m]);
        }
        const double v = f(x);
        if (k < 2) fv[k][sg] = v;
        fs += v;
        fa += fabs(v);
      }
      group(1 + k, fs, fa);
    }
    // (products and sums fused explicitly, so that every kernel built on
    // this evaluator rounds them alike)
    const double di = fabs(fma(R.ratio, fv[0][0] + fv[0][1], 2.0 * (1.0 - R.ratio) * f0) - (fv[1][0] + fv[1][1]));
    if (dvec != nullptr && valid) dvec[i * ld + r] = di;
    if (di > maxd) {
      maxd = di;
      ax = i;
    }
  }

  // pair sets 5 and 6, pair q = (i, j) in the order (0,1),(0,2),...,(D-2,D-1)
  constexpr int NPAIR = D * (D - 1) / 2;
#pragma unroll 1
  for (int q = s; q < NPAIR; q += S) {
    int i = 0, rem = q;
    while (rem >= D - 1 - i) {
      rem -= D - 1 - i;
      ++i;
    }
    const int j = i + 1 + rem;
    double ei[D], ej[D];
    if constexpr (HALF) {
      moved_coord(ei, h, i);
      moved_coord(ej, h, j);
    }
    // x at (gi, gj) on axes (i, j), the centre elsewhere
    const auto pair_point = [&](double gi, double gj) {
#pragma unroll
      for (int m = 0; m < D; ++m) {
        if constexpr (HALF)
          x[m] = fma(ei[m], gi, fma(ej[m], gj, c[m]));
        else
          x[m] = fma(h[m], m == i ? gi : (m == j ? gj : 0.0), c[m]);
      }
    };
    double fs = 0.0, fa = 0.0;
#pragma unroll
    for (int sg = 0; sg < 4; ++sg) {
      const double gi = (sg & 1) ? -R.t[0] : R.t[0], gj = (sg & 2) ? -R.t[0] : R.t[0];
      pair_point(gi, gj);
      const double v = f(x);
      fs += v;
      fa += fabs(v);
    }
    group(5, fs, fa);
    fs = 0.0;
    fa = 0.0;
#pragma unroll
    for (int o = 0; o < 2; ++o)
#pragma unroll
      for (int sg = 0; sg < 4; ++sg) {
        const double ti = o ? R.t[1] : R.t[0], tj = o ? R.t[0] : R.t[1];
        const double gi = (sg & 1) ? -ti : ti, gj = (sg & 2) ? -tj : tj;
        pair_point(gi, gj);
        const double v = f(x);
        fs += v;
        fa += fabs(v);
      }
    group(6, fs, fa);
  }

  // triple set 7, triples in lexicographic order
  constexpr int NTRI = D * (D - 1) * (D - 2) / 6;
#pragma unroll 1
  for (int q = s; q < NTRI; q += S) {
    int i = 0, rem = q;
    while (rem >= (D - 1 - i) * (D - 2 - i) / 2) {
      rem -= (D - 1 - i) * (D - 2 - i) / 2;
      ++i;
    }
    int j = i + 1;
    while (rem >= D - 1 - j) {
      rem -= D - 1 - j;
      ++j;
    }
    const int k = j + 1 + rem;
    double ei[D], ej[D], ek[D];
    if constexpr (HALF) {
      moved_coord(ei, h, i);
      moved_coord(ej, h, j);
      moved_coord(ek, h, k);
    }
    double fs = 0.0, fa = 0.0;
#pragma unroll 2
    for (int sg = 0; sg < 8; ++sg) {
      const double gi = (sg & 1) ? -R.t[0] : R.t[0], gj = (sg & 2) ? -R.t[0] : R.t[0],
                   gk = (sg & 4) ? -R.t[0] : R.t[0];
#pragma unroll
      for (int m = 0; m < D; ++m) {
        if constexpr (HALF)
          x[m] = fma(ei[m], gi, fma(ej[m], gj, fma(ek[m], gk, c[m])));
        else
          x[m] = fma(h[m], m == i ? gi : (m == j ? gj : (m == k ? gk : 0.0)), c[m]);
      }
      const double v = f(x);
      fs += v;
      fa += fabs(v);
    }
    group(7, fs, fa);
  }

  // corner set 8
  {
    double fs = 0.0, fa = 0.0;
    if constexpr (HALF && D > 5) {
      // corners k = s, s + 32, ... in increasing order: bits 0..4 of k are
      // the lane's s (those coordinates are set once), the next (up to
      // three) are unrolled, so an integrand that accumulates over the
      // coordinates in order computes its partial result over the first five
      // once for the unrolled corners
      constexpr int LT = D - 5;            // bits of k above the lane's five
      constexpr int LU = LT < 3 ? LT : 3;  // of which unrolled
      double cpu[LU], cmu[LU];
#pragma unroll
      for (int m = 0; m < 5; ++m) x[m] = fma(h[m], ((s >> m) & 1) ? -R.t[4] : R.t[4], c[m]);
#pragma unroll
      for (int b = 0; b < LU; ++b) {
        cpu[b] = fma(h[5 + b], R.t[4], c[5 + b]);
        cmu[b] = fma(h[5 + b], -R.t[4], c[5 + b]);
      }
#pragma unroll 1
      for (int g = 0; g < (1 << (LT - LU)); ++g) {
#pragma unroll
        for (int m = 5 + LU; m < D; ++m) x[m] = fma(h[m], ((g >> (m - 5 - LU)) & 1) ? -R.t[4] : R.t[4], c[m]);
#pragma unroll
        for (int u = 0; u < (1 << LU); ++u) {
#pragma unroll
          for (int b = 0; b < LU; ++b) x[5 + b] = ((u >> b) & 1) ? cmu[b] : cpu[b];
          const double v = f(x);
          fs += v;
          fa += fabs(v);
        }
      }
    } else {
#pragma unroll 2
      for (int k = s; k < (1 << D); k += S) {
#pragma unroll
        for (int m = 0; m < D; ++m) x[m] = fma(h[m], ((k >> m) & 1) ? -R.t[4] : R.t[4], c[m]);
        const double v = f(x);
        fs += v;
        fa += fabs(v);
      }
    }
    group(8, fs, fa);
  }

  if constexpr (HALF) {
#pragma unroll
    for (int q = 0; q < 5; ++q) sum[q] = half_sum(sum[q]);
    sa = half_sum(sa);
    half_argmax(maxd, ax);
  } else if (S > 1) {
    double* const sh = seg_lds;  // [W][7][64] doubles, then [W][64] ints
    int32_t* const shax = reinterpret_cast<int32_t*>(seg_lds + (size_t)W * 7 * kWave);
    if (s != 0) {
#pragma unroll
      for (int q = 0; q < 5; ++q) sh[(wave * 7 + q) * kWave + lane] = sum[q];
      sh[(wave * 7 + 5) * kWave + lane] = sa;
      sh[(wave * 7 + 6) * kWave + lane] = maxd;
      shax[wave * kWave + lane] = ax;
    }
    __syncthreads();
    if (s == 0) {
      for (int t = 1; t < S; ++t) {
        const int w2 = wave + t;
#pragma unroll
        for (int q = 0; q < 5; ++q) sum[q] += sh[(w2 * 7 + q) * kWave + lane];
        sa += sh[(w2 * 7 + 5) * kWave + lane];
        const double d2 = sh[(w2 * 7 + 6) * kWave + lane];
        const int a2 = shax[w2 * kWave + lane];
        if (d2 > maxd || (d2 == maxd && d2 > 0.0 && a2 < ax)) {
          maxd = d2;
          ax = a2;
        }
      }
    }
    __syncthreads();
  }
  if (ax < 0) ax = widest;

  // the largest normalised null rule in the span of two successive ones
  // (GPUQuadSample.cu:195-216, errcoeff {5, 1, 5})
  double e[4];
#pragma unroll
  for (int q = 1; q < 4; ++q) {
    double mx = 0.0;
#pragma unroll
    for (int t = 0; t < 9; ++t) mx = fmax(mx, fabs(fma(R.scale[t][q], sum[q], sum[q + 1])) * R.norm[t][q]);
    e[q] = mx;
  }
  RegionOut o;
  o.est = vol * sum[0];
  o.err = vol * ((5.0 * e[1] <= e[2] && 5.0 * e[2] <= e[3]) ? e[1] : 5.0 * fmax(fmax(e[1], e[2]), e[3]));
  o.ab = vol * sa;
  o.i5 = 0.0;
  o.axis = ax;
  return o;
}

// Outputs of a half-wavefront evaluation block b (8 regions, owner lane s
// = 0 of each): the per-region results, optionally the block's list of
// active regions (the split's test) in order, entry i = (position in the
// block << 4) | split axis of its i-th active region, and the block's
// partial sums and dense active count.
__device__ __forceinline__ void half_outputs(const EvalArgs& A, const RegionOut& o, int64_t r, bool valid,
                                             bool owner, uint32_t b)
{
  if (valid && owner) {
    A.est[r] = o.est;
    A.err[r] = o.err;
    A.axis[r] = (int8_t)o.axis;
    if (A.abssum) A.abssum[r] = o.ab;
    if (A.i5) A.i5[r] = o.i5;
  }
  if (A.act_out != nullptr) {
    __shared__ uint32_t s_mask;
    if (threadIdx.x == 0) s_mask = 0;
    __syncthreads();
    const int j = threadIdx.x / kHalfS;
    const bool act =
        valid && owner && isfinite(o.est) && isfinite(o.err) && region_active(o.est, o.err, A.epsrel, A.epsabs);
    if (act) atomicOr(&s_mask, 1u << j);
    __syncthreads();
    if (act) A.act_out[b * (kBlock / kHalfS) + __popc(s_mask & ((1u << j) - 1u))] = (uint8_t)((j << 4) | o.axis);
  }
  if (A.finalize) {
    Partial p;
    block_reduce_regions(owner, threadIdx.x / kHalfS, kBlock / kHalfS, o.est, o.err, o.ab, valid, A.loc_epsrel,
                         A.loc_epsabs, p);
    if (threadIdx.x == 0) {
      A.partials[b] = p;
      A.offsets[b] = p.nactive + (p.nnonfinite > 0 ? kBadFlag : 0);
    }
  }
}

// Half-wavefront form (S = kHalfS): 8 regions per 256-thread block.
template <int D, int FAM, int KEY>
__global__ __launch_bounds__(256) void k_eval_half(EvalArgs A, DevIntegrand P, RuleConst R)
{
  const int lane = threadIdx.x & (kWave - 1);
  const int s = threadIdx.x & (kHalfS - 1);
  const int64_t r = ((int64_t)blockIdx.x * kBlock + threadIdx.x) / kHalfS;
  const int64_t n = eval_len(A);
  if (block_beyond(A, n, (int64_t)blockIdx.x * (kBlock / kHalfS))) return;
  const bool valid = r < n;
  const int64_t rr = valid ? r : (n > 0 ? n - 1 : 0);
  const PlainIntegrand<D, FAM> f{P};
  RegionOut o;
  if constexpr (KEY == 9)
    o = seg_eval9<D, true>(f, A.lower, A.width, A.ld, r, valid, rr, kHalfS, s, 0, lane, 0, nullptr, R.r9, A.dvec);
  else
    o = seg_eval<D, true>(f, A.lower, A.width, A.ld, r, valid, rr, kHalfS, s, 0, lane, 0, nullptr, R, A.dvec);
  half_outputs(A, o, r, valid, s == 0, blockIdx.x);
}

template <int D, int FAM, int KEY>
void launch_half(const EvalArgs& A, const DevIntegrand& P, const RuleConst& R, hipStream_t st)
{
  const int64_t nb = (A.n + kBlock / kHalfS - 1) / (kBlock / kHalfS);
  hipExtLaunchKernelGGL((k_eval_half<D, FAM, KEY>), dim3((unsigned)nb), dim3(kBlock), 0, st, A.prof_start,
                        A.prof_stop, 0, A, P, R);
}

// The half-wavefront evaluation of a list whose regions are the children of
// the previous list's active regions, found here instead of by a separate
// split launch (lists of at most 8192 regions queued ahead of the host: one
// launch per iteration instead of two).  Block 0 finalises the previous
// iteration as k_split_fin's extra block does (its partials, the stop
// decision, the published state, the next list's slot).  Every other block
// scans the previous iteration's dense active counts, takes the same count
// test as the split (budget, buffers, non-finite values; the loop then
// stops and the block exits), and for each of its 8 regions r finds the
// parent: the (r/2)-th active region, in the previous block whose count
// scan covers r/2, at the position (and with the split axis) that block's
// list of active regions gives.  It bisects the
// parent along its split axis exactly as the split does (child r/2*2 takes
// the lower half), writes the child's bounds to this list, and evaluates
// the child from those values; results, active lists and partials follow
// as in k_eval_half.  Regions, their order and every result are those of
// the split followed by k_eval_half.
template <int D, int FAM>
__global__ __launch_bounds__(256) void k_eval_fused(EvalArgs A, FuseArgs F, DevIntegrand P, RuleConst R)
{
  __shared__ int64_t s_base[kOneLevel];
  __shared__ int64_t s_wsum[kBlock / kWave], s_wbad[kBlock / kWave];
  const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
  const int s = threadIdx.x & (kHalfS - 1);
  constexpr int kRpb = kBlock / kHalfS;  // regions per block, both lists
  EvalArgs RA = F.R;
  int64_t pn = RA.n;
  if (F.P.in != nullptr) {
    pn = F.P.in->n;
    if (pn == 0) return;  // the loop stopped before the previous iteration
    RA.fin_est_s = F.P.in->fin_s;
    RA.fin_est_c = F.P.in->fin_c;
    RA.fin_err_s = F.P.in->fer_s;
    RA.fin_err_c = F.P.in->fer_c;
    const int64_t left = F.P.left0 - pn * F.P.npts;
    const int64_t budget = left >= 0 ? left / (2 * F.P.npts) : -1;
    RA.max_active = budget < F.P.cap2 ? budget : F.P.cap2;
  }
  const uint32_t pnb = (uint32_t)((pn + kRpb - 1) / kRpb);
  if (blockIdx.x == 0) {
    Partial q;
    reduce_partials(RA.partials, pnb, nullptr, q);
    if (threadIdx.x == 0) {
      const IterState st = iteration_state(RA, q);
      write_state(RA, st);
      if (F.P.out != nullptr) {
        SpecSlot o;
        o.n = st.stop == 0 ? 2 * q.nactive : 0;
        o.pad = 0;
        o.fin_s = RA.fin_est_s;
        o.fin_c = RA.fin_est_c;
        o.fer_s = RA.fin_err_s;
        o.fer_c = RA.fin_err_c;
        dd_add(o.fin_s, o.fin_c, q.fest_s, q.fest_c);
        dd_add(o.fer_s, o.fer_c, q.ferr_s, q.ferr_c);
        *F.P.out = o;
      }
    }
    return;
  }
  const uint32_t b = blockIdx.x - 1;
  // exclusive scan of the previous iteration's dense counts, four per thread
  int64_t cnt[4], mine = 0, bad = 0;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const uint32_t j = 4 * threadIdx.x + i;
    cnt[i] = 0;
    if (j < pnb) {
      const int64_t v = RA.offsets[j];
      cnt[i] = v & kCountMask;
      bad += v >> 40;
    }
    mine += cnt[i];
  }
  long long incl = (long long)mine, nbad = (long long)bad;
#pragma unroll
  for (int off = 1; off < kWave; off <<= 1) {
    const long long v = __shfl_up(incl, off, kWave);
    if (lane >= off) incl += v;
  }
#pragma unroll
  for (int mask = 1; mask < kWave; mask <<= 1) nbad += __shfl_xor(nbad, mask, kWave);
  if (lane == kWave - 1) s_wsum[wave] = (int64_t)incl;
  if (lane == 0) s_wbad[wave] = (int64_t)nbad;
  __syncthreads();
  int64_t before = 0, nact = 0, nonfinite = 0;
  for (int w = 0; w < kBlock / kWave; ++w) {
    if (w < wave) before += s_wsum[w];
    nact += s_wsum[w];
    nonfinite += s_wbad[w];
  }
  int64_t run = before + (int64_t)incl - mine;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    if (4 * threadIdx.x + i < pnb) s_base[4 * threadIdx.x + i] = run;
    run += cnt[i];
  }
  __syncthreads();
  if (nonfinite != 0 || nact == 0 || nact > RA.max_active || nact > RA.alloc_active) return;
  const int64_t n = 2 * nact;
  if ((int64_t)b * kRpb >= n) {  // beyond this list
    if (threadIdx.x == 0) {
      Partial z;
      partial_zero(z);
      A.partials[b] = z;
      A.offsets[b] = 0;
    }
    return;
  }
  const int64_t r = (int64_t)b * kRpb + threadIdx.x / kHalfS;
  const bool valid = r < n;
  const int64_t rr = valid ? r : n - 1;
  // the parent: the (rr/2)-th active region of the previous list
  const int64_t k = rr >> 1;
  uint32_t lo_b = 0, hi_b = pnb;
  while (hi_b - lo_b > 1) {
    const uint32_t mid = (lo_b + hi_b) >> 1;
    if (s_base[mid] <= k) lo_b = mid;
    else hi_b = mid;
  }
  const uint32_t e = F.pact[lo_b * kRpb + (uint32_t)(k - s_base[lo_b])];
  const int64_t par = (int64_t)lo_b * kRpb + (e >> 4);
  const int a = (int)(e & 15u);
  double lo[D], wd[D];
#pragma unroll
  for (int m = 0; m < D; ++m) {
    lo[m] = F.plo[m * A.ld + par];
    wd[m] = F.pwd[m * A.ld + par];
  }
  const bool upper = (rr & 1) != 0;
#pragma unroll
  for (int m = 0; m < D; ++m)
    if (m == a) {
      const double hw = 0.5 * wd[m];
      if (upper) lo[m] = lo[m] + hw;
      wd[m] = hw;
    }
  if (valid) {
#pragma unroll
    for (int m = 0; m < D; ++m)
      if (s == m) {
        F.clo[m * A.ld + r] = lo[m];
        F.cwd[m * A.ld + r] = wd[m];
      }
  }
  const PlainIntegrand<D, FAM> f{P};
  const RegionOut o = seg_eval<D, true>(f, lo, wd, 1, r, valid, 0, kHalfS, s, 0, lane, 0, nullptr, R, A.dvec);
  half_outputs(A, o, r, valid, s == 0, b);
}

template <int D, int FAM>
void launch_fused(const EvalArgs& A, const FuseArgs& F, const DevIntegrand& P, const RuleConst& R, hipStream_t st)
{
  // the list's blocks and block 0 for the previous iteration's finalisation
  const int64_t nb = (A.n + kBlock / kHalfS - 1) / (kBlock / kHalfS) + 1;
  hipLaunchKernelGGL((k_eval_fused<D, FAM>), dim3((unsigned)nb), dim3(kBlock), 0, st, A, F, P, R);
}

template <int D, int FAM>
__global__ __launch_bounds__(512) void k_eval9_seg(EvalArgs A, DevIntegrand P, RuleConst R, int S)
{
  extern __shared__ double seg_lds[];
  const int W = S < 4 ? 4 : S;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int lane = threadIdx.x & (kWave - 1);
  const int g = wave / S, s = wave - g * S;
  const int64_t r = (int64_t)blockIdx.x * (kWave * (W / S)) + g * kWave + lane;
  const int64_t n = eval_len(A);
  if (block_beyond(A, n, (int64_t)blockIdx.x * (kWave * (W / S)))) return;
  const bool valid = r < n;
  const int64_t rr = valid ? r : (n > 0 ? n - 1 : 0);
  const PlainIntegrand<D, FAM> f{P};
  const RegionOut o = seg_eval9<D>(f, A.lower, A.width, A.ld, r, valid, rr, S, s, wave, lane, W, seg_lds, R.r9,
                                   A.dvec);
  const bool owner = s == 0;
  if (valid && owner) {
    A.est[r] = o.est;
    A.err[r] = o.err;
    A.axis[r] = (int8_t)o.axis;
    if (A.abssum) A.abssum[r] = o.ab;
    if (A.i5) A.i5[r] = o.i5;
  }
  if (A.finalize) {
    Partial p;
    block_reduce_regions(owner, g * kWave + lane, kWave * (W / S), o.est, o.err, o.ab, valid, A.loc_epsrel,
                         A.loc_epsabs, p);
    store_block_partial(A, p);
  }
}

template <int D, int FAM>
void launch_eval9(const EvalArgs& A, const DevIntegrand& P, const RuleConst& R, int S, hipStream_t st)
{
  if (S == kHalfS) return launch_half<D, FAM, 9>(A, P, R, st);
  if (S > 8) S = 8;  // LDS of the merge: 7 doubles + 1 int per lane and wave
  const int rpb = seg_regions_per_block(S);
  const int64_t nb = (A.n + rpb - 1) / rpb;
  hipExtLaunchKernelGGL((k_eval9_seg<D, FAM>), dim3((unsigned)nb), dim3(kWave * seg_waves(S)),
                        (uint32_t)seg9_lds_bytes(S), st, A.prof_start, A.prof_stop, 0, A, P, R, S);
}

// Launchers for one family over dimensions 2..10 (key 9: the degree-9 rule).
template <int FAM>
FusedLauncher family_fused_launcher(int dim)
{
  switch (dim) {
  case 2: return &launch_fused<2, FAM>;
  case 3: return &launch_fused<3, FAM>;
  case 4: return &launch_fused<4, FAM>;
  case 5: return &launch_fused<5, FAM>;
  case 6: return &launch_fused<6, FAM>;
  case 7: return &launch_fused<7, FAM>;
  case 8: return &launch_fused<8, FAM>;
  case 9: return &launch_fused<9, FAM>;
  case 10: return &launch_fused<10, FAM>;
  default: return nullptr;
  }
}

template <int FAM>
EvalLauncher family_launcher(int dim, int key = 7)
{
  if (key == 9) {
    switch (dim) {
    case 2: return &launch_eval9<2, FAM>;
    case 3: return &launch_eval9<3, FAM>;
    case 4: return &launch_eval9<4, FAM>;
    case 5: return &launch_eval9<5, FAM>;
    case 6: return &launch_eval9<6, FAM>;
    case 7: return &launch_eval9<7, FAM>;
    case 8: return &launch_eval9<8, FAM>;
    case 9: return &launch_eval9<9, FAM>;
    case 10: return &launch_eval9<10, FAM>;
    default: return nullptr;
    }
  }
  switch (dim) {
  case 2: return &launch_eval<2, FAM>;
  case 3: return &launch_eval<3, FAM>;
  case 4: return &launch_eval<4, FAM>;
  case 5: return &launch_eval<5, FAM>;
  case 6: return &launch_eval<6, FAM>;
  case 7: return &launch_eval<7, FAM>;
  case 8: return &launch_eval<8, FAM>;
  case 9: return &launch_eval<9, FAM>;
  case 10: return &launch_eval<10, FAM>;
  default: return nullptr;
  }
}

}  // namespace cuhre
