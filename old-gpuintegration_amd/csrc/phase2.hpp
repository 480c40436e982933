// phase2.hpp -- the reference's second phase: every region left by the
// breadth-first phase is refined on its own (quad/GPUquad/GPUQuadPhases.cu:
// 422-536, host side quad/GPUquad/GPUKernelquad.cu:670-911).
//
// The reference gives each region a 256-thread block with a region pool in
// shared memory, device-side malloc for the overflow and a serial max-error
// search.  Here each region is one wavefront (a 64-thread block).  A
// bisection evaluates both children at once: lanes 0-31 the left child and
// lanes 32-63 the right one, each half taking the rule's point sets round
// robin (32 segments) and merging by a xor butterfly.  The pool's keys (error
// and creation number per slot) of the first kP2LdsSlots slots sit in LDS,
// the rest (rarely used: most regions stop within a few dozen bisections) in
// the region's slice of global memory, so the next region to bisect -- the
// largest error, the earliest created among equal errors, as the
// reference's max-heap order -- is one strided scan plus a butterfly.  The
// small LDS footprint (3 KB per wavefront) keeps ~20 wavefronts per CU in
// flight.  The pool's geometry, estimates and axes sit in the region's own
// slice of global memory ([slot][lower D, width D]), written by lanes 0 and
// 32 and read back by the whole wave after a barrier.  A region stops when its own
// sum meets max(epsrel |RESULT|, epsabs) or after pool_max bisections.
// Per-region run lengths differ by orders of magnitude: the host evaluates
// every region once beforehand with the half-wavefront kernel (the same
// evaluator) and launches the regions by decreasing ratio of that first
// error to the region's tolerance (Phase2Args::order), so the longest runs
// start first and the hardware dispatcher fills the chip with the short
// ones around them.
#pragma once

#include "eval_kernels.hpp"

namespace cuhre {

struct Phase2Args {
  const double* lower;  // input list, SoA lower[m*ld + r]
  const double* width;
  int64_t ld;
  int64_t n;
  int64_t slots;  // pool slots per region (pool_max + 2)
  double* geo;    // [region][slot][lower D, width D]
  double* gkerr;  // [region][slot] pool keys of the slots past the LDS ones
  int32_t* gkseq;
  double* sest;   // [region][slot]
  int8_t* sax;    // [region][slot]
  double epsrel, epsabs;
  int32_t pool_max;
  double* res;  // per region of the input list
  double* err;
  int32_t* cnt;
  int8_t* act;
  int8_t* nonfinite;
  int8_t* zeroed;  // the region's error passed 1e10 and it was zeroed
  // rounding budget per region: sum of the abssums of every region the
  // pool evaluated, and sum over the RES updates of
  // 2 (|child 0| + |child 1| + |parent|) + |RES| (each update rounds three times)
  double* rab;
  double* rres;
  // optional: block b refines region order[b] (the regions expected to run
  // longest first, so that their chains do not start last)
  const int32_t* order;
  // optional: each region's first evaluation, made beforehand by the
  // half-wavefront kernel (the same evaluator, so the same values)
  const double* est0;
  const double* err0;
  const double* ab0;
  const int8_t* ax0;
};

// Pool slots whose keys (error, creation number) live in LDS.
constexpr int kP2LdsSlots = 256;
// Rows of 64 pool slots whose keys one step of the max-error scan loads.
constexpr int kScanBatch = 8;

__device__ __forceinline__ double read_lane_d(double v, int l)
{
  const long long b = __double_as_longlong(v);
  const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)b, l);
  const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(b >> 32), l);
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// Evaluates the region (lo, wd) with the rule selected by KEY, 32 segments
// per half wavefront; every lane of a half returns that half's result.
template <int D, int FAM, int KEY>
__device__ __forceinline__ RegionOut p2_eval(const PlainIntegrand<D, FAM>& f, const RuleConst& R,
                                             const double (&lo)[D], const double (&wd)[D], int s)
{
  if constexpr (KEY == 9)
    return seg_eval9<D, true>(f, lo, wd, 1, 0, true, 0, 32, s, 0, s, 0, nullptr, R.r9, nullptr);
  else
    // (without the moved-coordinate vectors: their registers cost this
    // kernel a wavefront per SIMD)
    return seg_eval<D, true, false, false>(f, lo, wd, 1, 0, true, 0, 32, s, 0, s, 0, nullptr, R, nullptr);
}

template <int D, int FAM, int KEY>
__global__ __launch_bounds__(64) void k_phase2(Phase2Args A, DevIntegrand P, RuleConst R)
{
  __shared__ double s_err[kP2LdsSlots];
  __shared__ int32_t s_seq[kP2LdsSlots];
  // the region of the input list this block refines, and its workspace slice
  const int64_t g = A.order != nullptr ? (int64_t)A.order[blockIdx.x] : (int64_t)blockIdx.x;
  double* const gerr = A.gkerr + g * A.slots;
  int32_t* const gseq = A.gkseq + g * A.slots;
  auto key_err = [&](int32_t t) { return t < kP2LdsSlots ? s_err[t] : gerr[t]; };
  auto key_seq = [&](int32_t t) { return t < kP2LdsSlots ? s_seq[t] : gseq[t]; };
  auto set_key = [&](int32_t t, double e, int32_t q) {
    if (t < kP2LdsSlots) {
      s_err[t] = e;
      s_seq[t] = q;
    } else {
      gerr[t] = e;
      gseq[t] = q;
    }
  };
  const int lane = threadIdx.x;
  const int half = lane >> 5;
  double* const geo = A.geo + g * A.slots * (2 * D);
  double* const sest = A.sest + g * A.slots;
  int8_t* const sax = A.sax + g * A.slots;
  const PlainIntegrand<D, FAM> f{P};

  double lo[D], wd[D];
#pragma unroll
  for (int m = 0; m < D; ++m) {
    lo[m] = A.lower[m * A.ld + g];
    wd[m] = A.width[m * A.ld + g];
  }
  // the input region's first evaluation (given, or by both halves here);
  // lane 0 keeps it in slot 0
  double e0, r0, rab;
  int ax0;
  if (A.est0 != nullptr) {
    e0 = A.est0[g];
    r0 = A.err0[g];
    rab = A.ab0[g];
    ax0 = A.ax0[g];
  } else {
    const RegionOut o = p2_eval<D, FAM, KEY>(f, R, lo, wd, lane & 31);
    e0 = read_lane_d(o.est, 0);
    r0 = read_lane_d(o.err, 0);
    rab = read_lane_d(o.ab, 0);
    ax0 = o.axis;
  }
  bool bad = !(isfinite(e0) && isfinite(r0));
  double rres = 0.0;
  if (lane == 0) {
#pragma unroll
    for (int m = 0; m < D; ++m) {
      geo[m] = lo[m];
      geo[D + m] = wd[m];
    }
    sest[0] = e0;
    sax[0] = (int8_t)ax0;
    set_key(0, r0, 0);
  }
  int32_t count = 1, next = 1;
  double RES = e0, ERR = r0;
  while (count <= A.pool_max && (count == 1 || ERR > fmax(A.epsrel * fabs(RES), A.epsabs))) {
    __syncthreads();  // the records written by lanes 0 and 32 are visible to the wave
    // the slot to bisect: largest error, earliest created among equal errors
    // (a pool of more than 64 slots is read kScanBatch rows of 64 slots at
    // a time, so that a deep pool's keys in global memory arrive together
    // instead of one row per round trip; the order among candidates is a
    // strict total order, so the result does not depend on the order of the
    // comparisons)
    double be = -INFINITY;
    int32_t bq = INT32_MAX, bs = -1;
    if (count <= kWave) {
      if (lane < count) {
        const double e = key_err(lane);
        if (e > be) {  // false for a NaN key, as in the batched form
          be = e;
          bq = key_seq(lane);
          bs = lane;
        }
      }
    } else {
      for (int32_t t0 = lane; t0 < count; t0 += kScanBatch * kWave) {
        double e[kScanBatch];
        int32_t q[kScanBatch];
#pragma unroll
        for (int u = 0; u < kScanBatch; ++u) {
          const int32_t t = t0 + u * kWave;
          e[u] = t < count ? key_err(t) : -INFINITY;
          q[u] = t < count ? key_seq(t) : INT32_MAX;
        }
#pragma unroll
        for (int u = 0; u < kScanBatch; ++u)
          if (e[u] > be || (e[u] == be && q[u] < bq)) {
            be = e[u];
            bq = q[u];
            bs = t0 + u * kWave;
          }
      }
    }
#pragma unroll
    for (int msk = 32; msk > 0; msk >>= 1) {
      const double e2 = shfl_xor_d(be, msk);
      const int32_t q2 = __shfl_xor(bq, msk, kWave), s2 = __shfl_xor(bs, msk, kWave);
      if (e2 > be || (e2 == be && q2 < bq)) {
        be = e2;
        bq = q2;
        bs = s2;
      }
    }
    int32_t a = __builtin_amdgcn_readfirstlane(bs);
    if (a < 0) a = 0;  // every error non-finite: the run is reported as such
    const int32_t b = count;
    const double pe = sest[a], pr = key_err(a);
    const int k = __builtin_amdgcn_readfirstlane((int)sax[a]);
    double plo[D], pwd[D], wk = 0.0;
#pragma unroll
    for (int m = 0; m < D; ++m) {
      plo[m] = geo[(int64_t)a * (2 * D) + m];
      pwd[m] = geo[(int64_t)a * (2 * D) + D + m];
      if (m == k) wk = pwd[m];
    }
    const double hw = 0.5 * wk;
    // lanes 0-31: the lower half along axis k (slot a), lanes 32-63 the upper (slot b)
#pragma unroll
    for (int m = 0; m < D; ++m) {
      wd[m] = m == k ? hw : pwd[m];
      lo[m] = (m == k && half) ? plo[m] + hw : plo[m];
    }
    // the segment index is loop-invariant; kept opaque so the compiler does
    // not hoist the per-lane point decode out of the loop (its results held
    // in registers across iterations took the kernel from 249 to 308 VGPRs,
    // one wavefront per SIMD instead of two)
    int seg = lane & 31;
    asm volatile("" : "+v"(seg));
    const RegionOut q = p2_eval<D, FAM, KEY>(f, R, lo, wd, seg);
    double ce[2] = {read_lane_d(q.est, 0), read_lane_d(q.est, 32)};
    double cr[2] = {read_lane_d(q.err, 0), read_lane_d(q.err, 32)};
    bad = bad || !(isfinite(ce[0]) && isfinite(cr[0]) && isfinite(ce[1]) && isfinite(cr[1]));
    // Cuba's refinement from the parent difference (GPUQuadPhases.cu:498-515)
    const double diff = fabs(0.25 * (ce[0] + ce[1] - pe));
    const double e = cr[0] + cr[1];
    if (e > 0) {
      const double c = 1 + 2 * diff / e;
      cr[0] *= c;
      cr[1] *= c;
    }
    cr[0] += diff;
    cr[1] += diff;
    ERR += cr[0] + cr[1] - pr;
    RES += ce[0] + ce[1] - pe;
    rab += read_lane_d(q.ab, 0) + read_lane_d(q.ab, 32);
    rres += 2.0 * (fabs(ce[0]) + fabs(ce[1]) + fabs(pe)) + fabs(RES);
    __syncthreads();  // every lane has read slot a before it is overwritten
    if ((lane & 31) == 0) {
      const int32_t slot = half ? b : a;
#pragma unroll
      for (int m = 0; m < D; ++m) {
        geo[(int64_t)slot * (2 * D) + m] = lo[m];
        geo[(int64_t)slot * (2 * D) + D + m] = wd[m];
      }
      sest[slot] = half ? ce[1] : ce[0];
      sax[slot] = (int8_t)q.axis;
      set_key(slot, half ? cr[1] : cr[0], next + half);
    }
    next += 2;
    ++count;
  }
  if (lane == 0) {
    int8_t active = ERR > fmax(A.epsrel * fabs(RES), A.epsabs) ? 1 : 0;
    int8_t zero = 0;
    if (ERR > 1e10) {  // the reference zeroes a diverging region and flags it
      RES = 0.0;
      ERR = 0.0;
      active = 1;
      zero = 1;
    }
    A.res[g] = RES;
    A.err[g] = ERR;
    A.cnt[g] = count;
    A.act[g] = active;
    A.nonfinite[g] = bad ? 1 : 0;
    A.zeroed[g] = zero;
    A.rab[g] = rab;
    A.rres[g] = rres;
  }
}

using Phase2Launcher = void (*)(const Phase2Args&, const DevIntegrand&, const RuleConst&, hipStream_t);

template <int D, int FAM, int KEY>
void launch_phase2(const Phase2Args& A, const DevIntegrand& P, const RuleConst& R, hipStream_t st)
{
  hipLaunchKernelGGL((k_phase2<D, FAM, KEY>), dim3((unsigned)A.n), dim3(kWave), 0, st, A,
                     P, R);
}

template <int FAM>
Phase2Launcher phase2_launcher(int dim, int key)
{
#define CUHRE_P2_CASES(K)                      \
  switch (dim) {                               \
  case 2: return &launch_phase2<2, FAM, K>;    \
  case 3: return &launch_phase2<3, FAM, K>;    \
  case 4: return &launch_phase2<4, FAM, K>;    \
  case 5: return &launch_phase2<5, FAM, K>;    \
  case 6: return &launch_phase2<6, FAM, K>;    \
  case 7: return &launch_phase2<7, FAM, K>;    \
  case 8: return &launch_phase2<8, FAM, K>;    \
  case 9: return &launch_phase2<9, FAM, K>;    \
  case 10: return &launch_phase2<10, FAM, K>;  \
  default: return nullptr;                     \
  }
  if (key == 9) {
    CUHRE_P2_CASES(9)
  }
  CUHRE_P2_CASES(7)
#undef CUHRE_P2_CASES
}

// Returns the Phase-2 launcher for (family, dim, key) or nullptr.
Phase2Launcher find_phase2_launcher(int family, int dim, int key);

}  // namespace cuhre
