// phase2.hpp -- the reference's second phase: every region left by the
// breadth-first phase is refined on its own (quad/GPUquad/GPUQuadPhases.cu:
// 422-536, host side quad/GPUquad/GPUKernelquad.cu:670-911).
//
// The reference gives each region a 256-thread block with a region pool in
// shared memory, device-side malloc for the overflow and a serial max-error
// search.  Here each region is one lane: all 64 lanes of a wavefront walk the
// same point sequence for their own current pair of children (so the rule
// evaluation is the same wave-uniform code as the breadth-first kernels),
// and each lane keeps its regions and a binary max-error heap in a per-lane
// slice of global memory (slot-major, lane-minor, so that lanes touching the
// same slot index touch consecutive addresses).  A lane stops when its own
// sum meets max(epsrel |RESULT|, epsabs) or after pool_max bisections; the
// wavefront runs while any lane is live.  The loop has no barrier, so lanes
// beyond the list return at once.
#pragma once

#include "eval_kernels.hpp"

namespace cuhre {

struct Phase2Args {
  const double* lower;  // input list, SoA lower[m*ld + r]
  const double* width;
  int64_t ld;
  int64_t n;
  int64_t L;  // lanes of the workspace (stride of every per-lane array)
  double* slo;  // [slot][m][lane]
  double* swd;
  double* sest;  // [slot][lane]
  double* serr;
  int8_t* sax;
  int32_t* sseq;
  int32_t* heap;  // [position][lane] -> slot
  double epsrel, epsabs;
  int32_t pool_max;
  double* res;  // per region of the input list
  double* err;
  int32_t* cnt;
  int8_t* act;
  int8_t* nonfinite;
};

// heap order: larger error first, then the region created first
__device__ __forceinline__ bool p2_before(const Phase2Args& A, int64_t g, int32_t a, int32_t b)
{
  const double ea = A.serr[(int64_t)a * A.L + g], eb = A.serr[(int64_t)b * A.L + g];
  return ea > eb || (ea == eb && A.sseq[(int64_t)a * A.L + g] < A.sseq[(int64_t)b * A.L + g]);
}

__device__ __forceinline__ void p2_push(const Phase2Args& A, int64_t g, int32_t& hn, int32_t slot)
{
  int32_t i = hn++;
  A.heap[(int64_t)i * A.L + g] = slot;
  while (i > 0) {
    const int32_t pa = (i - 1) / 2;
    const int32_t si = A.heap[(int64_t)i * A.L + g], sp = A.heap[(int64_t)pa * A.L + g];
    if (!p2_before(A, g, si, sp)) break;
    A.heap[(int64_t)i * A.L + g] = sp;
    A.heap[(int64_t)pa * A.L + g] = si;
    i = pa;
  }
}

__device__ __forceinline__ int32_t p2_pop(const Phase2Args& A, int64_t g, int32_t& hn)
{
  const int32_t top = A.heap[g];
  --hn;
  A.heap[g] = A.heap[(int64_t)hn * A.L + g];
  int32_t i = 0;
  for (;;) {
    const int32_t l = 2 * i + 1, r = l + 1;
    int32_t b = i;
    if (l < hn && p2_before(A, g, A.heap[(int64_t)l * A.L + g], A.heap[(int64_t)b * A.L + g])) b = l;
    if (r < hn && p2_before(A, g, A.heap[(int64_t)r * A.L + g], A.heap[(int64_t)b * A.L + g])) b = r;
    if (b == i) break;
    const int32_t t = A.heap[(int64_t)i * A.L + g];
    A.heap[(int64_t)i * A.L + g] = A.heap[(int64_t)b * A.L + g];
    A.heap[(int64_t)b * A.L + g] = t;
    i = b;
  }
  return top;
}

// Evaluates the region in workspace slot `slot` of lane g with the rule
// selected by KEY (one point segment: no block barrier).
template <int D, int FAM, int KEY>
__device__ __forceinline__ RegionOut p2_eval(const Phase2Args& A, const DevIntegrand& P, const RuleConst& R,
                                             int64_t g, int32_t slot)
{
  const PlainIntegrand<D, FAM> f{P};
  const double* lo = A.slo + (int64_t)slot * D * A.L;
  const double* wd = A.swd + (int64_t)slot * D * A.L;
  const int lane = threadIdx.x & (kWave - 1);
  if constexpr (KEY == 9)
    return seg_eval9<D>(f, lo, wd, A.L, g, true, g, 1, 0, 0, lane, 4, nullptr, R.r9, nullptr);
  else
    return seg_eval<D>(f, lo, wd, A.L, g, true, g, 1, 0, 0, lane, 4, nullptr, R, nullptr);
}

template <int D, int FAM, int KEY>
__global__ __launch_bounds__(256) void k_phase2(Phase2Args A, DevIntegrand P, RuleConst R)
{
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= A.n) return;  // no barrier below
  for (int m = 0; m < D; ++m) {
    A.slo[(int64_t)m * A.L + g] = A.lower[m * A.ld + g];
    A.swd[(int64_t)m * A.L + g] = A.width[m * A.ld + g];
  }
  RegionOut o = p2_eval<D, FAM, KEY>(A, P, R, g, 0);
  bool bad = !(isfinite(o.est) && isfinite(o.err));
  A.sest[g] = o.est;
  A.serr[g] = o.err;
  A.sax[g] = (int8_t)o.axis;
  A.sseq[g] = 0;
  int32_t hn = 0, next = 1, count = 1;
  p2_push(A, g, hn, 0);
  double RES = o.est, ERR = o.err;
  while (count <= A.pool_max && (count == 1 || ERR > fmax(A.epsrel * fabs(RES), A.epsabs))) {
    const int32_t a = p2_pop(A, g, hn), b = count;
    const double pe = A.sest[(int64_t)a * A.L + g], pr = A.serr[(int64_t)a * A.L + g];
    const int k = A.sax[(int64_t)a * A.L + g];
    for (int m = 0; m < D; ++m) {
      A.slo[((int64_t)b * D + m) * A.L + g] = A.slo[((int64_t)a * D + m) * A.L + g];
      A.swd[((int64_t)b * D + m) * A.L + g] = A.swd[((int64_t)a * D + m) * A.L + g];
    }
    const double hw = 0.5 * A.swd[((int64_t)a * D + k) * A.L + g];
    A.swd[((int64_t)a * D + k) * A.L + g] = hw;
    A.swd[((int64_t)b * D + k) * A.L + g] = hw;
    A.slo[((int64_t)b * D + k) * A.L + g] = A.slo[((int64_t)a * D + k) * A.L + g] + hw;
    double ce[2], cr[2];
    int cax[2];
    for (int c = 0; c < 2; ++c) {
      const RegionOut q = p2_eval<D, FAM, KEY>(A, P, R, g, c ? b : a);
      ce[c] = q.est;
      cr[c] = q.err;
      cax[c] = q.axis;
      bad = bad || !(isfinite(q.est) && isfinite(q.err));
    }
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
    A.sest[(int64_t)a * A.L + g] = ce[0];
    A.serr[(int64_t)a * A.L + g] = cr[0];
    A.sax[(int64_t)a * A.L + g] = (int8_t)cax[0];
    A.sseq[(int64_t)a * A.L + g] = next++;
    A.sest[(int64_t)b * A.L + g] = ce[1];
    A.serr[(int64_t)b * A.L + g] = cr[1];
    A.sax[(int64_t)b * A.L + g] = (int8_t)cax[1];
    A.sseq[(int64_t)b * A.L + g] = next++;
    p2_push(A, g, hn, a);
    p2_push(A, g, hn, b);
    ++count;
  }
  int8_t active = ERR > fmax(A.epsrel * fabs(RES), A.epsabs) ? 1 : 0;
  if (ERR > 1e10) {  // the reference zeroes a diverging region and flags it
    RES = 0.0;
    ERR = 0.0;
    active = 1;
  }
  A.res[g] = RES;
  A.err[g] = ERR;
  A.cnt[g] = count;
  A.act[g] = active;
  A.nonfinite[g] = bad ? 1 : 0;
}

using Phase2Launcher = void (*)(const Phase2Args&, const DevIntegrand&, const RuleConst&, hipStream_t);

template <int D, int FAM, int KEY>
void launch_phase2(const Phase2Args& A, const DevIntegrand& P, const RuleConst& R, hipStream_t st)
{
  const int64_t nb = (A.n + 255) / 256;
  hipLaunchKernelGGL((k_phase2<D, FAM, KEY>), dim3((unsigned)nb), dim3(256), 0, st, A, P, R);
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
