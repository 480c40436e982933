// engine.hip -- host loop of the plain adaptive cubature method, the
// classify/compact/bisect kernel and the C-ABI (include/cuhre.h).
//
// Reference path replaced: GPUcuhre<T>::integrate (quad/GPUquad/GPUquad.h:
// 611-727) -> GPUKernelCuhre::IntegrateFirstPhase / FirstPhaseIteration /
// GenerateActiveIntervals (quad/GPUquad/GPUKernelquad.cu:356-668).  The
// reference allocates and frees ~7 device buffers per iteration, reduces
// with four thrust calls and synchronises the host three times; here the
// region list lives in two preallocated SoA buffers in HBM (ping-pong), an
// iteration is two launches (evaluate+classify+reduce, then
// compact+bisect) and one 160-byte state copy, and the loop runs until the
// global test, the evaluation budget or the capacity stops it (no 2^14
// region Phase-1 cap and no Phase 2).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "common.hpp"
#include "context.hpp"
#include "eval_kernels.hpp"
#include "phase2.hpp"
#include "reduce.hpp"

namespace cuhre {

EvalLauncher launcher_gaussian(int, int);
EvalLauncher launcher_product_peak(int, int);
EvalLauncher launcher_oscillatory(int, int);
EvalLauncher launcher_corner_peak(int, int);
EvalLauncher launcher_continuous(int, int);
EvalLauncher launcher_polynomial(int, int);
EvalLauncher launcher_cosine_sum(int, int);
EvalLauncher launcher_table6(int, int);
EvalLauncher launcher_lclt(int, int);

FusedLauncher fused_gaussian(int);
FusedLauncher fused_product_peak(int);
FusedLauncher fused_oscillatory(int);
FusedLauncher fused_corner_peak(int);
FusedLauncher fused_continuous(int);
FusedLauncher fused_polynomial(int);
FusedLauncher fused_cosine_sum(int);
FusedLauncher fused_table6(int);

Phase2Launcher p2launcher_gaussian(int, int);
Phase2Launcher p2launcher_product_peak(int, int);
Phase2Launcher p2launcher_oscillatory(int, int);
Phase2Launcher p2launcher_corner_peak(int, int);
Phase2Launcher p2launcher_continuous(int, int);
Phase2Launcher p2launcher_polynomial(int, int);
Phase2Launcher p2launcher_cosine_sum(int, int);
Phase2Launcher p2launcher_table6(int, int);
Phase2Launcher p2launcher_lclt(int, int);

Phase2Launcher find_phase2_launcher(int family, int dim, int key)
{
  if (key == 0) key = 7;
  if (key != 7 && key != 9) return nullptr;
  switch (family) {
  case CUHRE_GAUSSIAN: return p2launcher_gaussian(dim, key);
  case CUHRE_PRODUCT_PEAK: return p2launcher_product_peak(dim, key);
  case CUHRE_OSCILLATORY: return p2launcher_oscillatory(dim, key);
  case CUHRE_CORNER_PEAK: return p2launcher_corner_peak(dim, key);
  case CUHRE_CONTINUOUS: return p2launcher_continuous(dim, key);
  case CUHRE_POLYNOMIAL: return p2launcher_polynomial(dim, key);
  case CUHRE_COSINE_SUM: return p2launcher_cosine_sum(dim, key);
  case CUHRE_TABLE6: return p2launcher_table6(dim, key);
  case CUHRE_LC_LT_T: return p2launcher_lclt(dim, key);
  default: return nullptr;
  }
}

EvalLauncher find_eval_launcher(int family, int dim, int key)
{
  if (key == 0) key = 7;
  if (key != 7 && key != 9) return nullptr;
  switch (family) {
  case CUHRE_GAUSSIAN: return launcher_gaussian(dim, key);
  case CUHRE_PRODUCT_PEAK: return launcher_product_peak(dim, key);
  case CUHRE_OSCILLATORY: return launcher_oscillatory(dim, key);
  case CUHRE_CORNER_PEAK: return launcher_corner_peak(dim, key);
  case CUHRE_CONTINUOUS: return launcher_continuous(dim, key);
  case CUHRE_POLYNOMIAL: return launcher_polynomial(dim, key);
  case CUHRE_COSINE_SUM: return launcher_cosine_sum(dim, key);
  case CUHRE_TABLE6: return launcher_table6(dim, key);
  case CUHRE_LC_LT_T: return launcher_lclt(dim, key);
  default: return nullptr;
  }
}

FusedLauncher find_fused_launcher(int family, int dim)
{
  switch (family) {
  case CUHRE_GAUSSIAN: return fused_gaussian(dim);
  case CUHRE_PRODUCT_PEAK: return fused_product_peak(dim);
  case CUHRE_OSCILLATORY: return fused_oscillatory(dim);
  case CUHRE_CORNER_PEAK: return fused_corner_peak(dim);
  case CUHRE_CONTINUOUS: return fused_continuous(dim);
  case CUHRE_POLYNOMIAL: return fused_polynomial(dim);
  case CUHRE_COSINE_SUM: return fused_cosine_sum(dim);
  case CUHRE_TABLE6: return fused_table6(dim);
  default: return nullptr;  // LC_LT_t: its iterations keep the split launch
  }
}

#define HIP_TRY(expr)                                                                   \
  do {                                                                                  \
    hipError_t e_ = (expr);                                                             \
    if (e_ != hipSuccess)                                                               \
      return fail(CUHRE_ERR_DEVICE, "%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_), \
                  __FILE__, __LINE__);                                                  \
  } while (0)

struct BatchEngine;
BatchEngine* batch_engine_new();
void batch_engine_delete(BatchEngine* e);

// ------------------------------------------------------------------ split



// Classify + stream-compact + bisect in one pass (replaces the reference's
// thrust::exclusive_scan, alignRegions and divideIntervalsGPU,
// quad/GPUquad/GPUKernelquad.cu:63-149, 356-505).  The rank of an active
// region inside its evaluation block comes from a wavefront ballot (plus
// the counts of earlier waves of the same block through LDS); the block's
// base offset was scanned by k_finish.  Child 0
// of the j-th active region goes to slot 2j and child 1 to 2j+1, so the
// two children of a region are adjacent and written with one 16-byte store
// per coordinate.
// Lists up to this many regions load a region's split inputs before the
// split's count scan (see k_split_fin).
constexpr int64_t kPrefetchSplit = 262144;

// The two children of a region with bounds (lo, wd), bisected along axis
// a, to slots 2 dst and 2 dst + 1 of the output list: one 16-byte store per
// coordinate and list.  The bounds were loaded for every coordinate first,
// so their loads are in flight together.
__device__ __forceinline__ void write_children(const SplitArgs& S, int64_t dst, int a, const double (&lo)[kMaxDim],
                                               const double (&wd)[kMaxDim])
{
#pragma unroll
  for (int m = 0; m < kMaxDim; ++m) {
    if (m >= S.dim) break;
    double2 L, W;
    if (m == a) {
      const double hw = 0.5 * wd[m];
      L = make_double2(lo[m], lo[m] + hw);
      W = make_double2(hw, hw);
    } else {
      L = make_double2(lo[m], lo[m]);
      W = make_double2(wd[m], wd[m]);
    }
    *reinterpret_cast<double2*>(S.out_lower + m * S.out_ld + 2 * dst) = L;
    *reinterpret_cast<double2*>(S.out_width + m * S.out_ld + 2 * dst) = W;
  }
}

__global__ __launch_bounds__(kBlock) void k_split(SplitArgs S)
{
  if (S.state != nullptr && S.state->stop != 0) return;
  __shared__ uint32_t wcnt[kBlock / kWave];
  const int64_t r = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  const bool valid = r < S.n;
  bool act = false;
  if (valid) {
    const double e = S.est[r], er = S.err[r];
    act = isfinite(e) && isfinite(er) && region_active(e, er, S.epsrel, S.epsabs);
    if (S.retire != nullptr && (S.retire_on == nullptr || *S.retire_on) && S.retire[r]) act = false;
  }
  const uint64_t mask = __ballot(act);
  const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
  const uint64_t below = lane ? (~0ull >> (kWave - lane)) : 0ull;
  int64_t rank;
  if (S.rpb >= kWave) {
    if (lane == 0) wcnt[wave] = (uint32_t)__popcll(mask);
    __syncthreads();
    const int wpb = S.rpb / kWave;
    const int first = wave - wave % wpb;
    rank = __popcll(mask & below);
    for (int w = first; w < wave; ++w) rank += wcnt[w];
  } else {
    const int g0 = lane - lane % S.rpb;
    const uint64_t grp = (S.rpb == kWave ? ~0ull : ((1ull << S.rpb) - 1ull)) << g0;
    rank = __popcll(mask & grp & below);
  }
  if (!act) return;
  const int64_t blk = r / S.rpb;
  const int64_t dst = S.offsets[blk] + (S.gbase != nullptr ? S.gbase[blk / kWave] : 0) + rank;
  const int a = S.axis[r];
  double lo[kMaxDim], wd[kMaxDim];
#pragma unroll
  for (int m = 0; m < kMaxDim; ++m)
    if (m < S.dim) {
      lo[m] = S.lower[m * S.ld + r];
      wd[m] = S.width[m * S.ld + r];
    }
  write_children(S, dst, a, lo, wd);
}

// The split with the iteration's finalisation folded in (one launch fewer
// per iteration).  part[0..np) are the evaluation's block partials, np <=
// kOneLevel (one level), or, after k_finish_groups, the sums of its groups
// of 64 blocks, np <= kWaveFinalize (two levels).  Every block but the last
// scans the np active counts (one level: the dense counts the evaluation
// blocks stored in offsets[]) into LDS for its own offsets and splits its
// regions; the last block (one beyond the list's blocks) reduces the
// partials in full, writes the group bases (two levels) for a re-run of the
// split, and writes and publishes the iteration state, in parallel with the
// split.
// Blocks skip the split when a count alone stops the loop (non-finite
// values, the budget, the buffers); when the global test is met the
// children written are unused.
// Queued ahead of the host (P.in set): the list length and the finished
// sums come from the device slot P.in, the number of partials and the
// levels of the finalisation follow from that length (the launch grids
// were sized for an upper bound), the budget test uses it, and the last
// block also writes the next list's slot: its length (0 when the loop stops, so
// that an iteration queued behind this one does nothing) and the finished
// sums after this iteration.
__global__ __launch_bounds__(kBlock) void k_split_fin(SplitArgs S, EvalArgs A, const Partial* part, uint32_t np,
                                                      int two_level, SpecArgs P)
{
  __shared__ int64_t s_base[kOneLevel];
  __shared__ int64_t s_wsum[kBlock / kWave], s_wbad[kBlock / kWave];
  __shared__ uint32_t wcnt[kBlock / kWave];
  const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
  int64_t n = S.n;
  if (P.in != nullptr) {
    n = P.in->n;
    if (n == 0) return;  // the loop stopped before this iteration
    const int64_t nb = (n + P.rpb - 1) / P.rpb;
    two_level = nb > (int64_t)kOneLevel;
    part = two_level ? A.gpart : A.partials;
    np = (uint32_t)(two_level ? (nb + kWave - 1) / kWave : nb);
    A.fin_est_s = P.in->fin_s;
    A.fin_est_c = P.in->fin_c;
    A.fin_err_s = P.in->fer_s;
    A.fin_err_c = P.in->fer_c;
    const int64_t left = P.left0 - n * P.npts;
    const int64_t budget = left >= 0 ? left / (2 * P.npts) : -1;
    A.max_active = budget < P.cap2 ? budget : P.cap2;
  }
  if (blockIdx.x == gridDim.x - 1) {
    Partial q;
    reduce_partials(part, np, two_level ? A.gbase : nullptr, q);
    if (threadIdx.x == 0) {
      const IterState st = iteration_state(A, q);
      write_state(A, st);
      if (P.out != nullptr) {
        SpecSlot o;
        o.n = st.stop == 0 ? 2 * q.nactive : 0;
        o.pad = 0;
        o.fin_s = A.fin_est_s;
        o.fin_c = A.fin_est_c;
        o.fer_s = A.fin_err_s;
        o.fer_c = A.fin_err_c;
        dd_add(o.fin_s, o.fin_c, q.fest_s, q.fest_c);
        dd_add(o.fer_s, o.fer_c, q.ferr_s, q.ferr_c);
        *P.out = o;
      }
    }
    return;
  }
  // Lists up to kPrefetchSplit regions are latency-bound: the region's est,
  // err, axis and bounds are loaded before the count scan so that their
  // latency overlaps it (larger lists load them after the stop test, and
  // the bounds only for active regions).
  const int64_t r = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  const bool valid = r < n, early = n <= kPrefetchSplit;
  double e = 0.0, er = 0.0, lo[kMaxDim], wd[kMaxDim];
  int a = 0;
  if (early && valid) {
    e = S.est[r];
    er = S.err[r];
    a = S.axis[r];
#pragma unroll
    for (int m = 0; m < kMaxDim; ++m)
      if (m < S.dim) {
        lo[m] = S.lower[m * S.ld + r];
        wd[m] = S.width[m * S.ld + r];
      }
  }
  // exclusive scan of the active counts, four per thread
  int64_t cnt[4], mine = 0, bad = 0;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const uint32_t j = 4 * threadIdx.x + i;
    cnt[i] = 0;
    if (j < np) {
      if (two_level) {
        cnt[i] = part[j].nactive;
        bad += part[j].nnonfinite;
      } else {
        const int64_t v = A.offsets[j];
        cnt[i] = v & kCountMask;
        bad += v >> 40;
      }
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
    if (4 * threadIdx.x + i < np) s_base[4 * threadIdx.x + i] = run;
    run += cnt[i];
  }
  __syncthreads();
  if (nonfinite != 0 || nact > A.max_active || nact > A.alloc_active) return;

  if (!early && valid) {
    e = S.est[r];
    er = S.err[r];
  }
  const bool act = valid && isfinite(e) && isfinite(er) && region_active(e, er, S.epsrel, S.epsabs);
  const uint64_t mask = __ballot(act);
  const uint64_t below = lane ? (~0ull >> (kWave - lane)) : 0ull;
  int64_t rank;
  if (S.rpb >= kWave) {
    if (lane == 0) wcnt[wave] = (uint32_t)__popcll(mask);
    __syncthreads();
    const int wpb = S.rpb / kWave;
    const int first = wave - wave % wpb;
    rank = __popcll(mask & below);
    for (int w = first; w < wave; ++w) rank += wcnt[w];
  } else {
    const int g0 = lane - lane % S.rpb;
    const uint64_t grp = (S.rpb == kWave ? ~0ull : ((1ull << S.rpb) - 1ull)) << g0;
    rank = __popcll(mask & grp & below);
  }
  if (!act) return;
  const int64_t blk = r / S.rpb;
  const int64_t dst = (two_level ? s_base[blk / kWave] + S.offsets[blk] : s_base[blk]) + rank;
  if (!early) {
    a = S.axis[r];
#pragma unroll
    for (int m = 0; m < kMaxDim; ++m)
      if (m < S.dim) {
        lo[m] = S.lower[m * S.ld + r];
        wd[m] = S.width[m * S.ld + r];
      }
  }
  write_children(S, dst, a, lo, wd);
}

// In place: the dense active counts of nb <= kOneLevel evaluation blocks
// (offsets[b] = count + kBadFlag when non-finite) to their exclusive scan,
// the offsets a re-run of the split takes.  One block, four per thread.
__global__ __launch_bounds__(kBlock) void k_scan_counts(int64_t* off, uint32_t nb)
{
  __shared__ int64_t s_w[kBlock / kWave];
  const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
  int64_t cnt[4], mine = 0;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const uint32_t j = 4 * threadIdx.x + i;
    cnt[i] = j < nb ? off[j] & kCountMask : 0;
    mine += cnt[i];
  }
  long long incl = (long long)mine;
#pragma unroll
  for (int o = 1; o < kWave; o <<= 1) {
    const long long v = __shfl_up(incl, o, kWave);
    if (lane >= o) incl += v;
  }
  if (lane == kWave - 1) s_w[wave] = (int64_t)incl;
  __syncthreads();
  int64_t run = (int64_t)incl - mine;
  for (int w = 0; w < wave; ++w) run += s_w[w];
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const uint32_t j = 4 * threadIdx.x + i;
    if (j < nb) off[j] = run;
    run += cnt[i];
  }
}

void launch_split(const SplitArgs& S, hipStream_t st)
{
  hipLaunchKernelGGL(k_split, dim3((unsigned)((S.n + kBlock - 1) / kBlock)), dim3(kBlock), 0, st, S);
}

// Active-region count per 256-region block (standalone split entry point).
__global__ __launch_bounds__(kBlock) void k_count(const double* est, const double* err, int64_t n,
                                                   double epsrel, double epsabs, int64_t* counts)
{
  __shared__ uint32_t wcnt[kBlock / kWave];
  const int64_t r = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  bool act = false;
  if (r < n) act = isfinite(est[r]) && isfinite(err[r]) && region_active(est[r], err[r], epsrel, epsabs);
  const uint64_t mask = __ballot(act);
  if ((threadIdx.x & (kWave - 1)) == 0) wcnt[threadIdx.x / kWave] = (uint32_t)__popcll(mask);
  __syncthreads();
  if (threadIdx.x == 0) {
    int64_t c = 0;
    for (int w = 0; w < kBlock / kWave; ++w) c += wcnt[w];
    counts[blockIdx.x] = c;
  }
}

// Deterministic two-pass FP64 sum with TwoSum-carried errors.
__global__ __launch_bounds__(kBlock) void k_sum_blocks(const double* x, int64_t n, int64_t per_block,
                                                        double2* part)
{
  __shared__ double2 lds[kBlock / kWave];
  const int64_t b0 = (int64_t)blockIdx.x * per_block;
  const int64_t b1 = b0 + per_block < n ? b0 + per_block : n;
  double s = 0.0, c = 0.0;
  for (int64_t i = b0 + threadIdx.x; i < b1; i += kBlock) dd_add(s, c, x[i], 0.0);
#pragma unroll
  for (int mask = 1; mask < kWave; mask <<= 1) {
    const double os = shfl_xor_d(s, mask), oc = shfl_xor_d(c, mask);
    dd_add(s, c, os, oc);
  }
  if ((threadIdx.x & (kWave - 1)) == 0) lds[threadIdx.x / kWave] = make_double2(s, c);
  __syncthreads();
  if (threadIdx.x == 0) {
    double2 t = lds[0];
    for (int w = 1; w < kBlock / kWave; ++w) dd_add(t.x, t.y, lds[w].x, lds[w].y);
    part[blockIdx.x] = t;
  }
}

__global__ __launch_bounds__(kBlock) void k_sum_final(const double2* part, int64_t nb, double* out)
{
  __shared__ double2 lds[kBlock / kWave];
  const int64_t chunk = (nb + kBlock - 1) / kBlock;
  const int64_t b0 = threadIdx.x * chunk;
  const int64_t b1 = b0 + chunk < nb ? b0 + chunk : nb;
  double s = 0.0, c = 0.0;
  for (int64_t b = b0; b < b1; ++b) dd_add(s, c, part[b].x, part[b].y);
#pragma unroll
  for (int mask = 1; mask < kWave; mask <<= 1) {
    const double os = shfl_xor_d(s, mask), oc = shfl_xor_d(c, mask);
    dd_add(s, c, os, oc);
  }
  if ((threadIdx.x & (kWave - 1)) == 0) lds[threadIdx.x / kWave] = make_double2(s, c);
  __syncthreads();
  if (threadIdx.x == 0) {
    double2 t = lds[0];
    for (int w = 1; w < kBlock / kWave; ++w) dd_add(t.x, t.y, lds[w].x, lds[w].y);
    out[0] = t.x + t.y;
  }
}

}  // namespace cuhre

using namespace cuhre;

// ------------------------------------------------------------------ context

namespace cuhre {

extern "C" __attribute__((visibility("hidden"))) BatchEngine* cuhre_ctx_batch(cuhre_context* c, hipStream_t* stream, int* device,
                                        int64_t* capacity)
{
  if (!c->batch) c->batch = batch_engine_new();
  *stream = c->stream;
  *device = c->device;
  *capacity = c->batch_capacity;
  return c->batch;
}

__global__ __launch_bounds__(kWave) void k_finish(EvalArgs A, uint32_t nb) { finish_iteration(A, nb); }
__global__ __launch_bounds__(kWave) void k_finish_groups(EvalArgs A, uint32_t nb, int rpb)
{
  finish_groups(A, nb, rpb);
}
__global__ __launch_bounds__(kMaxWaves * kWave) void k_finish_top(EvalArgs A, uint32_t G) { finish_top(A, G); }

void launch_finish(const EvalArgs& A, int64_t nb, hipStream_t st)
{
  if (nb <= (int64_t)kWaveFinalize) {
    hipLaunchKernelGGL(k_finish, dim3(1), dim3(kWave), 0, st, A, (uint32_t)nb);
    return;
  }
  const int64_t G = (nb + kWave - 1) / kWave;
  hipLaunchKernelGGL(k_finish_groups, dim3((unsigned)G), dim3(kWave), 0, st, A, (uint32_t)nb, 0);
  hipLaunchKernelGGL(k_finish_top, dim3(1), dim3(G <= (int64_t)kWaveFinalize ? kWave : kMaxWaves * kWave), 0, st,
                     A, (uint32_t)G);
}

int64_t blocks_for(int64_t n, int S)
{
  const int rpb = seg_regions_per_block(S);
  return (n + rpb - 1) / rpb;
}

static int ctx_free_lists(cuhre_context* c)
{
  for (int i = 0; i < 2; ++i) {
    if (c->lo[i]) (void)hipFree(c->lo[i]);
    if (c->wd[i]) (void)hipFree(c->wd[i]);
    c->lo[i] = c->wd[i] = nullptr;
  }
  if (c->est) (void)hipFree(c->est);
  if (c->err) (void)hipFree(c->err);
  if (c->abssum) (void)hipFree(c->abssum);
  if (c->axis) (void)hipFree(c->axis);
  if (c->partials) (void)hipFree(c->partials);
  if (c->offsets) (void)hipFree(c->offsets);
  c->est = c->err = c->abssum = nullptr;
  c->axis = nullptr;
  c->partials = nullptr;
  c->offsets = nullptr;
  c->cap = 0;
  return 0;
}

// Allocates list buffers for `cap` regions of `dim` coordinates.  When
// keep_n > 0 the first keep_n regions of list `cur` (and their est/err/axis,
// and keep_nb block offsets) are carried over into the new buffers.
int ctx_reserve(cuhre_context* c, int64_t cap, int dim, int64_t keep_n, int cur, int64_t keep_nb)
{
  if (cap <= c->cap && dim <= c->dim_alloc) return 0;
  // a longer list keeps the capacity it grew to; a list of more dimensions
  // is sized by the request (at the old capacity it may not fit beside the
  // old buffers), and without contents to carry the old buffers go first
  if (keep_n > 0 || dim <= c->dim_alloc) cap = std::max(cap, c->cap);
  dim = std::max(dim, c->dim_alloc);
  cap = (cap + 255) & ~255ll;
  if (keep_n == 0) ctx_free_lists(c);
  const int64_t nbmax = block_capacity(cap);
  double *lo[2], *wd[2], *est, *err, *ab;
  int8_t* ax;
  Partial* part;
  int64_t* off;
  const size_t lb = sizeof(double) * (size_t)cap * dim;
  if (hipMalloc(&lo[0], lb) != hipSuccess || hipMalloc(&lo[1], lb) != hipSuccess ||
      hipMalloc(&wd[0], lb) != hipSuccess || hipMalloc(&wd[1], lb) != hipSuccess ||
      hipMalloc(&est, sizeof(double) * cap) != hipSuccess ||
      hipMalloc(&err, sizeof(double) * cap) != hipSuccess ||
      hipMalloc(&ab, sizeof(double) * cap) != hipSuccess || hipMalloc(&ax, cap) != hipSuccess ||
      hipMalloc(&part, sizeof(Partial) * (nbmax + group_capacity(nbmax))) != hipSuccess ||
      hipMalloc(&off, sizeof(int64_t) * (nbmax + group_capacity(nbmax))) != hipSuccess) {
    (void)hipGetLastError();
    c->avail_bytes = -1.0;
    return fail(CUHRE_ERR_ALLOC, "cannot allocate a region list of %lld regions x %d dims",
                (long long)cap, dim);
  }
  if (keep_n > 0) {
    const int dk = c->dim_alloc;
    HIP_TRY(hipMemcpy2DAsync(lo[cur], sizeof(double) * cap, c->lo[cur], sizeof(double) * c->cap,
                             sizeof(double) * keep_n, dk, hipMemcpyDeviceToDevice, c->stream));
    HIP_TRY(hipMemcpy2DAsync(wd[cur], sizeof(double) * cap, c->wd[cur], sizeof(double) * c->cap,
                             sizeof(double) * keep_n, dk, hipMemcpyDeviceToDevice, c->stream));
    HIP_TRY(hipMemcpyAsync(est, c->est, sizeof(double) * keep_n, hipMemcpyDeviceToDevice, c->stream));
    HIP_TRY(hipMemcpyAsync(err, c->err, sizeof(double) * keep_n, hipMemcpyDeviceToDevice, c->stream));
    HIP_TRY(hipMemcpyAsync(ax, c->axis, keep_n, hipMemcpyDeviceToDevice, c->stream));
    HIP_TRY(hipMemcpyAsync(off, c->offsets, sizeof(int64_t) * keep_nb, hipMemcpyDeviceToDevice,
                           c->stream));
    // and the group bases of a two-level finalisation (behind the block offsets)
    const int64_t keep_ng = (keep_nb + kWave - 1) / kWave;
    HIP_TRY(hipMemcpyAsync(off + nbmax, c->offsets + block_capacity(c->cap), sizeof(int64_t) * keep_ng,
                           hipMemcpyDeviceToDevice, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
  }
  ctx_free_lists(c);
  c->lo[0] = lo[0];
  c->lo[1] = lo[1];
  c->wd[0] = wd[0];
  c->wd[1] = wd[1];
  c->est = est;
  c->err = err;
  c->abssum = ab;
  c->axis = ax;
  c->partials = part;
  c->offsets = off;
  c->cap = cap;
  c->dim_alloc = dim;
  c->avail_bytes = -1.0;
  return 0;
}

int ctx_reserve_aux(cuhre_context* c, int64_t cap, int dim)
{
  const int64_t need = cap * (1 + dim);
  if (need <= c->aux_cap) return 0;
  if (c->aux) (void)hipFree(c->aux);
  c->aux = nullptr;
  c->aux_cap = 0;
  HIP_TRY(hipMalloc(&c->aux, sizeof(double) * need));
  c->aux_cap = need;
  return 0;
}

// LC_LT_t tables are float on the device (the reference's textures): the
// no-allocation path uploads the caller's doubles and rounds them here.
__global__ __launch_bounds__(kBlock) void k_table_to_float(const double* src, float* dst, int64_t n)
{
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i < n) dst[i] = (float)src[i];
}

int make_dev_integrand(cuhre_context* c, const cuhre_integrand* f, int dim, DevIntegrand* P, char* table_dst,
                       size_t table_bytes)
{
  std::memset(P, 0, sizeof *P);
  P->family = f->family;
  P->dim = dim;
  P->exponent = f->exponent;
  P->nterms = f->nterms;
  if (f->nterms < 0 || f->nterms > CUHRE_MAXTERMS) return fail(CUHRE_ERR_ARG, "nterms out of range");
  for (int i = 0; i < kMaxDim; ++i) {
    P->a[i] = f->a[i];
    P->w[i] = f->w[i];
    P->b[i] = (i < dim && f->family == CUHRE_PRODUCT_PEAK) ? 1.0 / (f->a[i] * f->a[i]) : 0.0;
  }
  P->s0 = 2.0 * M_PI * f->u;
  for (int t = 0; t < CUHRE_MAXTERMS; ++t) {
    P->coef[t] = f->coef[t];
    for (int i = 0; i < kMaxDim; ++i) {
      const int e = f->powers[t][i];
      if (t < f->nterms && (e < 0 || e > 127)) return fail(CUHRE_ERR_ARG, "polynomial power out of range");
      P->powers[t][i] = (int8_t)e;
    }
  }
  if (f->family == CUHRE_CORNER_PEAK && f->exponent < 0)
    return fail(CUHRE_ERR_ARG, "corner peak exponent must be >= 0");
  if (f->family == CUHRE_TABLE6) {
    if (!f->table || f->table_nx < 2 || f->table_ny < 2)
      return fail(CUHRE_ERR_ARG, "table integrand needs a table of at least 2x2");
    const size_t cnt = (size_t)f->table_nx * f->table_ny;
    if (table_dst) {
      if (cnt * sizeof(double) > table_bytes)
        return fail(CUHRE_ERR_CAPACITY, "table of %zu values exceeds the static workspace's %zu bytes", cnt,
                    table_bytes);
      HIP_TRY(hipMemcpyAsync(table_dst, f->table, cnt * sizeof(double), hipMemcpyHostToDevice, c->stream));
      P->table = reinterpret_cast<const double*>(table_dst);
      P->nx = f->table_nx;
      P->ny = f->table_ny;
      return 0;
    }
    const bool same = c->table && c->table_host.size() == cnt &&
                      std::memcmp(c->table_host.data(), f->table, cnt * sizeof(double)) == 0;
    if (!same) {
      if (c->table) (void)hipFree(c->table);
      c->table = nullptr;
      c->table_host.assign(f->table, f->table + cnt);
      HIP_TRY(hipMalloc(&c->table, cnt * sizeof(double)));
      HIP_TRY(hipMemcpyAsync(c->table, f->table, cnt * sizeof(double), hipMemcpyHostToDevice, c->stream));
    }
    P->table = c->table;
    P->nx = f->table_nx;
    P->ny = f->table_ny;
  }
  if (f->family == CUHRE_LC_LT_T) {
    const int nx = f->table_nx, ny = f->table_ny;
    if (dim != 3) return fail(CUHRE_ERR_ARG, "the LC_LT_t integrand has 3 dimensions (lt, zt, lc)");
    if (!f->table || nx < 2 || ny < 2 || nx > 4096 || ny > 4096)
      return fail(CUHRE_ERR_ARG, "LC_LT_t needs node arrays of 2..4096 entries and five fields");
    const size_t cnt = (size_t)nx + ny + 5 * (size_t)nx * ny;
    // checks on the values as the device will hold them (rounded to float)
    auto f32 = [&](size_t k) { return (float)f->table[k]; };
    for (size_t k = 0; k < cnt; ++k)
      if (!std::isfinite(f32(k))) return fail(CUHRE_ERR_ARG, "LC_LT_t table value %zu is not finite", k);
    for (int i = 1; i < nx; ++i)
      if (!(f32(i) > f32(i - 1))) return fail(CUHRE_ERR_ARG, "LC_LT_t lt nodes must increase strictly");
    for (int j = 1; j < ny; ++j)
      if (!(f32(nx + j) > f32(nx + j - 1))) return fail(CUHRE_ERR_ARG, "LC_LT_t zt nodes must increase strictly");
    if (table_dst) {
      // doubles first, then their float rounding behind them
      const size_t dbytes = (cnt * sizeof(double) + 255) & ~(size_t)255;
      if (dbytes + cnt * sizeof(float) > table_bytes)
        return fail(CUHRE_ERR_CAPACITY, "LC_LT_t table of %zu values exceeds the static workspace's %zu bytes",
                    cnt, table_bytes);
      double* dsrc = reinterpret_cast<double*>(table_dst);
      float* dst = reinterpret_cast<float*>(table_dst + dbytes);
      HIP_TRY(hipMemcpyAsync(dsrc, f->table, cnt * sizeof(double), hipMemcpyHostToDevice, c->stream));
      hipLaunchKernelGGL(k_table_to_float, dim3((unsigned)((cnt + kBlock - 1) / kBlock)), dim3(kBlock), 0,
                         c->stream, dsrc, dst, (int64_t)cnt);
      HIP_TRY(hipGetLastError());
      P->ftab = dst;
      P->nx = nx;
      P->ny = ny;
      return 0;
    }
    std::vector<float> t32(cnt);
    for (size_t k = 0; k < cnt; ++k) t32[k] = f32(k);
    const bool same = c->ftable && c->ftable_host.size() == cnt &&
                      std::memcmp(c->ftable_host.data(), t32.data(), cnt * sizeof(float)) == 0;
    if (!same) {
      if (c->ftable) (void)hipFree(c->ftable);
      c->ftable = nullptr;
      c->ftable_host = t32;
      HIP_TRY(hipMalloc(&c->ftable, cnt * sizeof(float)));
      HIP_TRY(hipMemcpyAsync(c->ftable, c->ftable_host.data(), cnt * sizeof(float), hipMemcpyHostToDevice,
                             c->stream));
    }
    P->ftab = c->ftable;
    P->nx = nx;
    P->ny = ny;
  }
  return 0;
}

// Point segments per region: the smallest S with n*S >= 256k lanes, and at
// least 2 below 2M regions (measured on MI355X, 8-D product peak and
// oscillatory: S=16 is best up to 16k regions, S=4 at 64k-256k, S=2 at 1M).
// Regions per list: the context's setting, else what three quarters of the
// free HBM holds for two lists of this dimension (buffers the context
// already owns count as free).
int64_t list_capacity(cuhre_context* c, int dim)
{
  if (c->capacity > 0) return c->capacity;
  // the context's own list allocations move bytes between "free" and
  // "held", so the sum is measured (a runtime query of several us) only
  // when first needed and after the lists are reallocated
  if (c->avail_bytes < 0) {
    size_t fr = 0, tot = 0;
    if (hipMemGetInfo(&fr, &tot) != hipSuccess) return 64ll << 20;
    const double held = (double)c->cap * (4.0 * sizeof(double) * c->dim_alloc + 3 * sizeof(double) + 1);
    c->avail_bytes = (double)fr + held;
  }
  const double per = 4.0 * sizeof(double) * dim + 3 * sizeof(double) + 1 + sizeof(Partial) / 64.0 + 1;
  return (int64_t)(c->avail_bytes * 0.75 / per);
}

// Whether the family's rule evaluation unrolls its point loops at dim (kUnroll).
bool family_unrolls(int family, int dim)
{
  switch (family) {
  case CUHRE_GAUSSIAN: return Family<CUHRE_GAUSSIAN, 2>::unrolls(dim);
  case CUHRE_PRODUCT_PEAK: return Family<CUHRE_PRODUCT_PEAK, 2>::unrolls(dim);
  case CUHRE_OSCILLATORY: return Family<CUHRE_OSCILLATORY, 2>::unrolls(dim);
  case CUHRE_CORNER_PEAK: return Family<CUHRE_CORNER_PEAK, 2>::unrolls(dim);
  case CUHRE_CONTINUOUS: return Family<CUHRE_CONTINUOUS, 2>::unrolls(dim);
  case CUHRE_POLYNOMIAL: return Family<CUHRE_POLYNOMIAL, 2>::unrolls(dim);
  case CUHRE_COSINE_SUM: return Family<CUHRE_COSINE_SUM, 2>::unrolls(dim);
  case CUHRE_TABLE6: return Family<CUHRE_TABLE6, 6>::unrolls(dim);
  case CUHRE_LC_LT_T: return Family<CUHRE_LC_LT_T, 3>::unrolls(dim);
  default: return false;
  }
}

// Segment count of an evaluation launch over n regions (the point-segment
// sweep of the 8-D product peak and oscillatory integrand on MI355X): the
// half-wavefront form up to half_max_regions, above it the smallest S with
// n*S >= fill_lanes lanes, a quarter of that for families whose one-segment
// kernel unrolls its point loops (it is the fastest from 64k regions up);
// the degree-9 rule keeps S >= 2 below 2M regions.
int pick_segments(const cuhre_context* c, int64_t n, int key, int family, int dim)
{
  const int smax = key == 9 ? 8 : kMaxWaves;  // the degree-9 merge needs more LDS per wave
  // the half-wavefront form needs n/8 block partials
  const bool half_fits = (n + 7) / 8 <= block_capacity(c->cap);
  if (c->segments == kHalfS && half_fits) return kHalfS;
  if (c->segments > 0) return std::min(c->segments, smax);
  if (n <= c->half_max_regions && half_fits) return kHalfS;
  const int64_t lanes = key != 9 && family_unrolls(family, dim) ? c->fill_lanes / 4 : c->fill_lanes;
  int S = 1;
  while (S < smax && n * S < lanes) S *= 2;
  if (key == 9 && S == 1 && n < (2ll << 20)) S = 2;
  return S;
}

// The initial grid written on the device (lower = low + id * w with the
// product rounded first, width w = (high - low) / k, as the oracle's
// or_initial_grid), instead of two host-to-device copies per call.
struct GridArgs {
  double* lo;
  double* wd;
  int64_t ld, n;
  int32_t d, k;
  double lows[kMaxDim], highs[kMaxDim];
};

__global__ __launch_bounds__(kBlock) void k_initial_grid(GridArgs G)
{
#pragma clang fp contract(off)  // low + id * w rounded twice, as the oracle computes it
  const int64_t r = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (r >= G.n) return;
  int64_t q = r;
  for (int m = 0; m < G.d; ++m) {
    const int64_t id = q % G.k;
    q /= G.k;
    const double w = (G.highs[m] - G.lows[m]) / (double)G.k;
    G.lo[m * G.ld + r] = G.lows[m] + (double)id * w;
    G.wd[m * G.ld + r] = w;
  }
}

int upload_initial_grid(cuhre_context* c, int list, int d, const double* lows, const double* highs, int64_t n0)
{
  GridArgs G;
  G.lo = c->lo[list];
  G.wd = c->wd[list];
  G.ld = c->cap;
  G.n = n0;
  G.d = d;
  G.k = grid_k(d);
  for (int m = 0; m < kMaxDim; ++m) {
    G.lows[m] = m < d ? lows[m] : 0.0;
    G.highs[m] = m < d ? highs[m] : 0.0;
  }
  hipLaunchKernelGGL(k_initial_grid, dim3((unsigned)((n0 + kBlock - 1) / kBlock)), dim3(kBlock), 0, c->stream, G);
  return hipGetLastError() == hipSuccess ? 0 : fail(CUHRE_ERR_DEVICE, "initial grid launch failed");
}

int check_common(cuhre_context* ctx, const cuhre_integrand* f, int dim, int key)
{
  if (!ctx) return fail(CUHRE_ERR_ARG, "null context");
  int rc = check_integrand_args(f, dim, key);
  if (rc) return rc;
  if (!find_eval_launcher(f->family, dim, key))
    return fail(CUHRE_ERR_UNSUPPORTED, "family %d not compiled for dim %d", f->family, dim);
  HIP_TRY(hipSetDevice(ctx->device));
  return 0;
}

void fill_eval_args(cuhre_context* c, EvalArgs& A, int cur, int64_t n, int dim, double epsrel,
                           double epsabs)
{
  std::memset(&A, 0, sizeof A);
  A.lower = c->lo[cur];
  A.width = c->wd[cur];
  A.ld = c->cap;
  A.n = n;
  A.epsrel = epsrel;
  A.epsabs = epsabs;
  A.loc_epsrel = epsrel;
  A.loc_epsabs = epsabs;
  A.est = c->est;
  A.err = c->err;
  A.axis = c->axis;
  A.abssum = c->abssum;
  A.partials = c->partials;
  A.offsets = c->offsets;
  // the group sums of the two-level finalisation follow the block partials
  A.gpart = c->partials + block_capacity(c->cap);
  A.gbase = c->offsets + block_capacity(c->cap);
  A.state = c->state;
  (void)dim;
}

void fill_split_args(cuhre_context* c, SplitArgs& S, int cur, int64_t n, int dim, double epsrel,
                            double epsabs, int rpb, bool use_state)
{
  S.lower = c->lo[cur];
  S.width = c->wd[cur];
  S.ld = c->cap;
  S.n = n;
  S.est = c->est;
  S.err = c->err;
  S.axis = c->axis;
  S.epsrel = epsrel;
  S.epsabs = epsabs;
  S.offsets = c->offsets;
  S.rpb = rpb;
  S.dim = dim;
  S.out_lower = c->lo[1 - cur];
  S.out_width = c->wd[1 - cur];
  S.out_ld = c->cap;
  S.state = use_state ? c->state : nullptr;
  S.retire = nullptr;
  S.retire_on = nullptr;
  S.gbase = nullptr;
}

void dd_add_host(double& s1, double& c1, double s2, double c2)
{
  const double s = s1 + s2;
  const double bb = s - s1;
  const double e = (s1 - (s - bb)) + (s2 - bb);
  s1 = s;
  c1 = c1 + c2 + e;
}

}  // namespace cuhre

extern "C" {

}  // extern "C"
static int scratch_setup(int device);
extern "C" {

int cuhre_context_create(cuhre_context** out, int device, int64_t capacity_regions)
{
  if (!out) return fail(CUHRE_ERR_ARG, "null output pointer");
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
    (void)hipGetLastError();
    return fail(CUHRE_ERR_DEVICE, "no HIP device available");
  }
  if (device < 0 || device >= ndev) return fail(CUHRE_ERR_ARG, "device %d out of range (%d)", device, ndev);
  cuhre_context* c = new cuhre_context();
  c->device = device;
  if (capacity_regions > 0) {
    c->capacity = capacity_regions;
    c->batch_capacity = capacity_regions;
  }
  int rc = 0;
  if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
      hipMalloc(&c->state, sizeof(IterState)) != hipSuccess ||
      hipHostMalloc(&c->h_state, sizeof(IterState), hipHostMallocDefault) != hipSuccess ||
      hipMalloc(&c->spec, 2 * sizeof(SpecSlot)) != hipSuccess ||
      hipHostMalloc(&c->pub_state, 2 * sizeof(IterState), hipHostMallocCoherent | hipHostMallocMapped) != hipSuccess ||
      hipHostMalloc(&c->pub_seq, sizeof(uint32_t), hipHostMallocCoherent | hipHostMallocMapped) != hipSuccess ||
      hipHostGetDevicePointer((void**)&c->pub_state_dev, c->pub_state, 0) != hipSuccess ||
      hipHostGetDevicePointer((void**)&c->pub_seq_dev, c->pub_seq, 0) != hipSuccess ||
      hipMalloc(&c->sum_out, sizeof(double)) != hipSuccess) {
    rc = fail(CUHRE_ERR_DEVICE, "context setup failed: %s", hipGetErrorString(hipGetLastError()));
  }
  if (rc == 0) *c->pub_seq = 0;  // launches are numbered from 1
  if (rc == 0) {
    // two sets of per-block partials, dense counts and active-region lists
    // for the fused half-wavefront iterations
    const size_t part = al256(sizeof(Partial) * kOneLevel), cnt = al256(sizeof(int64_t) * kOneLevel);
    const size_t act = al256((size_t)kOneLevel * (kBlock / kHalfS));
    const size_t set = part + cnt + act;
    if (hipMalloc(&c->fuse_mem, 2 * set) != hipSuccess || hipMemset(c->fuse_mem, 0, 2 * set) != hipSuccess) {
      rc = fail(CUHRE_ERR_DEVICE, "context setup failed: %s", hipGetErrorString(hipGetLastError()));
    } else {
      for (int i = 0; i < 2; ++i) {
        char* b = c->fuse_mem + i * set;
        c->fuse[i].part = (Partial*)b;
        c->fuse[i].cnt = (int64_t*)(b + part);
        c->fuse[i].act = (uint8_t*)(b + part + cnt);
      }
    }
  }
  for (int i = 0; i < 4 && rc == 0; ++i)
    if (hipEventCreate(&c->ev[i]) != hipSuccess) rc = fail(CUHRE_ERR_DEVICE, "event create failed");
  // the static workspace's address on this device (the no-allocation entry
  // point); a failure here is reported again by that entry point
  if (rc == 0) (void)scratch_setup(device);
  if (rc != 0) {
    cuhre_context_destroy(c);
    return rc;
  }
  *out = c;
  return 0;
}

int cuhre_context_destroy(cuhre_context* c)
{
  if (!c) return 0;
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  for (cuhre_guide* g : c->guides) cuhre::guide_detach(g);
  c->guides.clear();
  ctx_free_lists(c);
  batch_engine_delete(c->batch);
  c->batch = nullptr;
  if (c->aux) (void)hipFree(c->aux);
  if (c->p2_ws) (void)hipFree(c->p2_ws);
  if (c->state) (void)hipFree(c->state);
  if (c->spec) (void)hipFree(c->spec);
  if (c->fuse_mem) (void)hipFree(c->fuse_mem);
  if (c->h_state) (void)hipHostFree(c->h_state);
  if (c->pub_state) (void)hipHostFree(c->pub_state);
  if (c->pub_seq) (void)hipHostFree(c->pub_seq);
  if (c->sum_part) (void)hipFree(c->sum_part);
  if (c->sum_out) (void)hipFree(c->sum_out);
  if (c->table) (void)hipFree(c->table);
  if (c->ftable) (void)hipFree(c->ftable);
  for (int i = 0; i < 4; ++i)
    if (c->ev[i]) (void)hipEventDestroy(c->ev[i]);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return 0;
}

int cuhre_context_set(cuhre_context* c, int profile, int point_segments)
{
  if (!c) return fail(CUHRE_ERR_ARG, "null context");
  if (point_segments != 0 && point_segments != 1 && point_segments != 2 && point_segments != 4 &&
      point_segments != 8 && point_segments != 16 && point_segments != kHalfS)
    return fail(CUHRE_ERR_ARG, "point_segments must be 0, 1, 2, 4, 8, 16 or 32");
  c->profile = profile;
  c->segments = point_segments;
  return 0;
}

int cuhre_context_queue_ahead(cuhre_context* c, int on)
{
  if (!c) return fail(CUHRE_ERR_ARG, "null context");
  c->queue_ahead = on != 0;
  return 0;
}

namespace {

// Waits for the launch numbered seq to publish its state (its finalisation
// writes the state to buffer seq & 1, then the number, to coherent host
// memory).  Meanwhile the launches queued behind it keep the GPU busy; the
// next one may already have published (number seq + 1, the other buffer).
int wait_published(cuhre_context* c, uint32_t seq, IterState* out)
{
  const auto reached = [&] { return (int32_t)(__atomic_load_n(c->pub_seq, __ATOMIC_ACQUIRE) - seq) >= 0; };
  for (uint64_t spin = 0;; ++spin) {
    if (reached()) {
      std::memcpy(out, (const void*)(c->pub_state + (seq & 1)), sizeof *out);
      return 0;
    }
    __builtin_ia32_pause();
    // the stream query (a runtime call) only guards against a launch that
    // never publishes; asked every ~2^18 spins it stays off the hot path
    if ((spin & 262143) == 262143) {
      const hipError_t q = hipStreamQuery(c->stream);
      if (q == hipSuccess && !reached())
        return fail(CUHRE_ERR_DEVICE, "evaluation ended without publishing its state");
      if (q != hipSuccess && q != hipErrorNotReady)
        return fail(CUHRE_ERR_DEVICE, "HIP error: %s", hipGetErrorString(q));
    }
  }
}

}  // namespace

}  // extern "C"

namespace {

// Options of the breadth-first loop: stop after max_iter iterations; with
// handoff_n > 0, stop once a split leaves handoff_n regions or more (or the
// iterations run out with regions left) and hand that list over.
struct LoopOpts {
  int64_t handoff_n = 0;
  int max_iter = 1000;
  // no-allocation mode: the context's list buffers already point into the
  // static workspace and must not grow; a run that outgrows them ends with
  // CUHRE_ERR_CAPACITY; tables go to table_dst
  bool noalloc = false;
  char* table_dst = nullptr;
  size_t table_bytes = 0;
};

struct LoopHandoff {
  int handed = 0;
  int cur = 0;
  int64_t n = 0;
  double fin_s = 0, fin_c = 0, fer_s = 0, fer_c = 0, fin_abs = 0, fin_aest = 0;
  int64_t nfin = 0;
};

int plain_loop(cuhre_context* c, const cuhre_integrand* f, int dim, double epsrel, double epsabs, int64_t maxeval,
               int key, const double* lows, const double* highs, cuhre_result* out, const LoopOpts& opt,
               LoopHandoff* ho)
{
  const auto t0 = std::chrono::steady_clock::now();
  if (!out) return fail(CUHRE_ERR_ARG, "null result");
  std::memset(out, 0, sizeof *out);
  int rc = check_common(c, f, dim, key);
  if (rc) return rc;
  if ((rc = check_bounds(dim, epsrel, epsabs, lows, highs))) return rc;

  DevIntegrand P;
  if ((rc = make_dev_integrand(c, f, dim, &P, opt.table_dst, opt.table_bytes))) return rc;
  const RuleConst R = make_rule(dim, key);
  const int64_t p = R.npts;
  const double rflops = region_flops(f, dim, key);

  const int64_t n0 = initial_grid_count(dim);
  const int64_t capacity = list_capacity(c, dim);
  if (n0 > capacity) return fail(CUHRE_ERR_CAPACITY, "initial grid exceeds capacity");
  const int64_t want = std::min<int64_t>(capacity, std::max<int64_t>(4 * n0, 1 << 16));
  if ((rc = ctx_reserve(c, want, dim))) return rc;
  int cur = 0;
  if ((rc = upload_initial_grid(c, cur, dim, lows, highs, n0))) return rc;

  double fin_s = 0.0, fin_c = 0.0, fer_s = 0.0, fer_c = 0.0, fin_abs = 0.0, fin_aest = 0.0;
  int64_t nfin = 0, n = n0;
  int result = CUHRE_NOT_CONVERGED;
  const EvalLauncher launch = find_eval_launcher(f->family, dim, key);
  // Queue-ahead: iteration it+1 (its evaluation, sized for the bound 2 n of
  // its list, and its split) enters the stream before the host has read
  // iteration it's state, so the GPU does not idle through the host's round
  // trip.  Its list length and finished sums come from the device slot the
  // split of iteration it writes (length 0 when the loop stops there).  The
  // segment count of every launch follows from that bound (2 n of the
  // previous list, the initial grid first) and the reductions from the real
  // length, so results do not depend on which launches were queued ahead.
  const bool ahead = c->queue_ahead && !c->profile && opt.handoff_n == 0;
  struct Launched {
    int S = 1;
    int64_t nb = 0;  // evaluation blocks of the launch
    uint32_t seq = 0;
    bool slot = false;       // its finalisation writes the next list's slot
    bool fuse_next = false;  // the next iteration's k_eval_fused splits and finalises it
    int lc = 0;              // its list
    int set = 0;             // its fused-chain buffers (c->fuse[set])
    EvalArgs A;              // its finalisation's arguments
    SpecArgs Q;
  };
  const auto fused = [](int64_t nb) {
    return nb <= (int64_t)kOneLevel || (nb + kWave - 1) / kWave <= (int64_t)kWaveFinalize;
  };
  const FusedLauncher fuse_launch = find_fused_launcher(f->family, dim);
  // Half-wavefront iterations queued ahead form a chain: each writes its
  // partials, dense counts, split axes and activity masks to one of two
  // buffer sets, and the next one (k_eval_fused) finds its regions from
  // them, finalising and splitting it in the same launch.
  const auto chains = [&](int S, int64_t bnd) {
    return ahead && fuse_launch != nullptr && key != 9 && c->segments == 0 && S == kHalfS && bnd <= c->cap / 2;
  };
  // Queues the evaluation and split of list lc: nl regions (exact) or at
  // most nl, the length then read on the device; bnd picks the segments;
  // nev: evaluations of the iterations before; iter: the iteration.  With
  // parent set, the evaluation is k_eval_fused over the parent's children.
  const auto queue = [&](int lc, int64_t nl, int64_t bnd, bool exact, int64_t nev, int iter, const Launched* parent,
                         Launched& L) -> int {
    L.S = pick_segments(c, bnd, key, f->family, dim);
    const int rpb = seg_regions_per_block(L.S);
    L.nb = blocks_for(nl, L.S);
    if (L.nb > 0x7fffffffll) return fail(CUHRE_ERR_CAPACITY, "grid too large");
    L.lc = lc;
    L.set = iter & 1;
    EvalArgs A;
    fill_eval_args(c, A, lc, nl, dim, epsrel, epsabs);
    A.finalize = 1;
    A.alloc_active = c->cap / 2;
    SpecArgs Q;
    std::memset(&Q, 0, sizeof Q);
    if (exact) {
      A.fin_est_s = fin_s;
      A.fin_est_c = fin_c;
      A.fin_err_s = fer_s;
      A.fin_err_c = fer_c;
      const int64_t left = maxeval - (nev + nl * p);
      const int64_t budget_active = left >= 0 ? left / (2 * p) : -1;
      A.max_active = std::min<int64_t>(budget_active, capacity / 2);
    } else {
      A.n_dev = &c->spec[lc].n;
      Q.in = &c->spec[lc];
      Q.left0 = maxeval - nev;
      Q.cap2 = capacity / 2;
      Q.npts = p;
    }
    const bool chain = chains(L.S, bnd);
    L.fuse_next = chain && iter + 1 < opt.max_iter && 2 * bnd <= c->half_max_regions &&
                  pick_segments(c, 2 * bnd, key, f->family, dim) == kHalfS;
    if (chain) {
      A.partials = c->fuse[L.set].part;
      A.offsets = c->fuse[L.set].cnt;
      A.act_out = c->fuse[L.set].act;
    }
    L.slot = ahead && fused(L.nb);
    if (L.slot) Q.out = &c->spec[1 - lc];
    Q.rpb = rpb;
    SplitArgs SA;
    fill_split_args(c, SA, lc, nl, dim, epsrel, epsabs, rpb, true);
    if (L.nb > (int64_t)kOneLevel) SA.gbase = A.gbase;
    // without profiling the state comes back through coherent host memory
    // (two buffers: the split queued behind may publish before the host reads)
    if (!c->profile) {
      L.seq = A.seq = ++c->seq;
      A.host_state = c->pub_state_dev + (A.seq & 1);
      A.host_seq = c->pub_seq_dev;
    } else {
      // the kernel's own start and end (no launch gap in the interval)
      A.prof_start = c->ev[0];
      A.prof_stop = c->ev[1];
    }
    if (parent != nullptr) {
      FuseArgs F;
      F.R = parent->A;
      F.P = parent->Q;
      F.plo = c->lo[parent->lc];
      F.pwd = c->wd[parent->lc];
      F.pact = c->fuse[parent->set].act;
      F.clo = c->lo[lc];
      F.cwd = c->wd[lc];
      fuse_launch(A, F, P, R, c->stream);
    } else {
      launch(A, P, R, L.S, c->stream);
    }
    A.prof_start = A.prof_stop = nullptr;
    L.A = A;
    L.Q = Q;
    if (L.fuse_next) {  // split and finalised by the next iteration's evaluation
      HIP_TRY(hipGetLastError());
      return 0;
    }
    // the split's blocks and one more for the finalisation
    const dim3 sgrid((unsigned)((nl + kBlock - 1) / kBlock + 1));
    const int64_t G = (L.nb + kWave - 1) / kWave;
    if (L.nb <= (int64_t)kOneLevel) {
      hipLaunchKernelGGL(k_split_fin, sgrid, dim3(kBlock), 0, c->stream, SA, A, (const Partial*)A.partials,
                         (uint32_t)L.nb, 0, Q);
    } else if (G <= (int64_t)kWaveFinalize) {
      hipLaunchKernelGGL(k_finish_groups, dim3((unsigned)G), dim3(kWave), 0, c->stream, A, (uint32_t)L.nb, rpb);
      hipLaunchKernelGGL(k_split_fin, sgrid, dim3(kBlock), 0, c->stream, SA, A, (const Partial*)A.gpart,
                         (uint32_t)G, 1, Q);
    } else {
      launch_finish(A, L.nb, c->stream);
      hipLaunchKernelGGL(k_split, dim3(sgrid.x - 1), dim3(kBlock), 0, c->stream, SA);
    }
    if (c->profile) HIP_TRY(hipEventRecord(c->ev[2], c->stream));
    HIP_TRY(hipGetLastError());
    return 0;
  };

  int64_t bound = n0;  // 2 n of the previous list (the initial grid first)
  Launched L, next;
  bool queued = false;  // iteration it is already in the stream (as next)
  for (int it = 0; it < opt.max_iter; ++it) {
    if (queued) L = next;
    else if ((rc = queue(cur, n, bound, true, out->neval, it, nullptr, L))) return rc;
    queued = false;
    // iteration it+1 is queued now; it must be when its evaluation is to
    // split and finalise iteration it (its state is published from there)
    if (L.fuse_next || (L.slot && it + 1 < opt.max_iter)) {
      const int S2 = pick_segments(c, 2 * n, key, f->family, dim);
      const int64_t nb2 = blocks_for(2 * n, S2);
      if (L.fuse_next || (fused(nb2) && nb2 <= block_capacity(c->cap))) {
        if ((rc = queue(1 - cur, 2 * n, 2 * n, false, out->neval + n * p, it + 1, L.fuse_next ? &L : nullptr,
                        next)))
          return rc;
        queued = true;
      }
    }
    IterState st;
    if (!c->profile) {
      if ((rc = wait_published(c, L.seq, &st))) return rc;
    } else {
      HIP_TRY(hipMemcpyAsync(c->h_state, c->state, sizeof(IterState), hipMemcpyDeviceToHost, c->stream));
      HIP_TRY(hipStreamSynchronize(c->stream));
      st = *c->h_state;
    }
    const int64_t left = maxeval - (out->neval + n * p);
    const int64_t budget_active = left >= 0 ? left / (2 * p) : -1;
    const int S = L.S;
    const int64_t nb = blocks_for(n, S);
    if (c->profile) {
      float ms_e = 0.f, ms_s = 0.f;
      HIP_TRY(hipEventElapsedTime(&ms_e, c->ev[0], c->ev[1]));
      HIP_TRY(hipEventElapsedTime(&ms_s, c->ev[1], c->ev[2]));
      out->ms_eval += ms_e;
      out->ms_split += ms_s;
      if (n >= out->largest_regions) {
        out->largest_regions = n;
        out->ms_eval_largest = ms_e;
      }
    }
    out->eval_launches++;
    out->eval_flops += rflops * (double)n;
    if (st.stop == 5 && opt.noalloc) {
      // cannot happen (max_active <= capacity / 2 = alloc_active stops first)
      result = fail(CUHRE_ERR_CAPACITY, "static workspace too small for the children of %lld regions",
                    (long long)st.tot.nactive);
      break;
    }
    if (st.stop == 5) {
      // list buffers too small for the children: grow (keeping this
      // iteration's list and results) and run the split again; an
      // iteration queued behind this one read length 0 and did nothing
      queued = false;
      const int64_t A2 = 2 * st.tot.nactive;
      const int64_t ncap = std::min<int64_t>(capacity, std::max<int64_t>(A2, 2 * c->cap));
      if ((rc = ctx_reserve(c, ncap, dim, n, cur, nb))) return rc;
      SplitArgs SA;
      fill_split_args(c, SA, cur, n, dim, epsrel, epsabs, seg_regions_per_block(S), false);
      if (nb > (int64_t)kOneLevel) SA.gbase = c->offsets + block_capacity(c->cap);
      if (c->profile) HIP_TRY(hipEventRecord(c->ev[1], c->stream));
      // one level: offsets[] still holds the blocks' dense counts
      if (nb <= (int64_t)kOneLevel)
        hipLaunchKernelGGL(k_scan_counts, dim3(1), dim3(kBlock), 0, c->stream, c->offsets, (uint32_t)nb);
      hipLaunchKernelGGL(k_split, dim3((unsigned)((n + kBlock - 1) / kBlock)), dim3(kBlock), 0, c->stream, SA);
      if (c->profile) HIP_TRY(hipEventRecord(c->ev[2], c->stream));
      HIP_TRY(hipGetLastError());
      if (c->profile) {
        HIP_TRY(hipStreamSynchronize(c->stream));
        float ms_s = 0.f;
        HIP_TRY(hipEventElapsedTime(&ms_s, c->ev[1], c->ev[2]));
        out->ms_split += ms_s;
      }
      st.stop = 0;
    }
    out->split_launches++;
    out->iterations = it + 1;
    out->nregions += n;
    out->neval += n * p;
    out->peak_regions = std::max(out->peak_regions, n);
    const double abs_contrib = fin_abs + st.tot.abs_all;
    const double aest_contrib = fin_aest + st.tot.aest_all;
    const int64_t ncontrib = nfin + n;
    dd_add_host(fin_s, fin_c, st.tot.fest_s, st.tot.fest_c);
    dd_add_host(fer_s, fer_c, st.tot.ferr_s, st.tot.ferr_c);
    fin_abs += st.tot.abs_fin;
    fin_aest += st.tot.aest_fin;
    nfin += n - st.tot.nactive;
    out->nfinished = nfin;
    out->integral = st.rG_s + st.rG_c;
    out->error = st.eG_s + st.eG_c;
    out->fp64_err_bound = fp64_rounding_bound(p, dim, ncontrib, abs_contrib, aest_contrib, out->integral);
    if (st.stop == 4) {
      result = fail(CUHRE_ERR_NONFINITE, "integrand returned a non-finite value in %lld region(s)",
                    (long long)st.tot.nnonfinite);
      break;
    }
    if (st.stop == 1) {
      result = CUHRE_OK;
      break;
    }
    if (st.stop == 3 && opt.noalloc && budget_active >= st.tot.nactive) {
      // the evaluation budget allowed the split; the static lists did not
      result = fail(CUHRE_ERR_CAPACITY,
                    "the run needs %lld regions, more than the %lld the static workspace holds at dim %d "
                    "(cuhre_integrate has no such limit)",
                    (long long)(2 * st.tot.nactive), (long long)capacity, dim);
      out->fail = 1;
      break;
    }
    if (st.stop == 2 || st.stop == 3) {
      result = CUHRE_NOT_CONVERGED;
      break;
    }
    bound = 2 * n;
    n = 2 * st.tot.nactive;
    cur = 1 - cur;
    if (opt.handoff_n > 0 && (n >= opt.handoff_n || it + 1 == opt.max_iter)) {
      ho->handed = 1;
      ho->cur = cur;
      ho->n = n;
      ho->fin_s = fin_s;
      ho->fin_c = fin_c;
      ho->fer_s = fer_s;
      ho->fer_c = fer_c;
      ho->fin_abs = fin_abs;
      ho->fin_aest = fin_aest;
      ho->nfin = nfin;
      break;
    }
  }
  // Queued ahead, the result is complete once the last state is read: what
  // the stream still holds writes only the context's own buffers (the tail
  // of the last split, an iteration that read length 0), and every later
  // use of the context is ordered behind it on its stream (destroy
  // synchronises).  The static workspace is shared by every context on the
  // device, so the no-allocation form drains its stream before returning.
  if (!ahead || opt.noalloc) HIP_TRY(hipStreamSynchronize(c->stream));
  if (result >= 0 || !opt.noalloc) out->fail = result > 0 ? result : 0;
  out->ms_total = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  return result;
}

}  // namespace

// Launch order of Phase 2.  A region's run ends when its own error meets
// its tolerance; its length grows with the ratio of its first error to that
// tolerance, so the regions are dispatched by decreasing log2 of the ratio
// (kP2Buckets buckets, counting sort): the longest chains start first and
// the short ones fill the chip around them instead of a long chain starting
// last.  Results are written per region, so the order changes only when
// each one runs.
constexpr int kP2Buckets = 64;

__global__ void k_p2_zero(uint32_t* p, int count)
{
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < count) p[t] = 0u;
}

__global__ __launch_bounds__(256) void k_p2_bucket(const double* est, const double* err, int64_t n, double epsrel,
                                                   double epsabs, uint8_t* bucket, uint32_t* hist)
{
  __shared__ uint32_t h[kP2Buckets];
  if (threadIdx.x < kP2Buckets) h[threadIdx.x] = 0u;
  __syncthreads();
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r < n) {
    // ratio <= 1 or NaN: bucket 0; above 2^62 (a zero tolerance): the last
    const double q = err[r] / fmax(epsrel * fabs(est[r]), epsabs);
    const int b = q > 1.0 ? (q <= 0x1p62 ? ilogb(q) + 1 : kP2Buckets - 1) : 0;
    bucket[r] = (uint8_t)b;
    atomicAdd(&h[b], 1u);
  }
  __syncthreads();
  if (threadIdx.x < kP2Buckets && h[threadIdx.x] != 0u) atomicAdd(&hist[threadIdx.x], h[threadIdx.x]);
}

// Each block ranks its regions within their buckets in LDS and reserves
// one range per bucket in the global order (one global atomic per block
// and bucket: the regions crowd into a few buckets).
__global__ __launch_bounds__(256) void k_p2_scatter(const uint8_t* bucket, int64_t n, const uint32_t* hist,
                                                    uint32_t* cursor, int32_t* order)
{
  __shared__ uint32_t base[kP2Buckets], cnt[kP2Buckets];
  if (threadIdx.x < kP2Buckets) cnt[threadIdx.x] = 0u;
  if (threadIdx.x == 0) {
    uint32_t s = 0;
    for (int b = kP2Buckets - 1; b >= 0; --b) {
      base[b] = s;
      s += hist[b];
    }
  }
  __syncthreads();
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int b = -1;
  uint32_t rank = 0;
  if (r < n) {
    b = bucket[r];
    rank = atomicAdd(&cnt[b], 1u);
  }
  __syncthreads();
  if (threadIdx.x < kP2Buckets && cnt[threadIdx.x] != 0u)
    base[threadIdx.x] += atomicAdd(&cursor[threadIdx.x], cnt[threadIdx.x]);
  __syncthreads();
  if (b >= 0) order[base[b] + rank] = (int32_t)r;
}

// Phase 2 of the reference's two-phase method on the n regions of list
// `cur` (SoA, ld = the context's capacity): the regions' first evaluation by
// the half-wavefront kernel, their launch order, then one wavefront per
// region (k_phase2), per-region outputs copied to the host.  The workspace
// is the context's and kept between calls.
struct Phase2Host {
  std::vector<char> buf;  // the per-region outputs, copied as one span
  const double *res, *err, *rab, *rres;
  const int32_t* cnt;
  const int8_t *act, *bad, *zero;
};

static int run_phase2(cuhre_context* c, const cuhre_integrand* f, int dim, int key, int cur, int64_t n,
                      double epsrel, double epsabs, int32_t pool_max, Phase2Host& H)
{
  int rc;
  const Phase2Launcher launch = find_phase2_launcher(f->family, dim, key);
  if (!launch) return fail(CUHRE_ERR_UNSUPPORTED, "no Phase-2 kernel for family %d, dim %d", f->family, dim);
  DevIntegrand P;
  if ((rc = make_dev_integrand(c, f, dim, &P))) return rc;
  const RuleConst R = make_rule(dim, key);
  const int64_t slots = (int64_t)pool_max + 2;
  const size_t per_slot = sizeof(double) * (2 * (size_t)dim + 2) + sizeof(int8_t) + sizeof(int32_t);
  const size_t ws = (size_t)n * (size_t)slots * per_slot;
  // outputs, then the first evaluations (est, err, abssum, axis), the
  // launch order with its buckets and the bucket counters
  const size_t outb = (size_t)n * (4 * sizeof(double) + sizeof(int32_t) + 3) +
                      (size_t)n * (3 * sizeof(double) + sizeof(int8_t) + sizeof(int32_t) + 1) +
                      2 * kP2Buckets * sizeof(uint32_t) + 16 * 256;
  const size_t need = ws + outb + 8 * 256;
  if (c->p2_bytes < need) {
    size_t free_b = 0, total_b = 0;
    if (c->p2_ws) {
      (void)hipFree(c->p2_ws);
      c->p2_ws = nullptr;
      c->p2_bytes = 0;
    }
    HIP_TRY(hipMemGetInfo(&free_b, &total_b));
    if (need > free_b / 4 * 3)
      return fail(CUHRE_ERR_CAPACITY, "Phase 2 needs %.1f GB for %lld regions x %d slots", need / 1e9, (long long)n,
                  pool_max);
    if (hipMalloc(&c->p2_ws, need) != hipSuccess) {
      (void)hipGetLastError();
      c->p2_ws = nullptr;
      return fail(CUHRE_ERR_ALLOC, "cannot allocate the Phase-2 workspace (%.1f GB)", need / 1e9);
    }
    c->p2_bytes = need;
  }
  Phase2Args A;
  A.lower = c->lo[cur];
  A.width = c->wd[cur];
  A.ld = c->cap;
  A.n = n;
  A.slots = slots;
  char* q = c->p2_ws;
  auto take = [&](size_t bytes) {
    char* r = q;
    q += (bytes + 255) & ~(size_t)255;
    return r;
  };
  A.geo = (double*)take(sizeof(double) * (size_t)n * slots * 2 * dim);
  A.sest = (double*)take(sizeof(double) * (size_t)n * slots);
  A.sax = (int8_t*)take((size_t)n * slots);
  A.gkerr = (double*)take(sizeof(double) * (size_t)n * slots);
  A.gkseq = (int32_t*)take(sizeof(int32_t) * (size_t)n * slots);
  A.res = (double*)take(sizeof(double) * n);
  A.err = (double*)take(sizeof(double) * n);
  A.cnt = (int32_t*)take(sizeof(int32_t) * n);
  A.act = (int8_t*)take(n);
  A.nonfinite = (int8_t*)take(n);
  A.zeroed = (int8_t*)take(n);
  A.rab = (double*)take(sizeof(double) * n);
  A.rres = (double*)take(sizeof(double) * n);
  A.epsrel = epsrel;
  A.epsabs = epsabs;
  A.pool_max = pool_max;
  A.order = nullptr;
  A.est0 = A.err0 = A.ab0 = nullptr;
  A.ax0 = nullptr;
  const EvalLauncher first = find_eval_launcher(f->family, dim, key);
  if (first != nullptr) {
    double* est0 = (double*)take(sizeof(double) * n);
    double* err0 = (double*)take(sizeof(double) * n);
    double* ab0 = (double*)take(sizeof(double) * n);
    int8_t* ax0 = (int8_t*)take(n);
    int32_t* order = (int32_t*)take(sizeof(int32_t) * n);
    uint8_t* bucket = (uint8_t*)take(n);
    uint32_t* hist = (uint32_t*)take(2 * kP2Buckets * sizeof(uint32_t));
    uint32_t* cursor = hist + kP2Buckets;
    EvalArgs E;
    std::memset(&E, 0, sizeof E);
    E.lower = A.lower;
    E.width = A.width;
    E.ld = A.ld;
    E.n = n;
    E.epsrel = E.loc_epsrel = epsrel;
    E.epsabs = E.loc_epsabs = epsabs;
    E.est = est0;
    E.err = err0;
    E.axis = ax0;
    E.abssum = ab0;
    E.finalize = 0;  // per-region outputs only
    first(E, P, R, kHalfS, c->stream);
    const unsigned nb = (unsigned)((n + 255) / 256);
    hipLaunchKernelGGL(k_p2_zero, dim3(1), dim3(2 * kP2Buckets), 0, c->stream, hist, 2 * kP2Buckets);
    hipLaunchKernelGGL(k_p2_bucket, dim3(nb), dim3(256), 0, c->stream, (const double*)est0, (const double*)err0, n,
                       epsrel, epsabs, bucket, hist);
    hipLaunchKernelGGL(k_p2_scatter, dim3(nb), dim3(256), 0, c->stream, (const uint8_t*)bucket, n,
                       (const uint32_t*)hist, cursor, order);
    A.order = order;
    A.est0 = est0;
    A.err0 = err0;
    A.ab0 = ab0;
    A.ax0 = ax0;
  }
  // the outputs res .. rres lie in this order in the workspace: one copy
  const char* const d0 = reinterpret_cast<const char*>(A.res);
  const size_t span = (size_t)(reinterpret_cast<const char*>(A.rres + n) - d0);
  H.buf.resize(span);
  const auto at = [&](const void* d) { return H.buf.data() + (reinterpret_cast<const char*>(d) - d0); };
  H.res = reinterpret_cast<const double*>(at(A.res));
  H.err = reinterpret_cast<const double*>(at(A.err));
  H.rab = reinterpret_cast<const double*>(at(A.rab));
  H.rres = reinterpret_cast<const double*>(at(A.rres));
  H.cnt = reinterpret_cast<const int32_t*>(at(A.cnt));
  H.act = reinterpret_cast<const int8_t*>(at(A.act));
  H.bad = reinterpret_cast<const int8_t*>(at(A.nonfinite));
  H.zero = reinterpret_cast<const int8_t*>(at(A.zeroed));
  launch(A, P, R, c->stream);
  if (hipGetLastError() != hipSuccess ||
      hipMemcpyAsync(H.buf.data(), d0, span, hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
      hipStreamSynchronize(c->stream) != hipSuccess)
    return fail(CUHRE_ERR_DEVICE, "Phase 2 failed: %s", hipGetErrorString(hipGetLastError()));
  return 0;
}

// ------------------------------------------------------- static workspace
// Device memory of cuhre_integrate_noalloc: one array per GPU (each device
// holds its own copy at its own address until the process ends), 64 MiB, the
// last 1 MiB reserved for integrand tables.  In host code `scratch` names a
// host-side stand-in, not device memory: kernels and copies only ever get the
// address hipGetSymbolAddress returns for the current device, kept in
// g_scratch_ptr.
static __device__ __align__(256) char scratch[kScratchBytes];

constexpr int kScratchDevices = 64;
static std::mutex g_scratch_mu;
static char* g_scratch_ptr[kScratchDevices];
static std::mutex g_scratch_call_mu[kScratchDevices];  // one no-allocation call per device at a time

// With `device` current: takes (once) that device's address of the array.
static int scratch_setup(int device)
{
  if (device < 0 || device >= kScratchDevices)
    return fail(CUHRE_ERR_UNSUPPORTED, "static workspace: device %d beyond %d", device, kScratchDevices);
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch_ptr[device]) return 0;
  void* ptr = nullptr;
  const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  if (e != hipSuccess || ptr == nullptr) {
    (void)hipGetLastError();
    return fail(CUHRE_ERR_DEVICE, "hipGetSymbolAddress(scratch) on device %d failed: %s", device,
                hipGetErrorString(e));
  }
  g_scratch_ptr[device] = static_cast<char*>(ptr);
  return 0;
}

extern "C" {

int cuhre_integrate(cuhre_context* c, const cuhre_integrand* f, int dim, double epsrel, double epsabs,
                    int64_t maxeval, int key, const double* lows, const double* highs, cuhre_result* out)
{
  LoopOpts opt;
  return plain_loop(c, f, dim, epsrel, epsabs, maxeval, key, lows, highs, out, opt, nullptr);
}

int cuhre_integrate_noalloc(cuhre_context* c, const cuhre_integrand* f, int dim, double epsrel, double epsabs,
                            int64_t maxeval, int key, const double* lows, const double* highs, cuhre_result* out)
{
  if (!c) return fail(CUHRE_ERR_ARG, "null context");
  if (!out) return fail(CUHRE_ERR_ARG, "null result");
  if (dim < 2 || dim > 10) return fail(CUHRE_ERR_UNSUPPORTED, "dimension %d outside 2..10", dim);
  HIP_TRY(hipSetDevice(c->device));
  int rc = scratch_setup(c->device);
  if (rc) return rc;
  std::lock_guard<std::mutex> call_lk(g_scratch_call_mu[c->device]);
  char* const base = g_scratch_ptr[c->device];
  const int64_t cap = scratch_capacity(dim);

  // the context's own list buffers step aside for the call
  double *lo0 = c->lo[0], *lo1 = c->lo[1], *wd0 = c->wd[0], *wd1 = c->wd[1];
  double *est = c->est, *err = c->err, *ab = c->abssum;
  int8_t* ax = c->axis;
  Partial* part = c->partials;
  int64_t* off = c->offsets;
  const int64_t cap0 = c->cap, capacity0 = c->capacity;
  const int dim0 = c->dim_alloc;

  char* q = base;
  auto take = [&](size_t bytes) {
    char* r = q;
    q += al256(bytes);
    return r;
  };
  const int64_t nb = block_capacity(cap), ng = group_capacity(nb);
  c->lo[0] = (double*)take(sizeof(double) * (size_t)cap * dim);
  c->lo[1] = (double*)take(sizeof(double) * (size_t)cap * dim);
  c->wd[0] = (double*)take(sizeof(double) * (size_t)cap * dim);
  c->wd[1] = (double*)take(sizeof(double) * (size_t)cap * dim);
  c->est = (double*)take(sizeof(double) * (size_t)cap);
  c->err = (double*)take(sizeof(double) * (size_t)cap);
  c->abssum = (double*)take(sizeof(double) * (size_t)cap);
  c->axis = (int8_t*)take((size_t)cap);
  c->partials = (Partial*)take(sizeof(Partial) * (size_t)(nb + ng));
  c->offsets = (int64_t*)take(sizeof(int64_t) * (size_t)(nb + ng));
  c->cap = cap;
  c->dim_alloc = dim;
  c->capacity = cap;

  LoopOpts opt;
  opt.noalloc = true;
  opt.table_dst = base + (kScratchBytes - kScratchTableBytes);
  opt.table_bytes = kScratchTableBytes;
  rc = plain_loop(c, f, dim, epsrel, epsabs, maxeval, key, lows, highs, out, opt, nullptr);
  // nothing may still use the workspace once the lock is released
  const hipError_t se = hipStreamSynchronize(c->stream);

  c->lo[0] = lo0;
  c->lo[1] = lo1;
  c->wd[0] = wd0;
  c->wd[1] = wd1;
  c->est = est;
  c->err = err;
  c->abssum = ab;
  c->axis = ax;
  c->partials = part;
  c->offsets = off;
  c->cap = cap0;
  c->capacity = capacity0;
  c->dim_alloc = dim0;
  if (rc == 0 && se != hipSuccess) return fail(CUHRE_ERR_DEVICE, "HIP error: %s", hipGetErrorString(se));
  return rc;
}

int cuhre_integrate_two_phase(cuhre_context* c, const cuhre_integrand* f, int dim, double epsrel, double epsabs,
                              int key, int64_t phase1_regions, int32_t pool_max, const double* lows,
                              const double* highs, cuhre_result* out)
{
  const auto t0 = std::chrono::steady_clock::now();
  if (!out) return fail(CUHRE_ERR_ARG, "null result");
  if (phase1_regions <= 0) phase1_regions = 1 << 14;  // the reference's FIRST_PHASE_MAXREGIONS
  if (pool_max <= 0) pool_max = 2048;                 // the reference's MAX_GLOBALPOOL_SIZE
  if (!c) return fail(CUHRE_ERR_ARG, "null context");
  if (pool_max > (1 << 24)) return fail(CUHRE_ERR_ARG, "pool_max %d above 2^24", pool_max);
  LoopOpts opt;
  opt.handoff_n = phase1_regions;
  opt.max_iter = 100;  // GPUKernelquad.cu:631
  LoopHandoff ho;
  const int64_t no_budget = INT64_MAX / 4;
  int rc = plain_loop(c, f, dim, epsrel, epsabs, no_budget, key, lows, highs, out, opt, &ho);
  if (rc < 0 || !ho.handed) {
    // converged in Phase 1 (or nothing left active): the loop's result
    out->ms_total = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return rc;
  }
  // Phase 2 on the ho.n regions of list ho.cur (split, not yet evaluated)
  const int64_t p = make_rule(dim, key).npts;
  const int64_t n = ho.n;
  Phase2Host H;
  if ((rc = run_phase2(c, f, dim, key, ho.cur, n, epsrel, epsabs, pool_max, H))) return rc;
  const double *hres = H.res, *herr = H.err, *hrab = H.rab, *hrres = H.rres;
  const int32_t* hcnt = H.cnt;
  const int8_t *hact = H.act, *hbad = H.bad, *hzero = H.zero;
  int result = 0;
  double I_s = ho.fin_s, I_c = ho.fin_c, E_s = ho.fer_s, E_c = ho.fer_c;
  double ab_tot = ho.fin_abs, res_round = 0.0, aest = ho.fin_aest;
  int64_t failed = 0, bad = 0, zeroed = 0;
  for (int64_t r = 0; r < n; ++r) {
    dd_add_host(I_s, I_c, hres[r], 0.0);
    dd_add_host(E_s, E_c, herr[r], 0.0);
    out->nregions += hcnt[r];
    out->neval += (int64_t)(hcnt[r] - 1) * 2 * p + p;
    failed += hact[r];
    bad += hbad[r];
    zeroed += hzero[r];
    ab_tot += hrab[r];
    res_round += hrres[r];
    aest += std::fabs(hres[r]);
  }
  if (bad) return fail(CUHRE_ERR_NONFINITE, "integrand returned a non-finite value in Phase 2 (%lld regions)",
                       (long long)bad);
  out->integral = I_s + I_c;
  out->error = E_s + E_c;
  out->nfinished = ho.nfin + n;
  out->nzeroed = (int32_t)std::min<int64_t>(zeroed, INT32_MAX);
  // rounding: every evaluated region's rule sum (an over-count: bisected
  // members left the sum), each Phase-2 region's RES update chain, and the
  // sum over the finished Phase-1 regions and the n Phase-2 results
  {
    const double u = 0x1p-53;
    // the budgets are themselves FP64 sums of at most pool_max + 1 and n terms
    const double fac = 1.0 + 4.0 * ((double)pool_max + (double)n + 2.0) * u;
    out->fp64_err_bound = fp64_rounding_bound(p, dim, ho.nfin + n, ab_tot, aest, out->integral) +
                          u * res_round * fac;
  }
  const bool ok = out->error <= std::max(epsrel * std::fabs(out->integral), epsabs);
  // a zeroed region's contribution is dropped (the reference's rule,
  // GPUQuadPhases.cu:519-531): the result is then not certified
  result = ok && zeroed == 0 ? CUHRE_OK : CUHRE_NOT_CONVERGED;
  // the reference's flag: Phase-2 regions that missed their test, 0 when
  // the global test holds (GPUquad.h:713-714)
  out->fail = ok && zeroed == 0 ? 0 : (int32_t)std::min<int64_t>(std::max<int64_t>(failed, 1), INT32_MAX);
  out->ms_total = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  return result;
}

int cuhre_phase2_regions(cuhre_context* c, const cuhre_integrand* f, int dim, int key, int64_t n,
                         const double* lower, const double* width, double epsrel, double epsabs, int32_t pool_max,
                         double* res, double* err, int32_t* count)
{
  int rc = check_common(c, f, dim, key);
  if (rc) return rc;
  if (n <= 0) return 0;
  if (!lower || !width || !res || !err || !count) return fail(CUHRE_ERR_ARG, "null buffer");
  if (pool_max <= 0) pool_max = 2048;
  if (pool_max > (1 << 24)) return fail(CUHRE_ERR_ARG, "pool_max %d above 2^24", pool_max);
  if ((rc = ctx_reserve(c, n, dim))) return rc;
  HIP_TRY(hipMemcpy2DAsync(c->lo[0], sizeof(double) * c->cap, lower, sizeof(double) * n, sizeof(double) * n, dim,
                           hipMemcpyHostToDevice, c->stream));
  HIP_TRY(hipMemcpy2DAsync(c->wd[0], sizeof(double) * c->cap, width, sizeof(double) * n, sizeof(double) * n, dim,
                           hipMemcpyHostToDevice, c->stream));
  Phase2Host H;
  if ((rc = run_phase2(c, f, dim, key == 0 ? 7 : key, 0, n, epsrel, epsabs, pool_max, H))) return rc;
  int64_t bad = 0;
  for (int64_t r = 0; r < n; ++r) {
    res[r] = H.res[r];
    err[r] = H.err[r];
    count[r] = H.cnt[r];
    bad += H.bad[r];
  }
  if (bad) return fail(CUHRE_ERR_NONFINITE, "integrand returned a non-finite value in Phase 2 (%lld regions)",
                       (long long)bad);
  return 0;
}

int cuhre_eval_regions(cuhre_context* c, const cuhre_integrand* f, int dim, int key, int64_t n,
                       const double* lower, const double* width, double* est, double* err,
                       int32_t* axis, double* abssum, double* i5, double* dvec)
{
  int rc = check_common(c, f, dim, key);
  if (rc) return rc;
  if (n <= 0) return 0;
  if (!lower || !width || !est || !err || !axis || !abssum) return fail(CUHRE_ERR_ARG, "null buffer");
  DevIntegrand P;
  if ((rc = make_dev_integrand(c, f, dim, &P))) return rc;
  const RuleConst R = make_rule(dim, key);
  if ((rc = ctx_reserve(c, n, dim))) return rc;
  const bool want_aux = i5 != nullptr || dvec != nullptr;
  if (want_aux && (rc = ctx_reserve_aux(c, c->cap, dim))) return rc;
  HIP_TRY(hipMemcpy2DAsync(c->lo[0], sizeof(double) * c->cap, lower, sizeof(double) * n,
                           sizeof(double) * n, dim, hipMemcpyHostToDevice, c->stream));
  HIP_TRY(hipMemcpy2DAsync(c->wd[0], sizeof(double) * c->cap, width, sizeof(double) * n,
                           sizeof(double) * n, dim, hipMemcpyHostToDevice, c->stream));
  const int S = pick_segments(c, n, key);
  EvalArgs A;
  fill_eval_args(c, A, 0, n, dim, 0.0, 0.0);
  A.finalize = 0;
  if (want_aux) {
    A.i5 = c->aux;
    A.dvec = c->aux + c->cap;
  }
  find_eval_launcher(f->family, dim, key)(A, P, R, S, c->stream);
  HIP_TRY(hipGetLastError());
  std::vector<int8_t> ax8((size_t)n);
  HIP_TRY(hipMemcpyAsync(est, c->est, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipMemcpyAsync(err, c->err, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipMemcpyAsync(abssum, c->abssum, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipMemcpyAsync(ax8.data(), c->axis, n, hipMemcpyDeviceToHost, c->stream));
  if (i5) HIP_TRY(hipMemcpyAsync(i5, c->aux, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
  if (dvec)
    HIP_TRY(hipMemcpy2DAsync(dvec, sizeof(double) * n, c->aux + c->cap, sizeof(double) * c->cap,
                             sizeof(double) * n, dim, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  for (int64_t r = 0; r < n; ++r) axis[r] = ax8[r];
  return 0;
}

int cuhre_split_regions(cuhre_context* c, int dim, int64_t n, const double* lower, const double* width,
                        const double* est, const double* err, const int32_t* axis, double epsrel,
                        double epsabs, double* out_lower, double* out_width, int64_t* n_active)
{
  if (!c) return fail(CUHRE_ERR_ARG, "null context");
  if (dim < 1 || dim > kMaxDim) return fail(CUHRE_ERR_ARG, "dimension %d out of range", dim);
  if (!n_active) return fail(CUHRE_ERR_ARG, "null n_active");
  *n_active = 0;
  if (n <= 0) return 0;
  if (!lower || !width || !est || !err || !axis || !out_lower || !out_width)
    return fail(CUHRE_ERR_ARG, "null buffer");
  for (int64_t r = 0; r < n; ++r)
    if (axis[r] < 0 || axis[r] >= dim) return fail(CUHRE_ERR_ARG, "axis[%lld] out of range", (long long)r);
  HIP_TRY(hipSetDevice(c->device));
  int rc;
  if ((rc = ctx_reserve(c, 2 * n, dim))) return rc;
  std::vector<int8_t> ax8((size_t)n);
  for (int64_t r = 0; r < n; ++r) ax8[r] = (int8_t)axis[r];
  HIP_TRY(hipMemcpy2DAsync(c->lo[0], sizeof(double) * c->cap, lower, sizeof(double) * n,
                           sizeof(double) * n, dim, hipMemcpyHostToDevice, c->stream));
  HIP_TRY(hipMemcpy2DAsync(c->wd[0], sizeof(double) * c->cap, width, sizeof(double) * n,
                           sizeof(double) * n, dim, hipMemcpyHostToDevice, c->stream));
  HIP_TRY(hipMemcpyAsync(c->est, est, sizeof(double) * n, hipMemcpyHostToDevice, c->stream));
  HIP_TRY(hipMemcpyAsync(c->err, err, sizeof(double) * n, hipMemcpyHostToDevice, c->stream));
  HIP_TRY(hipMemcpyAsync(c->axis, ax8.data(), n, hipMemcpyHostToDevice, c->stream));
  const int64_t nb = (n + kBlock - 1) / kBlock;
  hipLaunchKernelGGL(k_count, dim3((unsigned)nb), dim3(kBlock), 0, c->stream, c->est, c->err, n, epsrel,
                     epsabs, c->offsets);
  HIP_TRY(hipGetLastError());
  std::vector<int64_t> counts((size_t)nb);
  HIP_TRY(hipMemcpyAsync(counts.data(), c->offsets, sizeof(int64_t) * nb, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  int64_t run = 0;
  for (int64_t b = 0; b < nb; ++b) {
    const int64_t t = counts[b];
    counts[b] = run;
    run += t;
  }
  HIP_TRY(hipMemcpyAsync(c->offsets, counts.data(), sizeof(int64_t) * nb, hipMemcpyHostToDevice, c->stream));
  SplitArgs S;
  fill_split_args(c, S, 0, n, dim, epsrel, epsabs, kBlock, false);
  hipLaunchKernelGGL(k_split, dim3((unsigned)nb), dim3(kBlock), 0, c->stream, S);
  HIP_TRY(hipGetLastError());
  if (run > 0) {
    HIP_TRY(hipMemcpy2DAsync(out_lower, sizeof(double) * 2 * run, c->lo[1], sizeof(double) * c->cap,
                             sizeof(double) * 2 * run, dim, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipMemcpy2DAsync(out_width, sizeof(double) * 2 * run, c->wd[1], sizeof(double) * c->cap,
                             sizeof(double) * 2 * run, dim, hipMemcpyDeviceToHost, c->stream));
  }
  HIP_TRY(hipStreamSynchronize(c->stream));
  *n_active = run;
  return 0;
}

int cuhre_sum(cuhre_context* c, int64_t n, const double* x, double* out)
{
  if (!c || !out) return fail(CUHRE_ERR_ARG, "null argument");
  *out = 0.0;
  if (n <= 0) return 0;
  if (!x) return fail(CUHRE_ERR_ARG, "null input");
  HIP_TRY(hipSetDevice(c->device));
  int rc;
  if ((rc = ctx_reserve(c, n, 1))) return rc;
  HIP_TRY(hipMemcpyAsync(c->est, x, sizeof(double) * n, hipMemcpyHostToDevice, c->stream));
  const int64_t per = 4096;
  const int64_t nb = (n + per - 1) / per;
  if (c->sum_part) (void)hipFree(c->sum_part);
  c->sum_part = nullptr;
  HIP_TRY(hipMalloc(&c->sum_part, sizeof(double2) * nb));
  hipLaunchKernelGGL(k_sum_blocks, dim3((unsigned)nb), dim3(kBlock), 0, c->stream, c->est, n, per, c->sum_part);
  hipLaunchKernelGGL(k_sum_final, dim3(1), dim3(kBlock), 0, c->stream, c->sum_part, nb, c->sum_out);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipMemcpyAsync(out, c->sum_out, sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return 0;
}

}  // extern "C"
