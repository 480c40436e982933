// batch.hip -- many independent integrals in one region list.
//
// Batched form of the plain method (no reference counterpart; the reference
// integrates one functor per GPUcuhre::integrate call, quad/GPUquad/
// GPUquad.h:611-727).  Every instance runs exactly the plain loop -- same
// rule, same per-region classification, same global test on the instance's
// own sums, same evaluation budget -- but all instances share one SoA region
// list in HBM, tagged with an instance id and kept grouped by instance (the
// compaction is stable and children stay adjacent).  One iteration is six
// launches with fixed shapes that read the list length from device memory:
//
//   k_batch_eval    rule evaluation, grid-stride over 64-region groups, the
//                   lane's integrand parameters fetched by instance id;
//   k_chunk_reduce  one block per 8192-region chunk of an instance's slice:
//                   TwoSum-carried sums and active count of the chunk;
//   k_inst_combine  one block per instance: its chunks in order, global
//                   test, budget test, children count;
//   k_batch_scan    one block: next slice offsets of the instances (with a
//                   capacity cut) and the next list length;
//   k_tile_count /  per-256-region active counts, scanned within groups
//   k_group_scan    of 64 tiles, then the group totals;
//   k_batch_split   stable compaction + bisection into the other buffer.
//
// Because no launch depends on a host-side size, two iterations (one per
// ping-pong parity) are captured once in a hipGraph and replayed until the
// list is empty; finished instances cost nothing.  The evaluation launch
// shape (point segments, grid blocks) comes from cuhre_batch_autotune, which
// times the candidates on the batch's own integrands.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <vector>

#include "common.hpp"
#include "eval_kernels.hpp"
#include "reduce.hpp"

namespace cuhre {

RuleConst make_rule(int d, int key);
int gm_points(int d);
int set_error(int code, const char* msg);

// Regions per reduction chunk: an instance's slice is reduced by one block
// per chunk, then its chunks are combined in order by one block.
constexpr int64_t kChunk = 8192;

// Per-instance parameters and tolerances (uploaded once per call).
struct BatchParams {
  int32_t family, exponent;
  double a[kMaxDim], w[kMaxDim], b[kMaxDim];
  double s0;
  double epsrel, epsabs;
  int64_t maxeval;
  double lows[kMaxDim], highs[kMaxDim];
  double allot;  // shared-allotment share b (0: the plain per-region test)
};

// Per-instance loop state, updated on the device.
struct InstState {
  double fin_s, fin_c, fer_s, fer_c;  // finished-region sums (value, carry)
  double abs_fin, aest_fin;
  double integral, error, bound;
  int64_t neval, nregions, nfin, peak;
  int64_t start, count;  // slice of the current list
  int64_t next;          // children for the next iteration (2 * active)
  int32_t done, fail, iters, pad;
  double absI;           // |estimate| of the current iteration (allotment)
};

// Integrand of one lane in a batch: family FAM with the lane's instance
// parameters (held in registers).
template <int D, int FAM>
struct LaneIntegrand {
  static constexpr bool kUnroll = Family<FAM, D>::kUnroll;
  DevIntegrand P;
  __device__ __forceinline__ double operator()(const double (&x)[D]) const
  {
    return Family<FAM, D>::eval(P, x);
  }
};

struct BatchBuffers {
  double* lower;
  double* width;
  int32_t* inst;
  double* out_lower;
  double* out_width;
  int32_t* out_inst;
  int64_t ld;
  const int64_t* n;  // current list length
  int64_t* n_next;
  unsigned long long* n_peak;  // largest list length so far
  int64_t* flags;  // [0] overflow: parity + 1 of the iteration that ran out of room, [1] its need
  int32_t parity;
  int32_t allow_cut;  // at the memory limit: stop instances that do not fit
  double* est;
  double* err;
  double* abssum;
  double* vrel;      // region volume / instance box volume (allotment only)
  int8_t* axis;
  Partial* chunk_part;  // per chunk of an instance's slice
  double2* chunk_est;   // per chunk: estimate sum (the allotment's first pass)
  int64_t* chunk0;      // per instance: its first chunk; [count]: total chunks
  int64_t* nch;         // total chunks of the current list
  int64_t* tile_off;  // per 256-region tile: exclusive offset within its group of 64 tiles
  int64_t* grp;       // per group of 64 tiles: [0, ngrp) active totals, [gcap, gcap + ngrp) bases
  int64_t gcap;
  const BatchParams* params;
  InstState* state;
  int32_t count;
  int32_t dim;
  int32_t npts;
  int64_t cap;
};

// Per-region test of an instance: the reference's err > max(epsrel |est|,
// epsabs) (GPUQuadPhases.cu:63-73), or with the shared allotment b > 0
// err > max(epsrel ((1 - b) |est| + b |I| vrel), epsabs), I the instance's
// estimate in this iteration and vrel the region's share of its box.
__device__ __forceinline__ bool inst_region_active(const BatchParams& bp, double absI, double e, double er,
                                                   double vrel)
{
  if (!(isfinite(e) && isfinite(er))) return false;
  if (bp.allot > 0.0) {
    const double a = bp.epsrel * ((1.0 - bp.allot) * fabs(e) + bp.allot * absI * vrel);
    return er > fmax(a, bp.epsabs);
  }
  return region_active(e, er, bp.epsrel, bp.epsabs);
}

__device__ __forceinline__ bool batch_active(const BatchBuffers& B, const InstState& st, const BatchParams& bp,
                                             int64_t r)
{
  if (st.done) return false;
  return inst_region_active(bp, st.absI, B.est[r], B.err[r], bp.allot > 0.0 ? B.vrel[r] : 0.0);
}

// Initial k^d grid of every instance plus the instance states.
__global__ void k_batch_init(BatchBuffers B, int k, int64_t n0)
{
#pragma clang fp contract(off)  // low + id * w rounded twice, as the oracle computes it
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t total = n0 * B.count;
  const int64_t cpi = (n0 + kChunk - 1) / kChunk;  // chunks per instance
  if (r < B.count) {
    InstState st;
    memset(&st, 0, sizeof st);
    st.start = r * n0;
    st.count = n0;
    B.state[r] = st;
    B.chunk0[r] = r * cpi;
  }
  if (r == 0) {
    B.chunk0[B.count] = B.count * cpi;
    *B.nch = B.count * cpi;
  }
  if (r >= total) return;
  const int i = (int)(r / n0);
  int64_t q = r - (int64_t)i * n0;
  const BatchParams& bp = B.params[i];
  for (int m = 0; m < B.dim; ++m) {
    const int64_t id = q % k;
    q /= k;
    const double w = (bp.highs[m] - bp.lows[m]) / (double)k;
    B.lower[m * B.ld + r] = bp.lows[m] + (double)id * w;
    B.width[m * B.ld + r] = w;
  }
  B.inst[r] = i;
}

template <int D, int FAM, bool ONE>
__device__ __forceinline__ RegionOut batch_seg_eval(const BatchBuffers& B, const RuleConst& R,
                                                    int64_t r, bool valid, int64_t rr, int S, int s,
                                                    int wave, int lane, int W, double* seg_lds)
{
  const BatchParams& bp = B.params[B.inst[rr]];
  LaneIntegrand<D, FAM> f;
  f.P.family = FAM;
  f.P.exponent = bp.exponent;
#pragma unroll
  for (int m = 0; m < D; ++m) {
    f.P.a[m] = bp.a[m];
    f.P.w[m] = bp.w[m];
    f.P.b[m] = bp.b[m];
  }
  f.P.s0 = bp.s0;
  return seg_eval<D, false, ONE>(f, B.lower, B.width, B.ld, r, valid, rr, S, s, wave, lane, W, seg_lds, R, nullptr);
}

// One launch per family present in the batch.  Instances are ordered by
// family, so the families along the list are non-decreasing: a tile whose
// first and last regions bracket FAM may hold it, any other tile is skipped
// (a block-uniform test, which keeps the segment barriers convergent).
// Lanes of another family in a boundary tile compute but do not store.
// ONE: one segment (S = 1), the straight-line point loops of seg_eval.
template <int D, int FAM, bool ONE>
__global__ __launch_bounds__(kBlock) void k_batch_eval(BatchBuffers B, RuleConst R, int S)
{
  extern __shared__ double seg_lds[];
  const int W = S < 4 ? 4 : S;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int lane = threadIdx.x & (kWave - 1);
  const int g = wave / S, s = wave - g * S;
  const int64_t n = *B.n;
  const int64_t rpb = kWave * (W / S);
  const int64_t ntiles = (n + rpb - 1) / rpb;
  for (int64_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
    const int64_t t0 = t * rpb;
    const int64_t t1 = (t0 + rpb < n ? t0 + rpb : n) - 1;
    const int f_lo = B.params[B.inst[t0]].family;
    const int f_hi = B.params[B.inst[t1]].family;
    if (FAM < f_lo || FAM > f_hi) continue;
    const int64_t r = t0 + g * kWave + lane;
    const bool valid = r < n;
    const int64_t rr = valid ? r : n - 1;
    const RegionOut o = batch_seg_eval<D, FAM, ONE>(B, R, r, valid, rr, S, s, wave, lane, W, seg_lds);
    if (valid && s == 0) {
      const BatchParams& bp = B.params[B.inst[r]];
      if (bp.family == FAM) {
        B.est[r] = o.est;
        B.err[r] = o.err;
        B.abssum[r] = o.ab;
        B.axis[r] = (int8_t)o.axis;
        if (bp.allot > 0.0) {
          double v = 1.0;
#pragma unroll
          for (int m = 0; m < D; ++m) v *= B.width[m * B.ld + r] / (bp.highs[m] - bp.lows[m]);
          B.vrel[r] = v;
        }
      }
    }
  }
}

// The instance whose slice holds chunk c: the last i with chunk0[i] <= c
// (instances without chunks share chunk0 with the next one that has some).
__device__ __forceinline__ int chunk_owner(const int64_t* chunk0, int count, int64_t c)
{
  int lo = 0, hi = count - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (chunk0[mid] <= c) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// One block per chunk (grid-stride): the chunk's Partial, or with
// EST_ONLY its estimate sum alone (the allotment's test needs the
// instance's estimate before it can classify).
template <bool EST_ONLY>
__global__ __launch_bounds__(kBlock) void k_chunk_reduce(BatchBuffers B)
{
  __shared__ Partial ldsw[kBlock / kWave];
  if (B.flags[0] != 0) return;
  const int64_t nch = *B.nch;
  for (int64_t c = blockIdx.x; c < nch; c += gridDim.x) {
    const int i = chunk_owner(B.chunk0, B.count, c);
    const InstState& st = B.state[i];
    const BatchParams& bp = B.params[i];
    const int64_t r0 = st.start + (c - B.chunk0[i]) * kChunk;
    const int64_t end = st.start + st.count;
    const int64_t r1 = r0 + kChunk < end ? r0 + kChunk : end;
    const double absI = st.absI;
    Partial q;
    partial_zero(q);
    for (int64_t r = r0 + threadIdx.x; r < r1; r += kBlock) {
      const double e = B.est[r];
      if constexpr (EST_ONLY) {
        dd_add(q.est_s, q.est_c, e, 0.0);
      } else {
        const double er = B.err[r];
        const bool act = inst_region_active(bp, absI, e, er, bp.allot > 0.0 ? B.vrel[r] : 0.0);
        Partial one;
        partial_from_region_act(one, e, er, B.abssum[r], act);
        partial_add(q, one);
      }
    }
    block_reduce_partial(q, ldsw, kBlock / kWave);
    if (threadIdx.x == 0) {
      if constexpr (EST_ONLY) B.chunk_est[c] = make_double2(q.est_s, q.est_c);
      else B.chunk_part[c] = q;
    }
    __syncthreads();  // ldsw is written again by the next chunk
  }
}

// One block per instance: |estimate| of the iteration (finished so far plus
// every region of the iteration) for the allotment's test.
__global__ __launch_bounds__(kBlock) void k_inst_absI(BatchBuffers B)
{
  __shared__ Partial ldsw[kBlock / kWave];
  const int i = blockIdx.x;
  if (B.flags[0] != 0) return;
  const InstState& st = B.state[i];
  if (st.done) return;
  Partial q;
  partial_zero(q);
  for (int64_t k = B.chunk0[i] + threadIdx.x; k < B.chunk0[i + 1]; k += kBlock)
    dd_add(q.est_s, q.est_c, B.chunk_est[k].x, B.chunk_est[k].y);
  block_reduce_partial(q, ldsw, kBlock / kWave);
  if (threadIdx.x == 0) {
    double s = st.fin_s, c = st.fin_c;
    dd_add(s, c, q.est_s, q.est_c);
    B.state[i].absI = fabs(s + c);
  }
}

// One block per instance: its chunks' Partials in chunk order, then the
// instance's loop step (global test, budget, children count).
__global__ __launch_bounds__(kBlock) void k_inst_combine(BatchBuffers B)
{
  __shared__ Partial ldsw[kBlock / kWave];
  const int i = blockIdx.x;
  if (B.flags[0] != 0) return;
  InstState st = B.state[i];
  if (st.done) return;
  const BatchParams& bp = B.params[i];
  Partial q;
  partial_zero(q);
  for (int64_t k = B.chunk0[i] + threadIdx.x; k < B.chunk0[i + 1]; k += kBlock) partial_add(q, B.chunk_part[k]);
  block_reduce_partial(q, ldsw, kBlock / kWave);
  if (threadIdx.x != 0) return;
  const int64_t p = B.npts;
  double rs = st.fin_s, rc = st.fin_c, es = st.fer_s, ec = st.fer_c;
  dd_add(rs, rc, q.est_s, q.est_c);
  dd_add(es, ec, q.err_s, q.err_c);
  const double rG = rs + rc, eG = es + ec;
  st.iters += 1;
  st.nregions += st.count;
  st.neval += st.count * p;
  st.peak = st.count > st.peak ? st.count : st.peak;
  st.integral = rG;
  st.error = eG;
  st.bound = fp64_rounding_bound(p, B.dim, st.nfin + st.count, st.abs_fin + q.abs_all, st.aest_fin + q.aest_all, rG);
  dd_add(st.fin_s, st.fin_c, q.fest_s, q.fest_c);
  dd_add(st.fer_s, st.fer_c, q.ferr_s, q.ferr_c);
  st.abs_fin += q.abs_fin;
  st.aest_fin += q.aest_fin;
  st.nfin += st.count - q.nactive;
  st.next = 0;
  if (q.nnonfinite > 0) {
    st.done = 1;
    st.fail = CUHRE_ERR_NONFINITE;
  } else if (eG <= fmax(bp.epsrel * fabs(rG), bp.epsabs)) {
    st.done = 1;
    st.fail = 0;
  } else if (q.nactive == 0 || st.neval + 2 * q.nactive * p > bp.maxeval || st.iters >= 1000) {
    st.done = 1;
    st.fail = 1;
  } else {
    st.next = 2 * q.nactive;
  }
  B.state[i] = st;
}

// Slice offsets of the next list (exclusive scan of the children counts in
// instance order).  If the next list does not fit, the iteration is halted
// here (overflow flag, empty next list) so that the host can grow the
// buffers and finish it; only at the memory limit (allow_cut) do the
// instances from the first one that does not fit onward stop with fail = 1.
__global__ __launch_bounds__(1024) void k_batch_scan(BatchBuffers B)
{
  __shared__ int64_t sc[1024];
  __shared__ int s_cut;
  __shared__ unsigned long long s_tot;
  if (B.flags[0] != 0) return;
  const int T = blockDim.x;
  const int chunk = (B.count + T - 1) / T;
  const int i0 = min((int)threadIdx.x * chunk, B.count);
  const int i1 = min(i0 + chunk, B.count);
  int64_t mine = 0;
  for (int i = i0; i < i1; ++i) mine += B.state[i].done ? 0 : B.state[i].next;
  sc[threadIdx.x] = mine;
  if (threadIdx.x == 0) {
    s_cut = B.count;
    s_tot = 0ull;
  }
  __syncthreads();
  for (int off = 1; off < T; off <<= 1) {
    const int64_t v = (int)threadIdx.x >= off ? sc[threadIdx.x - off] : 0;
    __syncthreads();
    sc[threadIdx.x] += v;
    __syncthreads();
  }
  const int64_t total = sc[T - 1];
  if (total > B.cap && !B.allow_cut) {
    if (threadIdx.x == 0) {
      B.flags[0] = B.parity + 1;
      B.flags[1] = total;
      *B.n_next = 0;
    }
    return;
  }
  const int64_t excl = sc[threadIdx.x] - mine;
  int64_t run = excl;
  for (int i = i0; i < i1; ++i) {
    const int64_t c = B.state[i].done ? 0 : B.state[i].next;
    if (run + c > B.cap) {
      atomicMin(&s_cut, i);
      break;
    }
    run += c;
  }
  __syncthreads();
  const int cut = s_cut;
  run = excl;
  unsigned long long kept = 0ull;
  for (int i = i0; i < i1; ++i) {
    InstState& st = B.state[i];
    const int64_t c = st.done ? 0 : st.next;
    if (!st.done && i >= cut) {
      st.done = 1;
      st.fail = 1;
      st.next = 0;
    }
    if (!st.done) {
      st.start = run;
      st.count = c;
      kept += (unsigned long long)c;
    } else {
      st.count = 0;
    }
    run += c;
  }
  atomicAdd(&s_tot, kept);
  // chunk map of the next list: the chunk counts scanned in instance order
  int64_t my_ch = 0;
  for (int i = i0; i < i1; ++i)
    if (!B.state[i].done) my_ch += (B.state[i].count + kChunk - 1) / kChunk;
  __syncthreads();
  sc[threadIdx.x] = my_ch;
  __syncthreads();
  for (int off = 1; off < T; off <<= 1) {
    const int64_t v = (int)threadIdx.x >= off ? sc[threadIdx.x - off] : 0;
    __syncthreads();
    sc[threadIdx.x] += v;
    __syncthreads();
  }
  int64_t cb = sc[threadIdx.x] - my_ch;
  for (int i = i0; i < i1; ++i) {
    B.chunk0[i] = cb;
    if (!B.state[i].done) cb += (B.state[i].count + kChunk - 1) / kChunk;
  }
  if (threadIdx.x == 0) {
    *B.n_next = (int64_t)s_tot;
    if (s_tot > *B.n_peak) *B.n_peak = s_tot;
    B.chunk0[B.count] = sc[T - 1];
    *B.nch = sc[T - 1];
  }
}

// Active counts of the 256-region tiles, a group of 64 tiles per block
// iteration: tile by tile (coalesced loads, a ballot per wave), then one
// wavefront scans the group's 64 counts into offsets within the group and
// the group's total.
__global__ __launch_bounds__(kBlock) void k_tile_count(BatchBuffers B)
{
  __shared__ uint32_t wcnt[kBlock / kWave];
  __shared__ int64_t gcnt[kWave];
  if (B.flags[0] != 0) return;
  const int64_t n = *B.n;
  const int64_t ntiles = (n + kBlock - 1) / kBlock;
  const int64_t ngrp = (ntiles + kWave - 1) / kWave;
  const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
  for (int64_t g = blockIdx.x; g < ngrp; g += gridDim.x) {
    for (int j = 0; j < kWave; ++j) {
      const int64_t r = (g * kWave + j) * kBlock + threadIdx.x;
      bool act = false;
      if (r < n) {
        const int i = B.inst[r];
        act = batch_active(B, B.state[i], B.params[i], r);
      }
      const uint64_t mask = __ballot(act);
      if (lane == 0) wcnt[wave] = (uint32_t)__popcll(mask);
      __syncthreads();
      if (threadIdx.x == 0) gcnt[j] = (int64_t)wcnt[0] + wcnt[1] + wcnt[2] + wcnt[3];
      __syncthreads();
    }
    if (wave == 0) {
      const int64_t t = g * kWave + lane;
      const int64_t mine = gcnt[lane];
      long long incl = (long long)mine;
#pragma unroll
      for (int off = 1; off < kWave; off <<= 1) {
        const long long v = __shfl_up(incl, off, kWave);
        if (lane >= off) incl += v;
      }
      if (t < ntiles) B.tile_off[t] = (int64_t)incl - mine;
      if (lane == kWave - 1) B.grp[g] = (int64_t)incl;
    }
    __syncthreads();
  }
}

// Base offsets of the groups: one block scans the group totals in rounds of
// 1024 (coalesced), carrying the running total between rounds.
__global__ __launch_bounds__(1024) void k_group_scan(BatchBuffers B)
{
  __shared__ int64_t wsum[kMaxWaves];
  __shared__ int64_t s_carry;
  if (B.flags[0] != 0) return;
  const int64_t n = *B.n;
  const int64_t ntiles = (n + kBlock - 1) / kBlock;
  const int64_t ngrp = (ntiles + kWave - 1) / kWave;
  const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
  const int W = blockDim.x / kWave;
  if (threadIdx.x == 0) s_carry = 0;
  __syncthreads();
  for (int64_t g0 = 0; g0 < ngrp; g0 += blockDim.x) {
    const int64_t g = g0 + threadIdx.x;
    const int64_t mine = g < ngrp ? B.grp[g] : 0;
    long long incl = (long long)mine;
#pragma unroll
    for (int off = 1; off < kWave; off <<= 1) {
      const long long v = __shfl_up(incl, off, kWave);
      if (lane >= off) incl += v;
    }
    if (lane == kWave - 1) wsum[wave] = (int64_t)incl;
    __syncthreads();
    int64_t base = s_carry;
    for (int w = 0; w < wave; ++w) base += wsum[w];
    if (g < ngrp) B.grp[B.gcap + g] = base + (int64_t)incl - mine;
    __syncthreads();
    if (threadIdx.x == 0) {
      int64_t t = 0;
      for (int w = 0; w < W; ++w) t += wsum[w];
      s_carry += t;
    }
    __syncthreads();
  }
}

__global__ __launch_bounds__(kBlock) void k_batch_split(BatchBuffers B)
{
  __shared__ uint32_t wcnt[kBlock / kWave];
  if (B.flags[0] != 0) return;
  const int64_t n = *B.n;
  const int64_t ntiles = (n + kBlock - 1) / kBlock;
  const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
  const uint64_t below = lane ? (~0ull >> (kWave - lane)) : 0ull;
  for (int64_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
    const int64_t r = t * kBlock + threadIdx.x;
    bool act = false;
    int i = 0;
    if (r < n) {
      i = B.inst[r];
      act = batch_active(B, B.state[i], B.params[i], r);
    }
    const uint64_t mask = __ballot(act);
    if (lane == 0) wcnt[wave] = (uint32_t)__popcll(mask);
    __syncthreads();
    int64_t rank = __popcll(mask & below);
    for (int w = 0; w < wave; ++w) rank += wcnt[w];
    __syncthreads();
    if (!act) continue;
    const int64_t dst = B.tile_off[t] + B.grp[B.gcap + (t >> 6)] + rank;
    const int a = B.axis[r];
    // every coordinate's bounds loaded before the first store (one memory
    // round trip per region instead of one per coordinate)
    double lo[kMaxDim], wd[kMaxDim];
#pragma unroll
    for (int m = 0; m < kMaxDim; ++m)
      if (m < B.dim) {
        lo[m] = B.lower[m * B.ld + r];
        wd[m] = B.width[m * B.ld + r];
      }
#pragma unroll
    for (int m = 0; m < kMaxDim; ++m)
      if (m < B.dim) {
        double2 L, Wd;
        if (m == a) {
          const double hw = 0.5 * wd[m];
          L = make_double2(lo[m], lo[m] + hw);
          Wd = make_double2(hw, hw);
        } else {
          L = make_double2(lo[m], lo[m]);
          Wd = make_double2(wd[m], wd[m]);
        }
        *reinterpret_cast<double2*>(B.out_lower + m * B.ld + 2 * dst) = L;
        *reinterpret_cast<double2*>(B.out_width + m * B.ld + 2 * dst) = Wd;
      }
    *reinterpret_cast<int2*>(B.out_inst + 2 * dst) = make_int2(i, i);
  }
}

// Every region of the list split on axis (round mod d): a uniform
// refinement used to build autotuning samples.
__global__ __launch_bounds__(kBlock) void k_refine_all(BatchBuffers B, int64_t n, int axis)
{
  const int64_t r = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (r >= n) return;
  for (int m = 0; m < B.dim; ++m) {
    const double lo = B.lower[m * B.ld + r];
    const double w = B.width[m * B.ld + r];
    double2 L, Wd;
    if (m == axis) {
      const double hw = 0.5 * w;
      L = make_double2(lo, lo + hw);
      Wd = make_double2(hw, hw);
    } else {
      L = make_double2(lo, lo);
      Wd = make_double2(w, w);
    }
    *reinterpret_cast<double2*>(B.out_lower + m * B.ld + 2 * r) = L;
    *reinterpret_cast<double2*>(B.out_width + m * B.ld + 2 * r) = Wd;
  }
  const int i = B.inst[r];
  *reinterpret_cast<int2*>(B.out_inst + 2 * r) = make_int2(i, i);
}

template <int D>
static void launch_batch_eval(const BatchBuffers& B, const RuleConst& R, int S, int grid,
                              unsigned fam_mask, hipStream_t st)
{
  const dim3 g(grid), b(kWave * seg_waves(S));
  const size_t lds = seg_lds_bytes(S);
  if (S == 1) {
    if (fam_mask & (1u << CUHRE_GAUSSIAN))
      hipLaunchKernelGGL((k_batch_eval<D, CUHRE_GAUSSIAN, true>), g, b, lds, st, B, R, S);
    if (fam_mask & (1u << CUHRE_PRODUCT_PEAK))
      hipLaunchKernelGGL((k_batch_eval<D, CUHRE_PRODUCT_PEAK, true>), g, b, lds, st, B, R, S);
    if (fam_mask & (1u << CUHRE_OSCILLATORY))
      hipLaunchKernelGGL((k_batch_eval<D, CUHRE_OSCILLATORY, true>), g, b, lds, st, B, R, S);
    if (fam_mask & (1u << CUHRE_CORNER_PEAK))
      hipLaunchKernelGGL((k_batch_eval<D, CUHRE_CORNER_PEAK, true>), g, b, lds, st, B, R, S);
    if (fam_mask & (1u << CUHRE_CONTINUOUS))
      hipLaunchKernelGGL((k_batch_eval<D, CUHRE_CONTINUOUS, true>), g, b, lds, st, B, R, S);
    return;
  }
  if (fam_mask & (1u << CUHRE_GAUSSIAN))
    hipLaunchKernelGGL((k_batch_eval<D, CUHRE_GAUSSIAN, false>), g, b, lds, st, B, R, S);
  if (fam_mask & (1u << CUHRE_PRODUCT_PEAK))
    hipLaunchKernelGGL((k_batch_eval<D, CUHRE_PRODUCT_PEAK, false>), g, b, lds, st, B, R, S);
  if (fam_mask & (1u << CUHRE_OSCILLATORY))
    hipLaunchKernelGGL((k_batch_eval<D, CUHRE_OSCILLATORY, false>), g, b, lds, st, B, R, S);
  if (fam_mask & (1u << CUHRE_CORNER_PEAK))
    hipLaunchKernelGGL((k_batch_eval<D, CUHRE_CORNER_PEAK, false>), g, b, lds, st, B, R, S);
  if (fam_mask & (1u << CUHRE_CONTINUOUS))
    hipLaunchKernelGGL((k_batch_eval<D, CUHRE_CONTINUOUS, false>), g, b, lds, st, B, R, S);
}

using BatchEvalLauncher = void (*)(const BatchBuffers&, const RuleConst&, int, int, unsigned,
                                   hipStream_t);

BatchEvalLauncher batch_eval_launcher(int dim)
{
  switch (dim) {
  case 2: return &launch_batch_eval<2>;
  case 3: return &launch_batch_eval<3>;
  case 4: return &launch_batch_eval<4>;
  case 5: return &launch_batch_eval<5>;
  case 6: return &launch_batch_eval<6>;
  case 7: return &launch_batch_eval<7>;
  case 8: return &launch_batch_eval<8>;
  case 9: return &launch_batch_eval<9>;
  case 10: return &launch_batch_eval<10>;
  default: return nullptr;
  }
}

// Host side of the batch engine (state lives in the context).
struct BatchEngine {
  int64_t cap = 0;  // regions per list buffer
  int dim_alloc = 0;
  int count_alloc = 0;
  double* lo[2] = {nullptr, nullptr};
  double* wd[2] = {nullptr, nullptr};
  int32_t* inst[2] = {nullptr, nullptr};
  double *est = nullptr, *err = nullptr, *abssum = nullptr, *vrel = nullptr;
  int8_t* axis = nullptr;
  int64_t *tile_off = nullptr, *grp = nullptr;
  Partial* chunk_part = nullptr;
  double2* chunk_est = nullptr;
  int64_t* chunk0 = nullptr;  // per instance (+1), kept with the states
  int64_t* n = nullptr;  // [0,1] list lengths, [2] peak, [3] overflow, [4] needed length, [5] chunks
  BatchParams* params = nullptr;
  InstState* state = nullptr;
  std::vector<InstState> h_state;
  std::vector<BatchParams> h_params;
  std::vector<int> perm;  // list position k holds caller instance perm[k]
  unsigned fam_mask = 0;
  int allow_cut = 0;
  int allot = 0;  // the batch uses the shared allotment (extra first pass)
  int64_t h_n[6] = {0, 0, 0, 0, 0, 0};
  // cached graph of two iterations for one buffer set and launch shape
  hipGraphExec_t gexec = nullptr;
  hipGraph_t graph = nullptr;
  int g_count = -1, g_dim = -1, g_S = -1, g_grid = -1, g_cut = -1, g_allot = -1;
  unsigned g_mask = 0;
  int64_t g_cap = -1;
  // autotune cache
  int tuned_dim = -1, tuned_S = 0, tuned_grid = 0;
  // persistent form: queue head + staging offset, per-instance staging positions
  int64_t* p_head = nullptr;
  int64_t* p_stage = nullptr;
  int p_count = 0;
  std::vector<int64_t> h_stage, h_chunk0;
};

static void drop_graph(BatchEngine* e)
{
  if (e->gexec) (void)hipGraphExecDestroy(e->gexec);
  if (e->graph) (void)hipGraphDestroy(e->graph);
  e->gexec = nullptr;
  e->graph = nullptr;
  e->g_count = -1;
}

struct ListArrays {
  double* lo[2] = {nullptr, nullptr};
  double* wd[2] = {nullptr, nullptr};
  int32_t* inst[2] = {nullptr, nullptr};
  double *est = nullptr, *err = nullptr, *abssum = nullptr, *vrel = nullptr;
  int8_t* axis = nullptr;
  int64_t *tile_off = nullptr, *grp = nullptr;
  Partial* chunk_part = nullptr;
  double2* chunk_est = nullptr;
};

static void free_arrays(ListArrays& a)
{
  for (int p = 0; p < 2; ++p) {
    if (a.lo[p]) (void)hipFree(a.lo[p]);
    if (a.wd[p]) (void)hipFree(a.wd[p]);
    if (a.inst[p]) (void)hipFree(a.inst[p]);
  }
  for (void* q : {(void*)a.est, (void*)a.err, (void*)a.abssum, (void*)a.vrel, (void*)a.axis, (void*)a.grp,
                  (void*)a.tile_off, (void*)a.chunk_part, (void*)a.chunk_est})
    if (q) (void)hipFree(q);
  a = ListArrays();
}

// Chunks a list of cap regions split among count instances can have.
static int64_t chunk_capacity(int64_t cap, int count) { return cap / kChunk + count + 2; }

static bool alloc_arrays(ListArrays& a, int64_t cap, int dim, int count)
{
  const size_t lb = sizeof(double) * (size_t)cap * dim;
  const int64_t ntiles = cap / kBlock + 1;
  const int64_t gcap = ntiles / kWave + 2;
  bool ok = true;
  for (int p = 0; p < 2; ++p)
    ok = ok && hipMalloc(&a.lo[p], lb) == hipSuccess && hipMalloc(&a.wd[p], lb) == hipSuccess &&
         hipMalloc(&a.inst[p], sizeof(int32_t) * cap) == hipSuccess;
  ok = ok && hipMalloc(&a.est, sizeof(double) * cap) == hipSuccess &&
       hipMalloc(&a.err, sizeof(double) * cap) == hipSuccess &&
       hipMalloc(&a.abssum, sizeof(double) * cap) == hipSuccess &&
       hipMalloc(&a.vrel, sizeof(double) * cap) == hipSuccess && hipMalloc(&a.axis, cap) == hipSuccess &&
       hipMalloc(&a.grp, sizeof(int64_t) * 2 * gcap) == hipSuccess &&
       hipMalloc(&a.chunk_part, sizeof(Partial) * chunk_capacity(cap, count)) == hipSuccess &&
       hipMalloc(&a.chunk_est, sizeof(double2) * chunk_capacity(cap, count)) == hipSuccess &&
       hipMalloc(&a.tile_off, sizeof(int64_t) * ntiles) == hipSuccess;
  if (!ok) {
    (void)hipGetLastError();
    free_arrays(a);
  }
  return ok;
}

static ListArrays take_arrays(BatchEngine* e)
{
  ListArrays a;
  for (int p = 0; p < 2; ++p) {
    a.lo[p] = e->lo[p];
    a.wd[p] = e->wd[p];
    a.inst[p] = e->inst[p];
  }
  a.est = e->est;
  a.err = e->err;
  a.abssum = e->abssum;
  a.vrel = e->vrel;
  a.axis = e->axis;
  a.grp = e->grp;
  a.tile_off = e->tile_off;
  a.chunk_part = e->chunk_part;
  a.chunk_est = e->chunk_est;
  return a;
}

static void put_arrays(BatchEngine* e, const ListArrays& a)
{
  for (int p = 0; p < 2; ++p) {
    e->lo[p] = a.lo[p];
    e->wd[p] = a.wd[p];
    e->inst[p] = a.inst[p];
  }
  e->est = a.est;
  e->err = a.err;
  e->abssum = a.abssum;
  e->vrel = a.vrel;
  e->axis = a.axis;
  e->grp = a.grp;
  e->tile_off = a.tile_off;
  e->chunk_part = a.chunk_part;
  e->chunk_est = a.chunk_est;
}

static void free_batch(BatchEngine* e)
{
  drop_graph(e);
  ListArrays a = take_arrays(e);
  free_arrays(a);
  put_arrays(e, a);
  for (void* q : {(void*)e->n, (void*)e->params, (void*)e->state, (void*)e->chunk0, (void*)e->p_head,
                  (void*)e->p_stage})
    if (q) (void)hipFree(q);
  e->p_head = nullptr;
  e->p_stage = nullptr;
  e->p_count = 0;
  e->chunk0 = nullptr;
  e->n = nullptr;
  e->params = nullptr;
  e->state = nullptr;
  e->cap = 0;
  e->dim_alloc = 0;
  e->count_alloc = 0;
}

BatchEngine* batch_engine_new() { return new BatchEngine(); }
void batch_engine_delete(BatchEngine* e)
{
  if (!e) return;
  free_batch(e);
  delete e;
}

static size_t bytes_per_region(int dim)
{
  return (size_t)dim * 4 * sizeof(double) + 2 * sizeof(int32_t) + 4 * sizeof(double) + 1 +
         2 * sizeof(int64_t) / kBlock + 1;
}

// Regions per list buffer that three quarters of the device memory holds
// (counting what this engine already owns as free).
static int64_t memory_cap(BatchEngine* e, int dim)
{
  size_t fr = 0, tot = 0;
  if (hipMemGetInfo(&fr, &tot) != hipSuccess) return 1 << 20;
  const double held = (double)e->cap * (double)bytes_per_region(std::max(dim, e->dim_alloc));
  return (int64_t)(((double)fr + held) * 0.75 / (double)bytes_per_region(dim));
}

// Fresh buffers (contents not kept).
static int batch_reserve(BatchEngine* e, int64_t cap, int dim, int count)
{
  if (cap <= e->cap && dim <= e->dim_alloc && count <= e->count_alloc) return 0;
  cap = (std::max(cap, e->cap) + 255) & ~255ll;
  dim = std::max(dim, e->dim_alloc);
  count = std::max(count, e->count_alloc);
  free_batch(e);
  ListArrays a;
  bool ok = alloc_arrays(a, cap, dim, count);
  ok = ok && hipMalloc(&e->n, sizeof(int64_t) * 6) == hipSuccess &&
       hipMalloc(&e->params, sizeof(BatchParams) * count) == hipSuccess &&
       hipMalloc(&e->state, sizeof(InstState) * count) == hipSuccess &&
       hipMalloc(&e->chunk0, sizeof(int64_t) * (count + 1)) == hipSuccess;
  put_arrays(e, a);
  if (!ok) {
    (void)hipGetLastError();
    free_batch(e);
    return set_error(CUHRE_ERR_ALLOC, "cannot allocate the batch region list");
  }
  e->cap = cap;
  e->dim_alloc = dim;
  e->count_alloc = count;
  return 0;
}

// Larger list buffers keeping list `parity` (n regions) and its per-region
// results; the instance states and parameters stay where they are.
static int batch_grow(BatchEngine* e, int64_t cap, int parity, int64_t n, hipStream_t st)
{
  cap = (cap + 255) & ~255ll;
  ListArrays a;
  if (!alloc_arrays(a, cap, e->dim_alloc, e->count_alloc))
    return set_error(CUHRE_ERR_ALLOC, "cannot grow the batch list");
  bool ok = true;
  if (n > 0) {
    const size_t rb = sizeof(double) * (size_t)n;
    ok = hipMemcpy2DAsync(a.lo[parity], sizeof(double) * cap, e->lo[parity], sizeof(double) * e->cap, rb,
                          e->dim_alloc, hipMemcpyDeviceToDevice, st) == hipSuccess &&
         hipMemcpy2DAsync(a.wd[parity], sizeof(double) * cap, e->wd[parity], sizeof(double) * e->cap, rb,
                          e->dim_alloc, hipMemcpyDeviceToDevice, st) == hipSuccess &&
         hipMemcpyAsync(a.inst[parity], e->inst[parity], sizeof(int32_t) * n, hipMemcpyDeviceToDevice, st) ==
             hipSuccess &&
         hipMemcpyAsync(a.est, e->est, rb, hipMemcpyDeviceToDevice, st) == hipSuccess &&
         hipMemcpyAsync(a.err, e->err, rb, hipMemcpyDeviceToDevice, st) == hipSuccess &&
         hipMemcpyAsync(a.abssum, e->abssum, rb, hipMemcpyDeviceToDevice, st) == hipSuccess &&
         hipMemcpyAsync(a.vrel, e->vrel, rb, hipMemcpyDeviceToDevice, st) == hipSuccess &&
         hipMemcpyAsync(a.axis, e->axis, n, hipMemcpyDeviceToDevice, st) == hipSuccess;
  }
  ok = ok && hipStreamSynchronize(st) == hipSuccess;
  if (!ok) {
    free_arrays(a);
    return set_error(CUHRE_ERR_DEVICE, "batch list copy failed");
  }
  drop_graph(e);
  ListArrays old = take_arrays(e);
  free_arrays(old);
  put_arrays(e, a);
  e->cap = cap;
  return 0;
}

// ------------------------------------------------------------ persistent form
// One launch runs the instances' loops.  Each workgroup takes the next
// instance from an atomic queue head and runs its whole plain loop in its
// own region queue -- two ping-pong lists of Q regions (slice w of the list
// buffers: lists 0 and 1, with est/err/abssum/vrel/axis beside them), a
// region per lane, block barriers between the evaluation, the reduction,
// the classification and the split -- then takes the next instance.  The
// loop step is k_inst_combine's (same sums, tests and counters).  An
// instance whose children would not fit in Q regions is handed over: its
// list (before that iteration) is copied to a staging area in list 1 beyond
// the queues, with its sums, and the replayed iterations continue it; when
// the staging area is full the instance restarts there from its initial
// grid.  No block waits for another: the queue head and the staging offset
// are the only shared state.
struct PersistArgs {
  const BatchParams* params;
  InstState* state;
  double* lo[2];
  double* wd[2];
  double *est, *err, *abssum, *vrel;
  int8_t* axis;
  int32_t* head;       // the family's queue head (next instance - first)
  unsigned long long* stage_off;  // staging area used so far (all families)
  int64_t* stage;      // per instance: staging position of a handed-over list, -1 restart, -2 done
  int64_t ld;          // list leading dimension (the engine's capacity)
  int64_t Q;           // regions per queue list
  int64_t stage0, stage_cap;  // staging area [stage0, stage0 + stage_cap) of list 1
  int32_t first, last;  // the family's instances
  int32_t npts, k0;
};

// A value the caller knows to be equal in every lane, in scalar registers.
__device__ __forceinline__ double uniform_d(double v)
{
  const long long b = __double_as_longlong(v);
  const unsigned lo = (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)b);
  const unsigned hi = (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)(b >> 32));
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// Block-shared state of the instance a persistent workgroup runs (thread 0
// takes the loop steps; the other threads read n, |I| and the decision).
struct PersistShared {
  Partial part[kBlock / kWave];
  Partial tot;
  InstState st, before;
  int64_t run, n;
  double absI;
  long long pos;
  uint32_t wcnt[kBlock / kWave];
  int ctl;  // 0: next iteration, 1: finished, 2: handed over
  int inst;
};

template <int D, int FAM>
__device__ void persist_instance(const PersistArgs& A, const RuleConst& R, const BatchParams& bp, int i,
                                 PersistShared& S)
{
  const int tid = threadIdx.x, lane = tid & (kWave - 1), wave = tid / kWave;
  const int64_t Q = A.Q, p = A.npts, ld = A.ld, q0 = (int64_t)blockIdx.x * Q;
  double* const lo0 = A.lo[0] + q0;
  double* const lo1 = A.lo[1] + q0;
  double* const wd0 = A.wd[0] + q0;
  double* const wd1 = A.wd[1] + q0;
  double* const est = A.est + q0;
  double* const err = A.err + q0;
  double* const abss = A.abssum + q0;
  double* const vrel = A.vrel + q0;
  int8_t* const axis = A.axis + q0;
  // the instance's parameters are uniform across the block: scalar registers
  LaneIntegrand<D, FAM> f;
  f.P.family = FAM;
  f.P.exponent = __builtin_amdgcn_readfirstlane(bp.exponent);
  double vol[D];  // widths of the instance's box (the allotment's vrel)
#pragma unroll
  for (int m = 0; m < D; ++m) {
    f.P.a[m] = uniform_d(bp.a[m]);
    f.P.w[m] = uniform_d(bp.w[m]);
    f.P.b[m] = uniform_d(bp.b[m]);
    vol[m] = uniform_d(bp.highs[m] - bp.lows[m]);
  }
  f.P.s0 = uniform_d(bp.s0);
  const bool allot = bp.allot > 0.0;

  // initial k^D grid in list 0 (low + id w rounded twice, as k_batch_init)
  int64_t n = 1;
  for (int m = 0; m < D; ++m) n *= A.k0;
  {
#pragma clang fp contract(off)
    for (int64_t r = tid; r < n; r += kBlock) {
      int64_t q = r;
      for (int m = 0; m < D; ++m) {
        const int64_t id = q % A.k0;
        q /= A.k0;
        const double w = (bp.highs[m] - bp.lows[m]) / (double)A.k0;
        lo0[m * ld + r] = bp.lows[m] + (double)id * w;
        wd0[m * ld + r] = w;
      }
    }
  }
  if (tid == 0) memset(&S.st, 0, sizeof S.st);
  int cur = 0;
  const uint64_t below = lane ? (~0ull >> (kWave - lane)) : 0ull;
  for (;;) {
    __syncthreads();  // the list written by the previous step is visible
    double* const lo = cur ? lo1 : lo0;
    double* const wd = cur ? wd1 : wd0;
    double* const clo = cur ? lo0 : lo1;
    double* const cwd = cur ? wd0 : wd1;
    // evaluation: a region per lane, passes of 256
    Partial acc;
    partial_zero(acc);
#pragma unroll 1
    for (int64_t b = 0; b < n; b += kBlock) {
      const int64_t r = b + tid;
      const bool valid = r < n;
      const int64_t rr = valid ? r : n - 1;
      const RegionOut o = seg_eval<D, false, true>(f, lo, wd, ld, r, valid, rr, 1, 0, wave, lane, kBlock / kWave,
                                                   nullptr, R, nullptr);
      if (valid) {
        est[r] = o.est;
        err[r] = o.err;
        abss[r] = o.ab;
        axis[r] = (int8_t)o.axis;
        if (allot) {
          double v = 1.0;
#pragma unroll
          for (int m = 0; m < D; ++m) v *= wd[m * ld + r] / vol[m];
          vrel[r] = v;
        }
        dd_add(acc.est_s, acc.est_c, o.est, 0.0);
        dd_add(acc.err_s, acc.err_c, o.err, 0.0);
        acc.abs_all += o.ab;
        acc.aest_all += fabs(o.est);
        if (!(isfinite(o.est) && isfinite(o.err))) acc.nnonfinite += 1;
      }
    }
    block_reduce_partial(acc, S.part, kBlock / kWave);
    if (tid == 0) {
      // the iteration's sums (k_inst_combine's step, before the classification)
      S.tot = acc;
      S.before = S.st;
      InstState& st = S.st;
      double rs = st.fin_s, rc = st.fin_c;
      dd_add(rs, rc, acc.est_s, acc.est_c);
      double es = st.fer_s, ec = st.fer_c;
      dd_add(es, ec, acc.err_s, acc.err_c);
      const double rG = rs + rc;
      st.iters += 1;
      st.nregions += n;
      st.neval += n * p;
      st.peak = n > st.peak ? n : st.peak;
      st.integral = rG;
      st.error = es + ec;
      st.bound = fp64_rounding_bound(p, D, st.nfin + n, st.abs_fin + acc.abs_all, st.aest_fin + acc.aest_all, rG);
      S.absI = fabs(rG);
      S.run = 0;
    }
    __syncthreads();
    const double absI = S.absI;
    // classification (the iteration's finished sums), stable compaction and
    // bisection into the other list; children that would not fit are not
    // written (the instance is then handed over)
    Partial fa;
    partial_zero(fa);
#pragma unroll 1
    for (int64_t b = 0; b < n; b += kBlock) {
      const int64_t r = b + tid;
      bool act = false;
      if (r < n) {
        const double e = est[r], er = err[r];
        act = inst_region_active(bp, absI, e, er, allot ? vrel[r] : 0.0);
        if (!act) {
          dd_add(fa.fest_s, fa.fest_c, e, 0.0);
          dd_add(fa.ferr_s, fa.ferr_c, er, 0.0);
          fa.abs_fin += abss[r];
          fa.aest_fin += fabs(e);
        }
      }
      const uint64_t mask = __ballot(act);
      if (lane == 0) S.wcnt[wave] = (uint32_t)__popcll(mask);
      __syncthreads();
      int64_t rank = S.run + __popcll(mask & below), blk = 0;
      for (int w = 0; w < kBlock / kWave; ++w) {
        if (w < wave) rank += S.wcnt[w];
        blk += S.wcnt[w];
      }
      if (act && 2 * rank + 1 < Q) {
        const int a = axis[r];
        for (int m = 0; m < D; ++m) {
          const double l = lo[m * ld + r], w = wd[m * ld + r];
          double2 L, W;
          if (m == a) {
            const double hw = 0.5 * w;
            L = make_double2(l, l + hw);
            W = make_double2(hw, hw);
          } else {
            L = make_double2(l, l);
            W = make_double2(w, w);
          }
          *reinterpret_cast<double2*>(clo + m * ld + 2 * rank) = L;
          *reinterpret_cast<double2*>(cwd + m * ld + 2 * rank) = W;
        }
      }
      __syncthreads();  // every wave has read run and wcnt
      if (tid == 0) S.run += blk;
    }
    block_reduce_partial(fa, S.part, kBlock / kWave);
    if (tid == 0) {
      InstState& st = S.st;
      const Partial& tot = S.tot;
      const int64_t nact = S.run;
      dd_add(st.fin_s, st.fin_c, fa.fest_s, fa.fest_c);
      dd_add(st.fer_s, st.fer_c, fa.ferr_s, fa.ferr_c);
      st.abs_fin += fa.abs_fin;
      st.aest_fin += fa.aest_fin;
      st.nfin += n - nact;
      st.next = 0;
      int ctl = 0;
      if (tot.nnonfinite > 0) {
        st.done = 1;
        st.fail = CUHRE_ERR_NONFINITE;
        ctl = 1;
      } else if (st.error <= fmax(bp.epsrel * fabs(st.integral), bp.epsabs)) {
        st.done = 1;
        st.fail = 0;
        ctl = 1;
      } else if (nact == 0 || st.neval + 2 * nact * p > bp.maxeval || st.iters >= 1000) {
        st.done = 1;
        st.fail = 1;
        ctl = 1;
      } else if (2 * nact > Q) {  // hand over the list of this iteration, before its step
        st = S.before;
        ctl = 2;
      }
      S.n = 2 * nact;
      S.ctl = ctl;
    }
    __syncthreads();
    const int ctl = __builtin_amdgcn_readfirstlane(S.ctl);
    if (ctl == 1) {
      if (tid == 0) {
        S.st.count = 0;
        A.state[i] = S.st;
        A.stage[i] = -2;
      }
      __syncthreads();
      return;
    }
    if (ctl == 2) break;
    n = S.n;
    cur ^= 1;
  }
  // hand-over: one atomic reservation of n regions in the staging area
  if (tid == 0) {
    const long long pos = (long long)atomicAdd(A.stage_off, (unsigned long long)n);
    S.pos = pos + n <= A.stage_cap ? pos : -1;
  }
  __syncthreads();
  const long long pos = S.pos;
  if (pos >= 0) {
    const double* const lo = cur ? lo1 : lo0;
    const double* const wd = cur ? wd1 : wd0;
    for (int64_t r = tid; r < n; r += kBlock)
      for (int m = 0; m < D; ++m) {
        A.lo[1][m * ld + A.stage0 + pos + r] = lo[m * ld + r];
        A.wd[1][m * ld + A.stage0 + pos + r] = wd[m * ld + r];
      }
  }
  if (tid == 0) {
    if (pos < 0) memset(&S.st, 0, sizeof S.st);  // restarts from its initial grid
    S.st.count = pos >= 0 ? n : 0;
    A.state[i] = S.st;
    A.stage[i] = pos >= 0 ? pos : -1;
  }
  __syncthreads();
}

// One persistent kernel per family: the family's instances are the range
// [first, last) of the family-sorted list order, taken through its own
// queue head.
template <int D, int FAM>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(3, 3))) void k_batch_persist(PersistArgs A, RuleConst R)
{
  __shared__ PersistShared S;
  for (;;) {
    if (threadIdx.x == 0) S.inst = A.first + atomicAdd(A.head, 1);
    __syncthreads();
    const int i = __builtin_amdgcn_readfirstlane(S.inst);
    __syncthreads();  // read before thread 0 takes the next one
    if (i >= A.last) return;
    persist_instance<D, FAM>(A, R, A.params[i], i, S);
  }
}

// After the persistent launch: the instances still running, in instance
// order, into list 0 at their new starts (their lists from the staging
// area, or their initial grids), one block per instance.
__global__ __launch_bounds__(kBlock) void k_batch_gather(BatchBuffers B, const int64_t* stage, int64_t stage0,
                                                        int k)
{
#pragma clang fp contract(off)  // low + id * w rounded twice, as the oracle computes it
  const int i = blockIdx.x;
  const InstState& st = B.state[i];
  if (st.done) return;
  const int64_t s0 = st.start, c = st.count;
  const BatchParams& bp = B.params[i];
  for (int64_t r = threadIdx.x; r < c; r += kBlock) {
    if (stage[i] >= 0) {
      for (int m = 0; m < B.dim; ++m) {
        B.lower[m * B.ld + s0 + r] = B.out_lower[m * B.ld + stage0 + stage[i] + r];
        B.width[m * B.ld + s0 + r] = B.out_width[m * B.ld + stage0 + stage[i] + r];
      }
    } else {
      int64_t q = r;
      for (int m = 0; m < B.dim; ++m) {
        const int64_t id = q % k;
        q /= k;
        const double w = (bp.highs[m] - bp.lows[m]) / (double)k;
        B.lower[m * B.ld + s0 + r] = bp.lows[m] + (double)id * w;
        B.width[m * B.ld + s0 + r] = w;
      }
    }
    B.inst[s0 + r] = i;
  }
}

using PersistLauncher = void (*)(const PersistArgs&, const RuleConst&, int, hipStream_t);

template <int D, int FAM>
static void launch_persist(const PersistArgs& A, const RuleConst& R, int grid, hipStream_t st)
{
  hipLaunchKernelGGL((k_batch_persist<D, FAM>), dim3(grid), dim3(kBlock), 0, st, A, R);
}

// The family's persistent kernel at D and its resident blocks per CU.
template <int D>
static PersistLauncher persist_launcher_d(int family, int* per_cu)
{
  PersistLauncher l = nullptr;
  const void* k = nullptr;
  switch (family) {
  case CUHRE_GAUSSIAN: l = &launch_persist<D, CUHRE_GAUSSIAN>; k = (const void*)k_batch_persist<D, CUHRE_GAUSSIAN>; break;
  case CUHRE_PRODUCT_PEAK:
    l = &launch_persist<D, CUHRE_PRODUCT_PEAK>;
    k = (const void*)k_batch_persist<D, CUHRE_PRODUCT_PEAK>;
    break;
  case CUHRE_OSCILLATORY:
    l = &launch_persist<D, CUHRE_OSCILLATORY>;
    k = (const void*)k_batch_persist<D, CUHRE_OSCILLATORY>;
    break;
  case CUHRE_CORNER_PEAK:
    l = &launch_persist<D, CUHRE_CORNER_PEAK>;
    k = (const void*)k_batch_persist<D, CUHRE_CORNER_PEAK>;
    break;
  case CUHRE_CONTINUOUS:
    l = &launch_persist<D, CUHRE_CONTINUOUS>;
    k = (const void*)k_batch_persist<D, CUHRE_CONTINUOUS>;
    break;
  default: return nullptr;
  }
  int nb = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k, kBlock, 0) != hipSuccess) return nullptr;
  *per_cu = nb;
  return l;
}

static PersistLauncher persist_launcher(int dim, int family, int* per_cu)
{
  switch (dim) {
  case 2: return persist_launcher_d<2>(family, per_cu);
  case 3: return persist_launcher_d<3>(family, per_cu);
  case 4: return persist_launcher_d<4>(family, per_cu);
  case 5: return persist_launcher_d<5>(family, per_cu);
  case 6: return persist_launcher_d<6>(family, per_cu);
  case 7: return persist_launcher_d<7>(family, per_cu);
  case 8: return persist_launcher_d<8>(family, per_cu);
  case 9: return persist_launcher_d<9>(family, per_cu);
  case 10: return persist_launcher_d<10>(family, per_cu);
  default: return nullptr;
  }
}

static BatchBuffers make_buffers(BatchEngine* e, int parity, int count, int dim)
{
  BatchBuffers B;
  B.lower = e->lo[parity];
  B.width = e->wd[parity];
  B.inst = e->inst[parity];
  B.out_lower = e->lo[1 - parity];
  B.out_width = e->wd[1 - parity];
  B.out_inst = e->inst[1 - parity];
  B.ld = e->cap;
  B.n = e->n + parity;
  B.n_next = e->n + (1 - parity);
  B.n_peak = reinterpret_cast<unsigned long long*>(e->n + 2);
  B.flags = e->n + 3;
  B.parity = parity;
  B.allow_cut = e->allow_cut;
  B.est = e->est;
  B.err = e->err;
  B.abssum = e->abssum;
  B.vrel = e->vrel;
  B.axis = e->axis;
  B.tile_off = e->tile_off;
  B.grp = e->grp;
  B.gcap = (e->cap / kBlock + 1) / kWave + 2;
  B.chunk_part = e->chunk_part;
  B.chunk_est = e->chunk_est;
  B.chunk0 = e->chunk0;
  B.nch = e->n + 5;
  B.params = e->params;
  B.state = e->state;
  B.count = count;
  B.dim = dim;
  B.npts = gm_points(dim);
  B.cap = e->cap;
  return B;
}

static void enqueue_eval_reduce(BatchEngine* e, int parity, int count, int dim, const RuleConst& R, int S,
                                int grid, hipStream_t st)
{
  const BatchBuffers B = make_buffers(e, parity, count, dim);
  batch_eval_launcher(dim)(B, R, S, grid, e->fam_mask, st);
  if (e->allot) {
    hipLaunchKernelGGL(k_chunk_reduce<true>, dim3(grid), dim3(kBlock), 0, st, B);
    hipLaunchKernelGGL(k_inst_absI, dim3(count), dim3(kBlock), 0, st, B);
  }
  hipLaunchKernelGGL(k_chunk_reduce<false>, dim3(grid), dim3(kBlock), 0, st, B);
  hipLaunchKernelGGL(k_inst_combine, dim3(count), dim3(kBlock), 0, st, B);
}

static void enqueue_scan_split(BatchEngine* e, int parity, int count, int dim, int grid, hipStream_t st)
{
  const BatchBuffers B = make_buffers(e, parity, count, dim);
  hipLaunchKernelGGL(k_batch_scan, dim3(1), dim3(1024), 0, st, B);
  hipLaunchKernelGGL(k_tile_count, dim3(grid), dim3(kBlock), 0, st, B);
  hipLaunchKernelGGL(k_group_scan, dim3(1), dim3(1024), 0, st, B);
  hipLaunchKernelGGL(k_batch_split, dim3(grid), dim3(kBlock), 0, st, B);
}

static void enqueue_iteration(BatchEngine* e, int parity, int count, int dim, const RuleConst& R, int S,
                              int grid, hipStream_t st)
{
  enqueue_eval_reduce(e, parity, count, dim, R, S, grid, st);
  enqueue_scan_split(e, parity, count, dim, grid, st);
}

static int check_batch_args(int count, const cuhre_integrand* fs, int dim)
{
  if (count <= 0 || !fs) return set_error(CUHRE_ERR_ARG, "empty batch");
  if (dim < 2 || dim > 10) return set_error(CUHRE_ERR_UNSUPPORTED, "batch dimension outside 2..10");
  for (int i = 0; i < count; ++i) {
    const int f = fs[i].family;
    if (fs[i].dim != dim) return set_error(CUHRE_ERR_ARG, "instance dimension differs from dim");
    if (f != CUHRE_GAUSSIAN && f != CUHRE_PRODUCT_PEAK && f != CUHRE_OSCILLATORY &&
        f != CUHRE_CORNER_PEAK && f != CUHRE_CONTINUOUS)
      return set_error(CUHRE_ERR_UNSUPPORTED, "family not available in batches");
    if (f == CUHRE_CORNER_PEAK && fs[i].exponent < 0)
      return set_error(CUHRE_ERR_ARG, "corner peak exponent must be >= 0");
  }
  return 0;
}

// Parameters in list order: instances sorted (stably) by family.
static void fill_params(std::vector<BatchParams>& hp, std::vector<int>& perm, unsigned& mask, int count,
                        const cuhre_integrand* fs, int dim, const double* epsrel, const double* epsabs,
                        const int64_t* maxeval, const double* lows, const double* highs, double allot)
{
  perm.resize(count);
  for (int i = 0; i < count; ++i) perm[i] = i;
  std::stable_sort(perm.begin(), perm.end(), [&](int x, int y) { return fs[x].family < fs[y].family; });
  mask = 0;
  hp.assign(count, BatchParams{});
  for (int k = 0; k < count; ++k) {
    const int i = perm[k];
    mask |= 1u << fs[i].family;
    BatchParams& b = hp[k];
    std::memset(&b, 0, sizeof b);
    const cuhre_integrand& f = fs[i];
    b.family = f.family;
    b.exponent = f.exponent;
    for (int m = 0; m < kMaxDim; ++m) {
      b.a[m] = f.a[m];
      b.w[m] = f.w[m];
      b.b[m] = (m < dim && f.family == CUHRE_PRODUCT_PEAK) ? 1.0 / (f.a[m] * f.a[m]) : 0.0;
    }
    b.s0 = 2.0 * M_PI * f.u;
    b.epsrel = epsrel ? epsrel[i] : 1e-3;
    b.epsabs = epsabs ? epsabs[i] : 0.0;
    b.maxeval = maxeval ? maxeval[i] : 1000000000LL;
    b.allot = allot;
    for (int m = 0; m < dim; ++m) {
      b.lows[m] = lows ? lows[(size_t)i * dim + m] : 0.0;
      b.highs[m] = highs ? highs[(size_t)i * dim + m] : 1.0;
    }
  }
}

static int grid_k_batch(int d) { return d <= 4 ? 4 : (d <= 10 ? 2 : 1); }

// Uploads the batch and builds its initial list in buffer 0.
static int batch_setup(BatchEngine* e, hipStream_t st, int count, const cuhre_integrand* fs, int dim,
                       const double* epsrel, const double* epsabs, const int64_t* maxeval,
                       const double* lows, const double* highs, int64_t cap, int64_t* n0_out,
                       double allot = 0.0)
{
  const int k = grid_k_batch(dim);
  int64_t n0 = 1;
  for (int m = 0; m < dim; ++m) n0 *= k;
  const int64_t total = n0 * count;
  cap = std::max(cap, 2 * total);
  int rc = batch_reserve(e, cap, dim, count);
  if (rc) return rc;
  std::vector<BatchParams>& hp = e->h_params;
  fill_params(hp, e->perm, e->fam_mask, count, fs, dim, epsrel, epsabs, maxeval, lows, highs, allot);
  if (hipMemcpyAsync(e->params, hp.data(), sizeof(BatchParams) * count, hipMemcpyHostToDevice, st) !=
      hipSuccess)
    return set_error(CUHRE_ERR_DEVICE, "batch parameter upload failed");
  const BatchBuffers B = make_buffers(e, 0, count, dim);
  const int64_t nthreads = std::max<int64_t>(total, count);
  hipLaunchKernelGGL(k_batch_init, dim3((unsigned)((nthreads + 255) / 256)), dim3(256), 0, st, B, k, n0);
  e->h_n[0] = total;
  e->h_n[1] = 0;
  e->h_n[2] = total;
  e->h_n[3] = 0;
  e->h_n[4] = 0;
  e->h_n[5] = 0;  // the chunk count is written by the init kernel
  e->allot = allot > 0.0 ? 1 : 0;
  if (hipMemcpyAsync(e->n, e->h_n, 5 * sizeof(int64_t), hipMemcpyHostToDevice, st) != hipSuccess)
    return set_error(CUHRE_ERR_DEVICE, "batch setup copy failed");
  if (hipGetLastError() != hipSuccess) return set_error(CUHRE_ERR_DEVICE, "batch init launch failed");
  *n0_out = total;
  return 0;
}

static bool capture_graph(BatchEngine* e, int count, int dim, const RuleConst& R, int S, int G, hipStream_t st)
{
  const bool same = e->gexec && e->g_count == count && e->g_dim == dim && e->g_S == S && e->g_grid == G &&
                    e->g_cap == e->cap && e->g_mask == e->fam_mask && e->g_cut == e->allow_cut &&
                    e->g_allot == e->allot;
  if (same) return true;
  drop_graph(e);
  bool ok = false;
  if (hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal) == hipSuccess) {
    enqueue_iteration(e, 0, count, dim, R, S, G, st);
    enqueue_iteration(e, 1, count, dim, R, S, G, st);
    if (hipStreamEndCapture(st, &e->graph) == hipSuccess &&
        hipGraphInstantiate(&e->gexec, e->graph, nullptr, nullptr, 0) == hipSuccess) {
      ok = true;
      e->g_count = count;
      e->g_dim = dim;
      e->g_S = S;
      e->g_grid = G;
      e->g_cap = e->cap;
      e->g_mask = e->fam_mask;
      e->g_cut = e->allow_cut;
      e->g_allot = e->allot;
    }
  }
  (void)hipGetLastError();
  return ok;
}

// Families present in a batch (bit f for family f).
static unsigned batch_family_mask(int count, const cuhre_integrand* fs)
{
  unsigned m = 0;
  for (int i = 0; i < count; ++i) m |= 1u << fs[i].family;
  return m;
}

// Resident workgroups of the families' persistent kernels (the largest)
// and the list length their queues and staging area want: 2 x kPersistQ
// regions per workgroup.
constexpr int64_t kPersistQ = 16384;

static int persist_grid(int device, int dim, unsigned fam_mask)
{
  int ncu = 0, most = 0;
  if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess) return 0;
  for (int f = 0; f < 32; ++f) {
    if (!(fam_mask & (1u << f))) continue;
    int per_cu = 0;
    if (!persist_launcher(dim, f, &per_cu)) return 0;
    most = std::max(most, ncu * std::max(per_cu, 1));
  }
  return most;
}

static int64_t persist_regions(int device, int dim, unsigned fam_mask)
{
  return 2 * kPersistQ * persist_grid(device, dim, fam_mask);
}

// The persistent launches (one per family), then the instances they handed
// over gathered into list 0 in instance order with their states and chunk
// map, ready for the replayed iterations (an empty list when every instance
// finished).
static int run_persistent(BatchEngine* e, hipStream_t st, int device, int count, int dim, const RuleConst& R,
                          int* finished)
{
  const int gmax = std::min(persist_grid(device, dim, e->fam_mask), count);
  if (gmax <= 0) return 0;
  const int64_t Q = std::min<int64_t>(kPersistQ, (e->cap / 2 / gmax) & ~(int64_t)(kBlock - 1));
  if (Q < 1024) return 0;  // lists too short for useful queues: replayed iterations only
  if (e->p_count < count) {
    if (e->p_head) (void)hipFree(e->p_head);
    if (e->p_stage) (void)hipFree(e->p_stage);
    e->p_head = nullptr;
    e->p_stage = nullptr;
    e->p_count = 0;
    if (hipMalloc(&e->p_head, 16 * sizeof(int64_t)) != hipSuccess ||
        hipMalloc(&e->p_stage, sizeof(int64_t) * count) != hipSuccess) {
      (void)hipGetLastError();
      return set_error(CUHRE_ERR_ALLOC, "cannot allocate the persistent queue state");
    }
    e->p_count = count;
  }
  if (hipMemsetAsync(e->p_head, 0, 16 * sizeof(int64_t), st) != hipSuccess)
    return set_error(CUHRE_ERR_DEVICE, "persistent queue reset failed");
  PersistArgs A;
  A.params = e->params;
  A.state = e->state;
  for (int q = 0; q < 2; ++q) {
    A.lo[q] = e->lo[q];
    A.wd[q] = e->wd[q];
  }
  A.est = e->est;
  A.err = e->err;
  A.abssum = e->abssum;
  A.vrel = e->vrel;
  A.axis = e->axis;
  A.stage_off = reinterpret_cast<unsigned long long*>(e->p_head + 8);
  A.stage = e->p_stage;
  A.ld = e->cap;
  A.Q = Q;
  A.stage0 = (int64_t)gmax * Q;
  A.stage_cap = e->cap - A.stage0;
  A.npts = gm_points(dim);
  A.k0 = grid_k_batch(dim);
  // each family's instances are a contiguous range of the family-sorted order
  for (int k = 0; k < count;) {
    const int f = e->h_params[k].family;
    int j = k;
    while (j < count && e->h_params[j].family == f) ++j;
    int per_cu = 0, ncu = 0;
    const PersistLauncher launch = persist_launcher(dim, f, &per_cu);
    if (!launch || hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess)
      return set_error(CUHRE_ERR_DEVICE, "no persistent kernel for a family");
    A.head = reinterpret_cast<int32_t*>(e->p_head + f);
    A.first = k;
    A.last = j;
    launch(A, R, std::min(std::min(ncu * std::max(per_cu, 1), gmax), j - k), st);
    k = j;
  }
  if (hipGetLastError() != hipSuccess) return set_error(CUHRE_ERR_DEVICE, "persistent launch failed");
  e->h_state.resize(count);
  e->h_stage.resize(count);
  if (hipMemcpyAsync(e->h_state.data(), e->state, sizeof(InstState) * count, hipMemcpyDeviceToHost, st) !=
          hipSuccess ||
      hipMemcpyAsync(e->h_stage.data(), e->p_stage, sizeof(int64_t) * count, hipMemcpyDeviceToHost, st) !=
          hipSuccess ||
      hipStreamSynchronize(st) != hipSuccess)
    return set_error(CUHRE_ERR_DEVICE, "persistent result copy failed");
  // the handed-over instances' slices of list 0 in instance order
  int64_t n0 = 1;
  for (int m = 0; m < dim; ++m) n0 *= A.k0;
  e->h_chunk0.assign(count + 1, 0);
  int64_t run = 0, nch = 0;
  int done = 0;
  for (int i = 0; i < count; ++i) {
    InstState& s = e->h_state[i];
    e->h_chunk0[i] = nch;
    if (s.done) {
      ++done;
      s.count = 0;
      continue;
    }
    s.start = run;
    s.count = e->h_stage[i] >= 0 ? s.count : n0;
    run += s.count;
    nch += (s.count + kChunk - 1) / kChunk;
  }
  e->h_chunk0[count] = nch;
  *finished = done;
  if (run > e->cap) return set_error(CUHRE_ERR_CAPACITY, "handed-over lists exceed the list buffers");
  e->h_n[0] = run;
  e->h_n[1] = 0;
  e->h_n[2] = run;
  e->h_n[3] = 0;
  e->h_n[4] = 0;
  e->h_n[5] = nch;
  if (hipMemcpyAsync(e->state, e->h_state.data(), sizeof(InstState) * count, hipMemcpyHostToDevice, st) !=
          hipSuccess ||
      hipMemcpyAsync(e->chunk0, e->h_chunk0.data(), sizeof(int64_t) * (count + 1), hipMemcpyHostToDevice, st) !=
          hipSuccess ||
      hipMemcpyAsync(e->n, e->h_n, 6 * sizeof(int64_t), hipMemcpyHostToDevice, st) != hipSuccess)
    return set_error(CUHRE_ERR_DEVICE, "persistent hand-over upload failed");
  if (run > 0) {
    const BatchBuffers B = make_buffers(e, 0, count, dim);
    hipLaunchKernelGGL(k_batch_gather, dim3((unsigned)count), dim3(kBlock), 0, st, B, (const int64_t*)e->p_stage,
                       A.stage0, A.k0);
    if (hipGetLastError() != hipSuccess) return set_error(CUHRE_ERR_DEVICE, "hand-over gather launch failed");
  }
  return 0;
}

}  // namespace cuhre

using namespace cuhre;

// The context owns one batch engine (defined in engine.hip).
extern "C" BatchEngine* cuhre_ctx_batch(cuhre_context* c, hipStream_t* stream, int* device, int64_t* capacity);

extern "C" {

int cuhre_batch_autotune(cuhre_context* ctx, int32_t count, const cuhre_integrand* fs, int dim,
                         const double* lows, const double* highs, cuhre_batch_config* best,
                         double* table_ms)
{
  if (!ctx || !best) return set_error(CUHRE_ERR_ARG, "null argument");
  int rc = check_batch_args(count, fs, dim);
  if (rc) return rc;
  hipStream_t st;
  int device;
  int64_t capacity;
  BatchEngine* e = cuhre_ctx_batch(ctx, &st, &device, &capacity);
  if (hipSetDevice(device) != hipSuccess) return set_error(CUHRE_ERR_DEVICE, "hipSetDevice failed");
  // sample list: the batch's initial grids refined uniformly to >= 512k regions
  const int64_t want = 1 << 19;
  int64_t n0 = 0;
  rc = batch_setup(e, st, count, fs, dim, nullptr, nullptr, nullptr, lows, highs, 4 * want, &n0);
  if (rc) return rc;
  int64_t n = n0;
  int parity = 0, round = 0;
  while (n < want && 2 * n <= e->cap) {
    const BatchBuffers B = make_buffers(e, parity, count, dim);
    hipLaunchKernelGGL(k_refine_all, dim3((unsigned)((n + kBlock - 1) / kBlock)), dim3(kBlock), 0, st, B, n,
                       round % dim);
    n *= 2;
    parity = 1 - parity;
    ++round;
  }
  e->h_n[parity] = n;
  if (hipMemcpyAsync(e->n + parity, e->h_n + parity, sizeof n, hipMemcpyHostToDevice, st) != hipSuccess)
    return set_error(CUHRE_ERR_DEVICE, "autotune copy failed");
  const RuleConst R = make_rule(dim, 7);
  const BatchBuffers B = make_buffers(e, parity, count, dim);
  const int Ss[3] = {1, 2, 4};
  const int Gs[5] = {256, 512, 1024, 2048, 4096};
  hipEvent_t a, b;
  if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess)
    return set_error(CUHRE_ERR_DEVICE, "event create failed");
  float best_ms = 1e30f;
  int bs = 1, bg = 1024, idx = 0;
  for (int S : Ss)
    for (int G : Gs) {
      batch_eval_launcher(dim)(B, R, S, G, e->fam_mask, st);  // warm
      (void)hipEventRecord(a, st);
      for (int rep = 0; rep < 3; ++rep) batch_eval_launcher(dim)(B, R, S, G, e->fam_mask, st);
      (void)hipEventRecord(b, st);
      (void)hipEventSynchronize(b);
      float ms = 0.f;
      (void)hipEventElapsedTime(&ms, a, b);
      ms /= 3.0f;
      if (table_ms) table_ms[idx] = ms;
      ++idx;
      if (ms < best_ms) {
        best_ms = ms;
        bs = S;
        bg = G;
      }
    }
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  if (hipGetLastError() != hipSuccess) return set_error(CUHRE_ERR_DEVICE, "autotune launch failed");
  best->point_segments = bs;
  best->grid_blocks = bg;
  best->iterations_per_check = 2;
  best->use_graph = 1;
  best->persistent = 0;
  best->initial_capacity = 0;
  best->finish_budget = 0.0;
  e->tuned_dim = dim;
  e->tuned_S = bs;
  e->tuned_grid = bg;
  return 0;
}

int cuhre_integrate_batch(cuhre_context* ctx, int32_t count, const cuhre_integrand* fs, int dim,
                          const double* epsrel, const double* epsabs, const int64_t* maxeval,
                          const double* lows, const double* highs, const cuhre_batch_config* cfg,
                          cuhre_result* out, cuhre_batch_info* info)
{
  const auto t0 = std::chrono::steady_clock::now();
  if (!ctx || !out) return set_error(CUHRE_ERR_ARG, "null argument");
  int rc = check_batch_args(count, fs, dim);
  if (rc) return rc;
  hipStream_t st;
  int device;
  int64_t capacity;
  BatchEngine* e = cuhre_ctx_batch(ctx, &st, &device, &capacity);
  if (hipSetDevice(device) != hipSuccess) return set_error(CUHRE_ERR_DEVICE, "hipSetDevice failed");
  int S = cfg && cfg->point_segments > 0 ? cfg->point_segments : 0;
  int G = cfg && cfg->grid_blocks > 0 ? cfg->grid_blocks : 0;
  if (e->tuned_dim == dim) {
    S = S ? S : e->tuned_S;
    G = G ? G : e->tuned_grid;
  }
  S = S ? S : 1;
  G = G ? G : 2048;
  if (S != 1 && S != 2 && S != 4) return set_error(CUHRE_ERR_ARG, "batch point_segments must be 1, 2 or 4");
  const int per_check = cfg && cfg->iterations_per_check > 0 ? cfg->iterations_per_check : 2;
  const double allot = cfg ? cfg->finish_budget : 0.0;
  if (!(allot >= 0.0) || allot >= 1.0) return set_error(CUHRE_ERR_ARG, "batch finish_budget must be in [0, 1)");
  const bool use_graph = !cfg || cfg->use_graph != 0;
  const bool persistent = cfg && cfg->persistent != 0;
  // list capacity: an explicit engine capacity is a hard limit; otherwise the
  // list starts small and grows, up to three quarters of the device memory
  const int64_t limit = capacity > 0 ? capacity : memory_cap(e, dim);
  int64_t n0 = 0;
  const int64_t first = cfg && cfg->initial_capacity > 0 ? cfg->initial_capacity : (1 << 22);
  if (cfg && cfg->initial_capacity > 0 && e->cap > first) free_batch(e);  // honour a smaller start
  // the persistent form's queues (2 x 16384 regions per resident workgroup)
  // and its staging area share the list buffers
  const int64_t want = persistent ? std::max<int64_t>(first, persist_regions(device, dim, batch_family_mask(count, fs))) : first;
  rc = batch_setup(e, st, count, fs, dim, epsrel, epsabs, maxeval, lows, highs,
                   std::min<int64_t>(limit, std::max<int64_t>(e->cap, want)), &n0, allot);
  if (rc) return rc;
  e->allow_cut = e->cap >= limit ? 1 : 0;
  const RuleConst R = make_rule(dim, 7);
  int persist_finished = -1;
  if (persistent && (rc = run_persistent(e, st, device, count, dim, R, &persist_finished))) return rc;
  bool graph_ok = use_graph && capture_graph(e, count, dim, R, S, G, st);
  int iters = 0, grows = 0;
  const bool all_done = persist_finished >= 0 && e->h_n[0] == 0;  // nothing handed over
  for (int check = 0; check < 2000 && !all_done; ++check) {
    for (int k = 0; k < per_check; ++k) {
      if (graph_ok) {
        if (hipGraphLaunch(e->gexec, st) != hipSuccess) return set_error(CUHRE_ERR_DEVICE, "graph launch failed");
      } else {
        enqueue_iteration(e, 0, count, dim, R, S, G, st);
        enqueue_iteration(e, 1, count, dim, R, S, G, st);
      }
      iters += 2;
    }
    if (hipGetLastError() != hipSuccess) return set_error(CUHRE_ERR_DEVICE, "batch iteration launch failed");
    if (hipMemcpyAsync(e->h_n, e->n, sizeof e->h_n, hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
      return set_error(CUHRE_ERR_DEVICE, "batch sync failed");
    if (e->h_n[3] != 0) {
      // an iteration ran out of room after its reduction: grow, then finish it
      const int p = (int)e->h_n[3] - 1;
      const int64_t need = e->h_n[4];
      const int64_t ncap = std::min<int64_t>(limit, std::max<int64_t>(2 * e->cap, need + need / 4));
      if ((rc = batch_grow(e, ncap, p, e->h_n[p], st))) return rc;
      e->allow_cut = e->cap >= limit || ncap < need ? 1 : 0;
      e->h_n[3] = 0;
      e->h_n[4] = 0;
      if (hipMemcpyAsync(e->n + 3, e->h_n + 3, 2 * sizeof(int64_t), hipMemcpyHostToDevice, st) != hipSuccess)
        return set_error(CUHRE_ERR_DEVICE, "batch flag reset failed");
      enqueue_scan_split(e, p, count, dim, G, st);
      if (p == 0) enqueue_iteration(e, 1, count, dim, R, S, G, st);  // back in step with the graph
      ++grows;
      graph_ok = use_graph && capture_graph(e, count, dim, R, S, G, st);
      continue;
    }
    if (e->h_n[0] == 0) break;  // an even number of iterations leaves the list in buffer 0
  }
  e->h_state.resize(count);
  if (hipMemcpyAsync(e->h_state.data(), e->state, sizeof(InstState) * count, hipMemcpyDeviceToHost, st) !=
          hipSuccess ||
      hipStreamSynchronize(st) != hipSuccess)
    return set_error(CUHRE_ERR_DEVICE, "batch result copy failed");
  int worst = 0, nonconv = 0, maxit = 0;
  int64_t total_regions = 0;
  const int64_t p = gm_points(dim);
  for (int k = 0; k < count; ++k) {
    const InstState& s = e->h_state[k];
    cuhre_result& r = out[e->perm[k]];
    std::memset(&r, 0, sizeof r);
    r.integral = s.integral;
    r.error = s.error;
    r.fp64_err_bound = s.bound;
    r.nregions = s.nregions;
    r.neval = s.nregions * p;
    // per instance: 0 converged, 1 not converged, CUHRE_ERR_NONFINITE (-2)
    // when its integrand returned a non-finite value
    r.fail = s.fail != 0 ? s.fail : (s.done ? 0 : 1);
    r.iterations = s.iters;
    r.peak_regions = s.peak;
    r.nfinished = s.nfin;
    total_regions += s.nregions;
    maxit = std::max(maxit, s.iters);
    if (s.fail < 0) worst = s.fail;
    if (s.fail != 0 || !s.done) ++nonconv;
  }
  const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  if (info) {
    std::memset(info, 0, sizeof *info);
    info->iterations = maxit;
    info->launched_iterations = iters;
    info->total_regions = total_regions;
    info->total_neval = total_regions * p;
    info->not_converged = nonconv;
    info->ms_total = ms;
    info->persistent_finished = persist_finished;
    info->graph_used = graph_ok ? 1 : 0;
    info->point_segments = S;
    info->grid_blocks = G;
    info->peak_list = e->h_n[2];
    info->capacity = e->cap;
    info->grows = grows;
  }
  for (int i = 0; i < count; ++i) out[i].ms_total = ms;
  if (worst < 0)
    return set_error(worst, "an integrand returned a non-finite value in some instance (its result's fail "
                            "is -2; the other results are complete)");
  return nonconv ? CUHRE_NOT_CONVERGED : CUHRE_OK;
}

}  // extern "C"
