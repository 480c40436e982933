// reduce.hpp -- deterministic FP64 wave/block reductions and the
// finalisation of an iteration.
//
// The reference sums est/err with four thrust::reduce calls and a blocking
// copy per iteration (quad/GPUquad/GPUKernelquad.cu:565-577) and counts the
// active regions with thrust::exclusive_scan + two device-to-host copies
// (:364-386).  Here each evaluation block folds its regions into one Partial
// (block_reduce_regions: field by field in rows of 32 lanes from LDS) and
// stores it with its dense active count; one block launched behind the
// evaluation (the split's extra block in the plain loop, k_finish in the
// guided loop) reduces the partials in a fixed order (reduce_partials) and
// takes the stop decision, while the split's blocks scan the counts for
// their offsets, so an iteration needs no host sync.  (A fused last-block
// ticket was measured first: the acq_rel atomic on one counter serialises
// at ~23 ns per block, 12 us at 512 blocks and ~190 us at 8192, against ~3
// us for the extra launch.)  Sums are carried as (value, rounding error)
// pairs built with TwoSum.
#pragma once

#include "common.hpp"

namespace cuhre {

__device__ __forceinline__ void two_sum(double a, double b, double& s, double& e)
{
  s = a + b;
  double bb = s - a;
  e = (a - (s - bb)) + (b - bb);
}

// (s1,c1) += (s2,c2)
__device__ __forceinline__ void dd_add(double& s1, double& c1, double s2, double c2)
{
  double s, e;
  two_sum(s1, s2, s, e);
  s1 = s;
  c1 = c1 + c2 + e;
}

__device__ __forceinline__ double shfl_xor_d(double v, int mask)
{
  return __shfl_xor(v, mask, kWave);
}

__device__ __forceinline__ void partial_zero(Partial& p)
{
  p.est_s = p.est_c = p.err_s = p.err_c = 0.0;
  p.fest_s = p.fest_c = p.ferr_s = p.ferr_c = 0.0;
  p.abs_all = p.abs_fin = p.aest_all = p.aest_fin = 0.0;
  p.nactive = p.nnonfinite = 0;
}

__device__ __forceinline__ void partial_add(Partial& a, const Partial& b)
{
  dd_add(a.est_s, a.est_c, b.est_s, b.est_c);
  dd_add(a.err_s, a.err_c, b.err_s, b.err_c);
  dd_add(a.fest_s, a.fest_c, b.fest_s, b.fest_c);
  dd_add(a.ferr_s, a.ferr_c, b.ferr_s, b.ferr_c);
  a.abs_all += b.abs_all;
  a.abs_fin += b.abs_fin;
  a.aest_all += b.aest_all;
  a.aest_fin += b.aest_fin;
  a.nactive += b.nactive;
  a.nnonfinite += b.nnonfinite;
}

__device__ __forceinline__ void partial_shfl_add(Partial& p, int mask)
{
  Partial q;
  q.est_s = shfl_xor_d(p.est_s, mask);
  q.est_c = shfl_xor_d(p.est_c, mask);
  q.err_s = shfl_xor_d(p.err_s, mask);
  q.err_c = shfl_xor_d(p.err_c, mask);
  q.fest_s = shfl_xor_d(p.fest_s, mask);
  q.fest_c = shfl_xor_d(p.fest_c, mask);
  q.ferr_s = shfl_xor_d(p.ferr_s, mask);
  q.ferr_c = shfl_xor_d(p.ferr_c, mask);
  q.abs_all = shfl_xor_d(p.abs_all, mask);
  q.abs_fin = shfl_xor_d(p.abs_fin, mask);
  q.aest_all = shfl_xor_d(p.aest_all, mask);
  q.aest_fin = shfl_xor_d(p.aest_fin, mask);
  q.nactive = (int64_t)__shfl_xor((long long)p.nactive, mask, kWave);
  q.nnonfinite = (int64_t)__shfl_xor((long long)p.nnonfinite, mask, kWave);
  // combine in lane order so both partners compute the same bits
  if ((__lane_id() & mask) == 0) {
    partial_add(p, q);
  } else {
    partial_add(q, p);
    p = q;
  }
}

// Region classification of the reference: active iff err > MaxErr(est)
// (quad/GPUquad/GPUQuadPhases.cu:63-73, MaxErr in quad/util/cudaUtil.h:218).
__device__ __forceinline__ bool region_active(double est, double err, double epsrel,
                                              double epsabs)
{
  return err > fmax(epsrel * fabs(est), epsabs);
}

// One region's contribution to a Partial, its classification given (act:
// still active; a region with a non-finite est or err counts in nnonfinite
// and in the finished sums).
__device__ __forceinline__ void partial_from_region_act(Partial& p, double est, double err, double ab, bool act)
{
  partial_zero(p);
  const bool finite = isfinite(est) && isfinite(err);
  p.est_s = est;
  p.err_s = err;
  p.abs_all = ab;
  p.aest_all = fabs(est);
  if (!act) {
    p.fest_s = est;
    p.ferr_s = err;
    p.abs_fin = ab;
    p.aest_fin = fabs(est);
  }
  p.nactive = act ? 1 : 0;
  p.nnonfinite = finite ? 0 : 1;
}

// Reduce one Partial per thread to thread 0 of a block of W waves.  Waves
// whose lanes hold no region (contributes == false, wave-uniform) skip the
// shuffles and add zeros; thread 0 adds the waves 0, wstep, 2 wstep, ...
// (the others are known to hold nothing).
__device__ __forceinline__ void block_reduce_partial(Partial& p, Partial* ldsw, int W,
                                                     bool contributes = true, int wstep = 1)
{
  if (contributes) {
#pragma unroll
    for (int mask = 1; mask < kWave; mask <<= 1) partial_shfl_add(p, mask);
  } else {
    partial_zero(p);
  }
  const int wave = threadIdx.x / kWave;
  if (__lane_id() == 0) ldsw[wave] = p;
  __syncthreads();
  if (threadIdx.x == 0) {
    Partial t = ldsw[0];
    for (int w = wstep; w < W; w += wstep) partial_add(t, ldsw[w]);
    p = t;
  }
}

// Sums of an evaluation block's regions (at most 256, in slots 0..nslots-1
// of LDS) into thread 0's Partial, for blocks of at least 256 threads.  The
// owner thread of each slot stores the region's est, err and volume *
// sum|w7 f| (zeros for a slot beyond the list) and whether it stays active;
// then eight rows of 32 threads add one field each (est, err, their
// finished parts, the two |.| sums and their finished parts): slots k,
// k+32, ... in order into 32 sums, combined pairwise (partners k xor 1, 2,
// 4, 8, 16).  The counts come from wavefront ballots.  A region counts as
// active when its est and err are finite and err > max(epsrel |est|,
// epsabs); one with a non-finite value counts in nnonfinite and, like every
// inactive region, in the finished sums.
struct RegionSlots {
  double est[kBlock], err[kBlock], ab[kBlock];
  uint8_t fin[kBlock];  // the region is finished (its values count in the finished sums)
};

__device__ __forceinline__ void block_reduce_regions(bool owner, int slot, int nslots, double est, double err,
                                                     double ab, bool valid, double epsrel, double epsabs, Partial& out)
{
  __shared__ RegionSlots L;
  __shared__ double s_f[8][2];
  __shared__ int64_t s_wact[kMaxWaves], s_wbad[kMaxWaves];
  const int wave = threadIdx.x / kWave, lane = threadIdx.x & (kWave - 1);
  const bool finite = isfinite(est) && isfinite(err);
  const bool act = finite && region_active(est, err, epsrel, epsabs);
  if (owner) {
    L.est[slot] = valid ? est : 0.0;
    L.err[slot] = valid ? err : 0.0;
    L.ab[slot] = valid ? ab : 0.0;
    L.fin[slot] = valid && !act;
  }
  const uint64_t ma = __ballot(owner && valid && act), mb = __ballot(owner && valid && !finite);
  if (lane == 0) {
    s_wact[wave] = __popcll(ma);
    s_wbad[wave] = __popcll(mb);
  }
  __syncthreads();
  if (threadIdx.x < 8 * 32) {  // whole wavefronts
    const int f = threadIdx.x >> 5, k = threadIdx.x & 31;
    double v = 0.0, c = 0.0;
    for (int j = k; j < nslots; j += 32) {
      const bool keep = (f != 2 && f != 3 && f != 5 && f != 7) || L.fin[j];
      const double e = L.est[j];
      double x = f == 0 || f == 2 ? e : f == 1 || f == 3 ? L.err[j] : f == 4 || f == 5 ? L.ab[j] : fabs(e);
      x = keep ? x : 0.0;
      if (f < 4) dd_add(v, c, x, 0.0);
      else v += x;
    }
#pragma unroll
    for (int m = 1; m < 32; m <<= 1) {
      const double vo = shfl_xor_d(v, m), co = shfl_xor_d(c, m);
      if (f < 4) dd_add(v, c, vo, co);
      else v += vo;
    }
    if (k == 0) {
      s_f[f][0] = v;
      s_f[f][1] = c;
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    Partial t;
    t.est_s = s_f[0][0];
    t.est_c = s_f[0][1];
    t.err_s = s_f[1][0];
    t.err_c = s_f[1][1];
    t.fest_s = s_f[2][0];
    t.fest_c = s_f[2][1];
    t.ferr_s = s_f[3][0];
    t.ferr_c = s_f[3][1];
    t.abs_all = s_f[4][0];
    t.abs_fin = s_f[5][0];
    t.aest_all = s_f[6][0];
    t.aest_fin = s_f[7][0];
    t.nactive = t.nnonfinite = 0;
    const int W = blockDim.x / kWave;
    for (int w = 0; w < W; ++w) {
      t.nactive += s_wact[w];
      t.nnonfinite += s_wbad[w];
    }
    out = t;
  }
}

// The iteration's state from its total q: the sums with the finished
// regions before it and the global stop decision.
__device__ __forceinline__ IterState iteration_state(const EvalArgs& A, const Partial& q)
{
  IterState st;
  st.tot = q;
  st.rG_s = A.fin_est_s;
  st.rG_c = A.fin_est_c;
  dd_add(st.rG_s, st.rG_c, q.est_s, q.est_c);
  st.eG_s = A.fin_err_s;
  st.eG_c = A.fin_err_c;
  dd_add(st.eG_s, st.eG_c, q.err_s, q.err_c);
  const double rG = st.rG_s + st.rG_c, eG = st.eG_s + st.eG_c;
  int stop = 0;
  if (q.nnonfinite > 0) stop = 4;
  else if (eG <= fmax(A.epsrel * fabs(rG), A.epsabs)) stop = 1;
  else if (q.nactive == 0) stop = 2;
  else if (q.nactive > A.max_active) stop = 3;
  else if (q.nactive > A.alloc_active) stop = 5;
  st.stop = stop;
  st.pad = 0;
  return st;
}

// One thread writes the state (and publishes it to the host when asked).
__device__ __forceinline__ void write_state(const EvalArgs& A, const IterState& st)
{
  *A.state = st;
  if (A.host_state != nullptr) {
    // publish to the host, which polls host_seq instead of synchronising
    *A.host_state = st;
    __hip_atomic_store(A.host_seq, A.seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
  }
}

__device__ __forceinline__ void write_iteration_state(const EvalArgs& A, const Partial& q)
{
  write_state(A, iteration_state(A, q));
}

// Launches up to this many blocks are finalised by one wavefront, larger
// ones in groups of 64 blocks.
constexpr uint32_t kWaveFinalize = 4 * kWave;

// Evaluation launches of up to this many blocks are finalised by the split
// in one level: every evaluation block also stores its active count, plus
// kBadFlag when it holds a non-finite value, into offsets[block], a dense
// array the split's blocks scan (a finalisation kernel later overwrites it
// with offsets).
constexpr uint32_t kOneLevel = 1024;
constexpr int64_t kBadFlag = 1ll << 40;
constexpr int64_t kCountMask = kBadFlag - 1;

// Called by every evaluation block after its thread 0 holds the block's
// Partial in p.
__device__ __forceinline__ void store_block_partial(const EvalArgs& A, const Partial& p)
{
  if (threadIdx.x == 0) {
    A.partials[blockIdx.x] = p;
    A.offsets[blockIdx.x] = p.nactive + (p.nnonfinite > 0 ? kBadFlag : 0);
  }
}

// Reduces the n partials src[0..n) with one block of T = 64 W threads
// (W <= 16 wavefronts): dst[b] = exclusive scan of the active counts (unless
// dst is null), the total lands in thread 0.  The counts are summed exactly
// over contiguous chunks.  The sums follow one order whatever T is: field
// by field, partials k, k+32, k+64, ... are added in order into 32 sums
// (k = 0..31), which are then combined pairwise (partners k xor 1, 2, 4, 8,
// 16 in a 32-lane row); the eight double fields go to 8 rows of 32
// threads, so the whole block works on short chains (loads issued 16 at a
// time) instead of shuffling the partial's 14 fields through a 64-lane
// tree.  Every thread returns (callers may barrier).
__device__ __forceinline__ void reduce_partials(const Partial* src, uint32_t n, int64_t* dst, Partial& total)
{
  constexpr int kFields = 8;  // est, err, fest, ferr (value + carry), abs_all, abs_fin, aest_all, aest_fin
  __shared__ double s_f[kFields][2];
  __shared__ int64_t s_cnt[kMaxWaves], s_bad[kMaxWaves];
  const uint32_t T = blockDim.x, W = T / kWave;
  const uint32_t lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
  const uint32_t chunk = (n + T - 1) / T;
  const uint32_t b0 = threadIdx.x * chunk < n ? threadIdx.x * chunk : n;
  const uint32_t b1 = b0 + chunk < n ? b0 + chunk : n;
  int64_t mine = 0, bad = 0;
  for (uint32_t b = b0; b < b1; ++b) {
    mine += src[b].nactive;
    bad += src[b].nnonfinite;
  }
  long long incl = (long long)mine, nbad = (long long)bad;
#pragma unroll
  for (int off = 1; off < kWave; off <<= 1) {
    const long long v = __shfl_up(incl, off, kWave);
    if ((int)lane >= off) incl += v;
  }
#pragma unroll
  for (int mask = 1; mask < kWave; mask <<= 1) nbad += __shfl_xor(nbad, mask, kWave);
  if (lane == kWave - 1) s_cnt[wave] = (int64_t)incl;
  if (lane == 0) s_bad[wave] = (int64_t)nbad;
  constexpr uint32_t kRows = 32, kBatch = 16;
  for (uint32_t base = 0; base < kFields * kRows; base += T) {
    const uint32_t idx = base + threadIdx.x;
    const uint32_t f = idx / kRows, k = idx % kRows;
    double v = 0.0, c = 0.0;
    if (f < kFields) {
      const bool dd = f < 4;  // uniform in a wavefront (two rows of one kind)
      const uint32_t off = dd ? 2 * f : 4 + f;
      for (uint32_t j0 = k; j0 < n; j0 += kBatch * kRows) {
        double a[kBatch], b[kBatch];
#pragma unroll
        for (uint32_t u = 0; u < kBatch; ++u) {
          const uint32_t j = j0 + u * kRows;
          a[u] = b[u] = 0.0;
          if (j < n) {
            const double* q = reinterpret_cast<const double*>(src + j) + off;
            a[u] = q[0];
            if (dd) b[u] = q[1];
          }
        }
#pragma unroll
        for (uint32_t u = 0; u < kBatch; ++u) {
          if (j0 + u * kRows < n) {
            if (dd) dd_add(v, c, a[u], b[u]);
            else v += a[u];
          }
        }
      }
    }
#pragma unroll
    for (int m = 1; m < (int)kRows; m <<= 1) {
      const double vo = shfl_xor_d(v, m), co = shfl_xor_d(c, m);
      // TwoSum's error term is exact, so both partners get the same bits
      if (f < 4) dd_add(v, c, vo, co);
      else v += vo;
    }
    if (f < kFields && k == 0) {
      s_f[f][0] = v;
      s_f[f][1] = c;
    }
  }
  __syncthreads();
  int64_t run = (int64_t)incl - mine;
  for (uint32_t w = 0; w < wave; ++w) run += s_cnt[w];
  if (dst != nullptr)
    for (uint32_t b = b0; b < b1; ++b) {
      dst[b] = run;
      run += src[b].nactive;
    }
  if (threadIdx.x == 0) {
    Partial t;
    t.est_s = s_f[0][0];
    t.est_c = s_f[0][1];
    t.err_s = s_f[1][0];
    t.err_c = s_f[1][1];
    t.fest_s = s_f[2][0];
    t.fest_c = s_f[2][1];
    t.ferr_s = s_f[3][0];
    t.ferr_c = s_f[3][1];
    t.abs_all = s_f[4][0];
    t.abs_fin = s_f[5][0];
    t.aest_all = s_f[6][0];
    t.aest_fin = s_f[7][0];
    t.nactive = t.nnonfinite = 0;
    for (uint32_t w = 0; w < W; ++w) {
      t.nactive += s_cnt[w];
      t.nnonfinite += s_bad[w];
    }
    total = t;
  }
}

// Body of k_finish (launches of at most kWaveFinalize blocks, one
// wavefront): the block partials straight to the state.
__device__ __forceinline__ void finish_iteration(const EvalArgs& A, uint32_t nb)
{
  Partial q;
  reduce_partials(A.partials, nb, A.offsets, q);
  if (threadIdx.x == 0) write_iteration_state(A, q);
}

// Larger launches, first level (k_finish_groups, one wavefront per group
// of 64 blocks): the group's sum and the offsets of its blocks within it.
// One block reading all partials of a large launch took ~40 us at 8k blocks
// (its load stream is serial per thread); in groups they are read in
// parallel.
// Queued ahead of the host (A.n_dev set, rpb regions per evaluation block):
// nothing to do when the list is empty (the loop stopped before it) or
// when its real length needs only one level (the split then scans the
// blocks' dense counts, which must stay in offsets[]).
__device__ __forceinline__ void finish_groups(const EvalArgs& A, uint32_t nb, int rpb = 0)
{
  if (A.n_dev != nullptr) {
    const int64_t n = *A.n_dev;
    if (n == 0 || (rpb > 0 && (n + rpb - 1) / rpb <= (int64_t)kOneLevel)) return;
  }
  const uint32_t lane = threadIdx.x;
  const uint32_t b = blockIdx.x * kWave + lane;
  Partial q;
  if (b < nb) q = A.partials[b];
  else partial_zero(q);
  const int64_t mine = q.nactive;
  long long incl = (long long)mine;
#pragma unroll
  for (int off = 1; off < kWave; off <<= 1) {
    const long long v = __shfl_up(incl, off, kWave);
    if ((int)lane >= off) incl += v;
  }
  if (b < nb) A.offsets[b] = (int64_t)incl - mine;
#pragma unroll
  for (int mask = 1; mask < kWave; mask <<= 1) partial_shfl_add(q, mask);
  if (lane == 0) A.gpart[blockIdx.x] = q;
}

// Second level (k_finish_top, one block): the G group sums to the state
// and the groups' base offsets (the split adds gbase[b/64] to offsets[b]).
__device__ __forceinline__ void finish_top(const EvalArgs& A, uint32_t G)
{
  Partial q;
  reduce_partials(A.gpart, G, A.gbase, q);
  if (threadIdx.x == 0) write_iteration_state(A, q);
}

// Queues k_finish for an evaluation launch of nb blocks.
void launch_finish(const EvalArgs& A, int64_t nb, hipStream_t st);

}  // namespace cuhre
