// common.hpp -- definitions shared by the host engine and the gfx950 kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "cuhre.h"

namespace cuhre {

constexpr int kMaxDim = CUHRE_MAXDIM;
constexpr int kBlock = 256;  // threads per block in every kernel
constexpr int kWave = 64;    // CDNA wavefront
constexpr int kMaxWaves = 16;  // largest block: 1024 threads

// Device copy of an integrand: the C-ABI description plus constants derived
// on the host once per call (1/a^2 for the product peak, 2 pi u for the
// oscillatory family) so every evaluation point reuses them.
struct DevIntegrand {
  int32_t family;
  int32_t dim;
  int32_t exponent;
  int32_t nterms;
  double a[kMaxDim];
  double w[kMaxDim];
  double b[kMaxDim];  // product peak: 1/(a*a)
  double s0;          // oscillatory: 2*pi*u
  double coef[CUHRE_MAXTERMS];
  int8_t powers[CUHRE_MAXTERMS][kMaxDim];
  const double* table;  // device pointer (TABLE6)
  const float* ftab;    // device pointer (LC_LT_T): nodes, then five fields
  int32_t nx, ny;
};

// Genz-Malik degree-7/5 rule constants, weights normalised to sum to one
// (est = volume * sum w f).  Built on the host by rule.cpp.
// The reference's degree-9 rule (KEY 9): nine generator sets, a basic rule
// and four null rules (weights per set and rule), radii t[] in half-width
// units, and the null-rule combination constants.
struct Rule9Const {
  double t[5];
  double w[9][5];
  double scale[9][5], norm[9][5];
  double ratio;  // (t2/t1)^2, fourth differences
};

struct RuleConst {
  double l2, l3, l4, l5;
  double w7c, w7a2, w7a3, w7p, w7v;  // centre, axis l2, axis l3, pair l4, corner l5
  double w5c, w5a2, w5a3, w5p;       // embedded degree-5 rule (no corners)
  double ratio;                      // (l2/l3)^2 = 1/7
  int32_t npts;
  int32_t key;                       // 7: the rule above; 9: r9
  Rule9Const r9;
};

// Per-block partial sums written by the evaluation kernel.  The *_s/*_c
// pairs are value + carried rounding error (TwoSum), so the final sums are
// accurate to about one ulp whatever the block count.
struct Partial {
  double est_s, est_c, err_s, err_c;      // all regions of the iteration
  double fest_s, fest_c, ferr_s, ferr_c;  // regions retired as finished
  double abs_all, abs_fin;                // sum volume*sum|w7 f|
  double aest_all, aest_fin;              // sum |est|
  int64_t nactive, nnonfinite;
};

// Iteration outcome, written by the last evaluation block.
struct IterState {
  Partial tot;        // sums over the iteration's regions
  double rG_s, rG_c;  // finished-so-far + all regions of this iteration
  double eG_s, eG_c;
  int32_t stop;  // 1 converged, 2 no active region, 3 budget, 4 nonfinite, 5 grow
  int32_t pad;
};

struct EvalArgs {
  const double* lower;  // SoA, lower[m*ld + r]
  const double* width;
  int64_t ld;
  int64_t n;
  double epsrel, epsabs;          // global test
  double loc_epsrel, loc_epsabs;  // per-region test (active iff err > max(loc_epsrel*|est|, loc_epsabs))
  double* est;
  double* err;
  int8_t* axis;
  double* abssum;  // optional
  double* i5;      // optional
  double* dvec;    // optional, dvec[m*ld + r]
  Partial* partials;
  int64_t* offsets;  // exclusive scan of active counts per block
  Partial* gpart;    // per group of 64 blocks (launches of many blocks)
  int64_t* gbase;
  IterState* state;
  double fin_est_s, fin_est_c, fin_err_s, fin_err_c;
  int64_t max_active;    // stop (eval budget / capacity) above this count
  int64_t alloc_active;  // grow the list buffers above this count
  int32_t finalize;    // 0: per-region outputs only (no partials/state)
  uint32_t seq;        // published with the state when host_state is set
  IterState* host_state;  // optional: coherent host copy of *state
  uint32_t* host_seq;     // written (system-scope release) after host_state
  const int64_t* n_dev;   // when set: the list length, read on the device (n is then
                          // an upper bound: the launch of a queued-ahead iteration)
  uint8_t* act_out;       // optional (half-wavefront form): per block of 8, its active
                          // regions in order, (position << 4) | split axis
  // host side only: when set, the evaluation kernel's own start and end are
  // recorded in these events (hipExtLaunchKernel), so a timed launch does
  // not include the gap before it reaches the GPU
  hipEvent_t prof_start, prof_stop;
};

// Device-side loop state of a list parity when iterations are queued ahead
// of the host: the list's length and the finished-region sums before it.
struct SpecSlot {
  int64_t n, pad;
  double fin_s, fin_c, fer_s, fer_c;  // est and err of the finished regions (value, carry)
};

// What the split of a queued-ahead iteration needs on the device: its
// list's slot, the next list's slot, the evaluation budget left before the
// iteration (maxeval minus the evaluations of earlier iterations) and the
// regions per evaluation block (for the number of block partials).
struct SpecArgs {
  const SpecSlot* in;  // nullptr: not queued ahead
  SpecSlot* out;
  int64_t left0, cap2;  // budget before the iteration; capacity / 2
  int64_t npts;
  int32_t rpb, pad;
};

// A half-wavefront evaluation that first does the previous iteration's
// split for its own regions (k_eval_fused): the previous iteration's
// finalisation arguments (R: its block partials in R.partials, its dense
// active counts in R.offsets, the state to publish; P: its queued-ahead
// arguments, P.in null when its length is R.n) and its list (bounds, and
// per block of 8 regions the list of its active ones, act_out above).
struct FuseArgs {
  EvalArgs R;
  SpecArgs P;
  const double* plo;
  const double* pwd;
  const uint8_t* pact;
  double* clo;  // this iteration's list (the children's bounds are written there)
  double* cwd;
};

// FP64 rounding bound of a result that sums n regions of a p-point rule in
// d dimensions: it bounds |I - sum_r V_r sum_k w_k f_k| for the FP64
// integrand values f_k the kernels obtained (the integrand's own rounding
// belongs to the integrand).  A = sum of the regions' V sum |w f| (abssum),
// E = sum |est|.  Standard model, u = 2^-53:
//   rule sums: every term passes through at most p roundings (an FMA chain,
//     or segment chains merged by adds), so |s^ - s| <= g_p sum |w f|;
//   est = fl(V^ s^) with V^ the product of d widths: g_{p+d} V sum |w f|;
//   K = p + d + 2 and the factor 1 + 4(K + n)u absorb the rounding of the
//     abssums themselves and of their plain FP64 total;
//   regions: (value, TwoSum error) pairs over any summation tree of depth
//     <= n: u|T| (<= 2u|I|) plus the compensation's own rounding,
//     g_{2n} * n u * E = 2 n^2 u^2 / (1 - 2nu) E.
// (Higham, Accuracy and Stability of Numerical Algorithms, 2nd ed., 3.1 and
// 4.2; Ogita, Rump and Oishi 2005, Sum2.)
__host__ __device__ inline double fp64_rounding_bound(int64_t p, int d, int64_t n, double A, double E, double I)
{
  const double u = 0x1p-53;
  const double K = (double)p + (double)d + 2.0;
  const double gK = K * u / (1.0 - K * u);
  const double nn = (double)n;
  return gK * A * (1.0 + 4.0 * (K + nn) * u) + 2.0 * u * fabs(I) + 2.0 * nn * nn * u * u / (1.0 - 2.0 * nn * u) * E;
}

// Launches the rule evaluation of A.n regions with S point segments.
using EvalLauncher = void (*)(const EvalArgs&, const DevIntegrand&, const RuleConst&, int S,
                              hipStream_t);

// Returns the launcher for (family, dim, rule key 0/7 or 9) or nullptr.
EvalLauncher find_eval_launcher(int family, int dim, int key = 7);

// Launches k_eval_fused over at most A.n regions (the degree-7/5 rule);
// nullptr where the family has no such kernel.
using FusedLauncher = void (*)(const EvalArgs&, const FuseArgs&, const DevIntegrand&, const RuleConst&, hipStream_t);
FusedLauncher find_fused_launcher(int family, int dim);

// A pseudo-random permutation of [0, n) for the order of training samples,
// computed per row with no stored order: a bijection of [0, 2^bits) built
// from four rounds of a key addition, a multiplication by an odd constant
// and a xor-shift (each invertible modulo 2^bits), walked along its cycles
// until it lands below n.  The walk ends (the cycle through a start below n
// comes back to it), and the first value below n after each start is itself
// a bijection of [0, n) (cycle walking, Black and Rogaway 2002).
struct Shuffle {
  uint64_t n, mask;
  uint64_t key[4];
  int shift;
};

__host__ __device__ inline uint64_t shuffle_index(const Shuffle& P, uint64_t x)
{
  const auto round = [&](uint64_t k) {
    x = (x + k) & P.mask;
    x = (x * 0xD1342543DE82EF95ull) & P.mask;
    x ^= x >> P.shift;
  };
  do {
    round(P.key[0]);
    round(P.key[1]);
    round(P.key[2]);
    round(P.key[3]);
  } while (x >= P.n);
  return x;
}

// The permutation of [0, n) (n >= 1) drawn from `seed`.
inline Shuffle make_shuffle(uint64_t n, uint64_t seed)
{
  Shuffle P;
  int bits = 1;
  while (bits < 63 && (1ull << bits) < n) ++bits;
  P.n = n;
  P.mask = (1ull << bits) - 1;
  P.shift = (bits + 1) / 2;
  uint64_t s = seed;
  for (int i = 0; i < 4; ++i) {
    s += 0x9E3779B97F4A7C15ull;
    uint64_t z = s;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    P.key[i] = z ^ (z >> 31);
  }
  return P;
}

}  // namespace cuhre
