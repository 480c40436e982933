// Host check of the library's device-free logic, built with AddressSanitizer
// and UndefinedBehaviorSanitizer (make -C old-gpuintegration_amd sanitize)
// and run by tests/test_sanitize.py: the rule tables of both rules in every
// dimension, the flop model, the rounding bound, the initial grid and static
// workspace arithmetic, the training order's permutation, the argument
// checks with their error codes and messages.  Exits non-zero on the first failed check; a sanitizer report
// aborts the program.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>
#include <vector>

#include "context.hpp"

using namespace cuhre;

static int g_fail = 0;

#define CHECK(cond)                                                       \
  do {                                                                    \
    if (!(cond)) {                                                        \
      std::fprintf(stderr, "check failed: %s (line %d)\n", #cond, __LINE__); \
      g_fail = 1;                                                         \
    }                                                                     \
  } while (0)

static cuhre_integrand integrand(int family, int dim)
{
  cuhre_integrand f;
  std::memset(&f, 0, sizeof f);
  f.family = family;
  f.dim = dim;
  for (int i = 0; i < CUHRE_MAXDIM; ++i) {
    f.a[i] = 1.0 + 0.5 * i;
    f.w[i] = 0.1 * i;
  }
  f.exponent = dim + 1;
  return f;
}

int main()
{
  // rules: point counts and normalised weights, both keys, every dimension
  for (int d = 2; d <= 10; ++d) {
    const RuleConst R7 = make_rule(d, 7);
    CHECK(R7.npts == gm_points(d) && cuhre_points_per_region(d, 0) == R7.npts);
    const double n = d, corners = (double)(1 << d);
    const double w7 = R7.w7c + 2 * n * (R7.w7a2 + R7.w7a3) + 2 * n * (n - 1) * R7.w7p + corners * R7.w7v;
    const double w5 = R7.w5c + 2 * n * (R7.w5a2 + R7.w5a3) + 2 * n * (n - 1) * R7.w5p;
    CHECK(std::fabs(w7 - 1.0) < 1e-13 && std::fabs(w5 - 1.0) < 1e-13);
    const RuleConst R9 = make_rule(d, 9);
    CHECK(R9.key == 9 && R9.npts == r9_points(d) && cuhre_points_per_region(d, 9) == R9.npts);
    for (int q = 1; q < 4; ++q)
      for (int t = 0; t < 9; ++t) CHECK(std::isfinite(R9.r9.scale[t][q]) && R9.r9.norm[t][q] > 0.0);
    // flop model: positive and growing with the rule
    for (int fam = 1; fam <= 7; ++fam) {
      const cuhre_integrand f = integrand(fam, d);
      CHECK(region_flops(&f, d, 7) > 0.0 && region_flops(&f, d, 9) > region_flops(&f, d, 7));
    }
    CHECK(initial_grid_count(d) == (d <= 4 ? (int64_t)std::pow(4.0, d) : (int64_t)1 << d));
    // static workspace: what it holds fits, one more 256 regions does not
    const int64_t cap = scratch_capacity(d);
    CHECK(cap > 0 && cap % 256 == 0 && cuhre_noalloc_capacity(d) == cap);
    CHECK(scratch_list_bytes(cap, d) <= kScratchBytes - kScratchTableBytes);
    CHECK(scratch_list_bytes(cap + 256, d) > kScratchBytes - kScratchTableBytes);
  }
  CHECK(cuhre_points_per_region(1, 0) == -1 && cuhre_points_per_region(17, 7) == -1);
  CHECK(cuhre_points_per_region(5, 3) == -1 && cuhre_noalloc_capacity(1) == -1 && cuhre_noalloc_capacity(11) == -1);
  CHECK(ipow_flops(0, 3) == 0.0 && ipow_flops(8, 7) == 3.0 && ipow_flops(8, 5) == 4.0);

  // rounding bound: monotone in its inputs, finite at large counts
  const double b0 = fp64_rounding_bound(401, 8, 1000, 1.0, 1.0, 1.0);
  CHECK(b0 > 0.0 && fp64_rounding_bound(401, 8, 2000, 1.0, 1.0, 1.0) > b0);
  CHECK(fp64_rounding_bound(401, 8, 1000, 2.0, 1.0, 1.0) > b0);
  CHECK(std::isfinite(fp64_rounding_bound(1105, 8, (int64_t)1 << 40, 1e10, 1e10, 1e10)));

  // argument checks: codes and messages
  const cuhre_integrand f3 = integrand(CUHRE_GAUSSIAN, 3);
  const double lo[3] = {0, 0, 0}, hi[3] = {1, 1, 1}, bad[3] = {0, 1, 0};
  const double inf = std::numeric_limits<double>::infinity(), nan = std::numeric_limits<double>::quiet_NaN();
  const double hinf[3] = {1, inf, 1};
  CHECK(check_integrand_args(&f3, 3, 0) == 0 && check_integrand_args(&f3, 3, 9) == 0);
  CHECK(check_integrand_args(nullptr, 3, 0) == CUHRE_ERR_ARG);
  CHECK(std::strstr(cuhre_last_error(), "null integrand") != nullptr);
  CHECK(check_integrand_args(&f3, 1, 0) == CUHRE_ERR_UNSUPPORTED);
  CHECK(check_integrand_args(&f3, 11, 0) == CUHRE_ERR_UNSUPPORTED);
  CHECK(check_integrand_args(&f3, 4, 0) == CUHRE_ERR_ARG);
  CHECK(check_integrand_args(&f3, 3, 5) == CUHRE_ERR_UNSUPPORTED);
  CHECK(std::strstr(cuhre_last_error(), "rule key 5") != nullptr);
  CHECK(check_bounds(3, 1e-6, 0.0, lo, hi) == 0);
  CHECK(check_bounds(3, 1e-6, 0.0, nullptr, hi) == CUHRE_ERR_ARG);
  CHECK(check_bounds(3, -1.0, 0.0, lo, hi) == CUHRE_ERR_ARG);
  CHECK(check_bounds(3, nan, 0.0, lo, hi) == CUHRE_ERR_ARG);
  CHECK(check_bounds(3, 1e-6, nan, lo, hi) == CUHRE_ERR_ARG);
  CHECK(check_bounds(3, 1e-6, 0.0, lo, bad) == CUHRE_ERR_ARG);
  CHECK(std::strstr(cuhre_last_error(), "axis 0") != nullptr);
  CHECK(check_bounds(3, 1e-6, 0.0, lo, hinf) == CUHRE_ERR_ARG);
  // a long message is cut at the buffer, not overrun
  char longmsg[2000];
  std::memset(longmsg, 'x', sizeof longmsg - 1);
  longmsg[sizeof longmsg - 1] = 0;
  CHECK(set_error(CUHRE_ERR_DEVICE, longmsg) == CUHRE_ERR_DEVICE && std::strlen(cuhre_last_error()) < 600);

  // training order: a permutation of [0, n) for every n, and the values the
  // Python restatement (oracle/mlp_oracle.py) gives
  for (const uint64_t n : {1ull, 2ull, 3ull, 64ull, 1000ull, 65537ull}) {
    const Shuffle P = make_shuffle(n, n * 977);
    std::vector<char> seen(n, 0);
    for (uint64_t i = 0; i < n; ++i) {
      const uint64_t v = shuffle_index(P, i);
      CHECK(v < n && !seen[v]);
      if (v < n) seen[v] = 1;
    }
  }
  {
    const Shuffle P = make_shuffle(1000, 12345);
    const uint64_t want[6] = {195, 337, 220, 19, 681, 514};
    for (int i = 0; i < 6; ++i) CHECK(shuffle_index(P, (uint64_t)i) == want[i]);
  }

  std::printf(g_fail ? "host checks FAILED\n" : "host checks ok\n");
  return g_fail;
}
