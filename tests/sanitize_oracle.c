/* Runs every entry of the CPU oracle on small problems under
 * AddressSanitizer and UndefinedBehaviorSanitizer (make -C oracle sanitize;
 * tests/test_sanitize.py): both rules in several dimensions, every family,
 * the breadth-first loop with its checks and finish policies, the two-phase
 * method, the heap, the closed forms, the grid and split, and the edge cases
 * (empty split, non-finite integrand, tiny budget and capacity).  Exits
 * non-zero on a failed check; a sanitizer report aborts it. */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "cuhre_oracle.h"

static int g_fail = 0;
#define CHECK(c)                                                          \
  do {                                                                    \
    if (!(c)) {                                                           \
      fprintf(stderr, "check failed: %s (line %d)\n", #c, __LINE__);      \
      g_fail = 1;                                                         \
    }                                                                     \
  } while (0)

static or_integrand make(int family, int d)
{
  or_integrand f;
  memset(&f, 0, sizeof f);
  f.family = family;
  f.dim = d;
  for (int i = 0; i < d; ++i) {
    f.a[i] = 1.5 + 0.7 * i;
    f.w[i] = 0.2 + 0.05 * i;
  }
  f.u = 0.3;
  f.exponent = d + 1;
  f.nterms = 2;
  f.coef[0] = 1.0;
  f.coef[1] = -0.5;
  f.powers[0][0] = 2;
  f.powers[1][d - 1] = 3;
  return f;
}

int main(void)
{
  static double table[16 * 16];
  for (int i = 0; i < 16 * 16; ++i) table[i] = sin(0.37 * i);
  const double lo8[8] = {0}, hi8[8] = {1, 1, 1, 1, 1, 1, 1, 1};
  /* region evaluation, both rules, every family, several dimensions */
  for (int d = 2; d <= 8; d += 3) {
    or_rule r7;
    or_rule9 r9;
    CHECK(or_rule_init(&r7, d) == 0 && or_rule9_init(&r9, d) == 0);
    double lo[8], wd[8], dv[8], sums[5];
    for (int m = 0; m < d; ++m) {
      lo[m] = 0.1 * m / d;
      wd[m] = 0.3;
    }
    for (int fam = 1; fam <= 7; ++fam) {
      or_integrand f = make(fam, d);
      double est, err, ab, i5, out[6];
      int ax;
      or_region_eval(&r7, &f, lo, wd, &est, &err, &ax, &ab, &i5, dv);
      CHECK(isfinite(est) && err >= 0 && ax >= 0 && ax < d);
      or_region_eval9(&r9, &f, lo, wd, &est, &err, &ax, &ab, dv, sums);
      CHECK(isfinite(est) && err >= 0 && ax >= 0 && ax < d);
      or_region_check(7, &f, lo, wd, out);
      or_region_check(9, &f, lo, wd, out);
      double rlo;
      (void)or_fq_residual(&f, lo, &rlo);
    }
    double e3[3];
    CHECK(or_r9_error(&r9, sums, e3) >= 0.0);
    or_rule_free(&r7);
  }
  /* loops on a 3-D Gaussian and the table integrand */
  or_integrand g = make(OR_GAUSSIAN, 3);
  const double lo3[3] = {0, 0, 0}, hi3[3] = {1, 1, 1};
  or_result r;
  or_check chk;
  for (int th = 1; th <= 2; ++th) {
    CHECK(or_integrate_bfs(&g, lo3, hi3, 1e-5, 0.0, 1000000, 100000, th, &r) == 0);
    CHECK(or_integrate_bfs_key(&g, lo3, hi3, 1e-5, 0.0, 1000000, 100000, th, 9, &r) == 0);
    CHECK(or_integrate_bfs_check(&g, lo3, hi3, 1e-5, 0.0, 1000000, 100000, th, 7, 1, &r, &chk) == 0);
    CHECK(fabs(r.integral - (chk.istar_hi + chk.istar_lo)) <= r.bound);
    CHECK(or_integrate_bfs_allot(&g, lo3, hi3, 1e-5, 0.0, 1000000, 100000, th, 0.5, &r) >= 0);
    CHECK(or_integrate_bfs_normalized(&g, lo3, hi3, 1e-5, 0.0, 1000000, 100000, th, 0.25, &r) >= 0);
    CHECK(or_integrate_two_phase(&g, lo3, hi3, 1e-6, 0.0, 9, 256, 16, th, &r) >= 0);
    CHECK(or_integrate_two_phase(&g, lo3, hi3, 1e-6, 0.0, 7, 64, 4, th, &r) >= 0);
  }
  CHECK(or_integrate_heap(&g, lo3, hi3, 1e-5, 0.0, 1000000, &r) == 0);
  or_integrand t = make(OR_TABLE6, 6);
  t.table = table;
  t.table_nx = 16;
  t.table_ny = 16;
  const double lo6[6] = {0, 0, 0.1, 0, 0, 0}, hi6[6] = {1, 1, 1, 1, 1, 1};
  CHECK(or_integrate_bfs(&t, lo6, hi6, 1e-2, 0.0, 2000000, 20000, 2, &r) >= 0);
  double ph, pl;
  CHECK(or_exact(&t, lo6, hi6, &ph, &pl) == 0);
  for (int fam = 1; fam <= 7; ++fam) {
    or_integrand f = make(fam, 4);
    const double lo4[4] = {0, 0, 0, 0}, hi4[4] = {1, 1, 1, 1};
    (void)or_exact(&f, lo4, hi4, &ph, &pl);
    double out[5];
    or_phase2_region(fam == 4 ? 7 : 9, &f, lo4, hi4, 1e-4, 0.0, 8, out);
    CHECK(out[2] >= 2);
  }
  /* edge cases: tiny budget, tiny capacity, non-finite integrand, empty split */
  CHECK(or_integrate_bfs(&g, lo3, hi3, 1e-12, 0.0, 64 * 33, 100000, 1, &r) == 1);
  CHECK(or_integrate_bfs(&g, lo3, hi3, 1e-12, 0.0, 1000000, 70, 1, &r) == 1);
  or_integrand bad = make(OR_CORNER_PEAK, 3);
  for (int i = 0; i < 3; ++i) bad.a[i] = -1.0;
  bad.exponent = 700;
  CHECK(or_integrate_bfs(&bad, lo3, hi3, 1e-6, 0.0, 1000000, 100000, 1, &r) == -2);
  const int64_t n8 = or_initial_grid(8, lo8, hi8, NULL, NULL);
  double* L = malloc(sizeof(double) * n8 * 8);
  double* Wd = malloc(sizeof(double) * n8 * 8);
  double* L2 = malloc(sizeof(double) * 2 * n8 * 8);
  double* W2 = malloc(sizeof(double) * 2 * n8 * 8);
  int32_t* act = calloc(n8, sizeof(int32_t));
  int32_t* axs = calloc(n8, sizeof(int32_t));
  or_initial_grid(8, lo8, hi8, L, Wd);
  CHECK(or_split(8, n8, L, Wd, act, axs, L2, W2) == 0);
  for (int64_t i = 0; i < n8; i += 3) {
    act[i] = 1;
    axs[i] = (int32_t)(i % 8);
  }
  CHECK(or_split(8, n8, L, Wd, act, axs, L2, W2) == (n8 + 2) / 3);
  free(L);
  free(Wd);
  free(L2);
  free(W2);
  free(act);
  free(axs);
  CHECK(or_fp64_bound(401, 8, 1000, 1.0, 1.0, 1.0) > 0.0);
  printf(g_fail ? "oracle checks FAILED\n" : "oracle checks ok\n");
  return g_fail;
}
