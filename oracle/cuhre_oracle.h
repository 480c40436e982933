/*
 * cuhre_oracle.h -- CPU FP64 restatement of the adaptive cubature path.
 *
 * TEST INFRASTRUCTURE ONLY.  This code is the checker: only tests/, the
 * smoke() entry and bench.py's CPU-baseline leg may load it.  The product
 * library never links or calls it.
 *
 * What it restates (reference file:line, paths relative to the reference):
 *   - per-region rule evaluation, fourth-difference axis choice and its tie
 *     rules: quad/GPUquad/GPUQuadSample.cu:99-220 (the rule itself is the
 *     Genz-Malik degree-7/5 pair, Genz & Malik 1980, not the reference's
 *     hard-wired degree-9 rule);
 *   - region classification err > max(epsrel*|est|, epsabs):
 *     quad/GPUquad/GPUQuadPhases.cu:60-78, quad/util/cudaUtil.h:218;
 *   - the breadth-first loop, global error test and counters:
 *     quad/GPUquad/GPUKernelquad.cu:507-668;
 *   - initial k^d grid: quad/GPUquad/GPUKernelquad.cu:23-61, 300-316;
 *   - bisection: quad/GPUquad/GPUKernelquad.cu:97-149;
 *   - final flag: quad/GPUquad/GPUquad.h:713-714.
 * Closed forms of every integrand family are evaluated in __float128.
 */
#ifndef CUHRE_ORACLE_H
#define CUHRE_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define OR_MAXDIM 16
#define OR_MAXTERMS 8

enum {
  OR_GAUSSIAN = 1,      /* exp(-sum a_i^2 (x_i-w_i)^2)                     */
  OR_PRODUCT_PEAK = 2,  /* prod 1/(a_i^-2 + (x_i-w_i)^2)                   */
  OR_OSCILLATORY = 3,   /* cos(2 pi u + sum a_i x_i)                       */
  OR_CORNER_PEAK = 4,   /* (1 + sum a_i x_i)^-k, k = exponent              */
  OR_CONTINUOUS = 5,    /* exp(-sum a_i |x_i - w_i|)                       */
  OR_POLYNOMIAL = 6,    /* sum_t c_t prod_i x_i^e_ti                       */
  OR_COSINE_SUM = 7,    /* -(1/2) sum cos(10 x_i)/0.054402111088937        */
  OR_TABLE6 = 8,        /* B(x1,x2) x3^-1.5 exp(-(x4-.5)^2/.02) (1+x5 x6)  */
  OR_LC_LT_T = 9        /* the reference's LC_LT_t in x = (lt, zt, lc)      */
};

typedef struct {
  int32_t family;
  int32_t dim;
  double a[OR_MAXDIM];
  double w[OR_MAXDIM];
  double u;
  int32_t exponent;
  int32_t nterms;
  double coef[OR_MAXTERMS];
  int32_t powers[OR_MAXTERMS][OR_MAXDIM];
  const double* table; /* TABLE6: nx*ny values, table[i*ny + j] = B(i/(nx-1), j/(ny-1));
                          LC_LT_T: lt nodes (nx), zt nodes (ny), then tau, mu, sigma,
                          fmsk, fprj, each [j*nx + i], all taken as float */
  int32_t table_nx, table_ny;
} or_integrand;

/* Genz-Malik degree-7 rule with embedded degree-5 rule on [-1,1]^d, weights
 * normalised to sum to 1 (so est = volume * sum w_k f(x_k)). */
typedef struct {
  int dim, npts;
  double* g;   /* npts*dim generator offsets, point k row g[k*dim ..] */
  double* w7;  /* npts */
  double* w5;  /* npts (0 for the 2^d corner points) */
} or_rule;

int or_rule_init(or_rule* r, int dim);
void or_rule_free(or_rule* r);
int or_gm_npoints(int dim);

double or_f(const or_integrand* f, const double* x);

/* One region.  lower/width are the region's d-vectors.  dvec (may be NULL)
 * receives the d fourth differences; abssum receives volume*sum|w7 f|. */
void or_region_eval(const or_rule* r, const or_integrand* f,
                    const double* lower, const double* width,
                    double* est, double* err, int* axis, double* abssum,
                    double* i5, double* dvec);

/* The reference's degree-9 rule (KEY 9): Cuba's Cuhre rule of degree 9 with
 * four null rules and nine generator sets, as the reference hard-wires it
 * (quad/GPUquad/GPUQuadRule.cu:30-226, quad/quad.h:80-97); error from the
 * null rules with errcoeff {5, 1, 5} (quad/GPUquad/GPUQuadSample.cu:195-216).
 * Points: centre; +-t_k e_i for four radii; (+-t1, +-t1) on pairs i<j;
 * (+-t1, +-t2) on ordered pairs i != j; (+-t1)^3 on triples i<j<k; the 2^d
 * corners at +-t5.  t_k are in half-width units.  Weights sum to 1 in rule
 * 0 (est = volume * sum_k w_k f(x_k)). */
#define OR_R9_SETS 9
#define OR_R9_RULES 5
typedef struct {
  int dim, npts;
  double t[5];
  double w[OR_R9_SETS][OR_R9_RULES];
  double scale[OR_R9_SETS][OR_R9_RULES], norm[OR_R9_SETS][OR_R9_RULES];
  int count[OR_R9_SETS];
} or_rule9;

int or_rule9_init(or_rule9* r, int dim);
int or_r9_npoints(int dim);
/* One region under the degree-9 rule; sums[5] (may be NULL) receives the
 * five rule sums (rule 0 = the estimate's, before the volume factor). */
void or_region_eval9(const or_rule9* r, const or_integrand* f,
                     const double* lower, const double* width,
                     double* est, double* err, int* axis, double* abssum,
                     double* dvec, double* sums);

/* The degree-9 rule's error (before the volume factor) from its five rule
 * sums; e_out (may be NULL) receives e1, e2, e3. */
double or_r9_error(const or_rule9* r, const double* sums, double* e_out);

typedef struct {
  double integral, error, bound;
  int64_t nregions, neval;
  int32_t fail, iterations;
  int64_t peak_regions;
  int64_t nfinished;
} or_result;

/* Breadth-first loop (the plain method).  nthreads>1 uses OpenMP over the
 * regions of an iteration.  Returns 0 (converged), 1 (not converged), <0
 * error (-2 nonfinite integrand value, -3 allocation, -1 bad argument).
 * bound = or_fp64_bound over the regions the result sums. */
int or_integrate_bfs(const or_integrand* f, const double* lows,
                     const double* highs, double epsrel, double epsabs,
                     int64_t maxeval, int64_t capacity, int nthreads,
                     or_result* out);
/* The same loop with the rule chosen by key (0 or 7: GM7/5, 9: degree 9). */
int or_integrate_bfs_key(const or_integrand* f, const double* lows,
                         const double* highs, double epsrel, double epsabs,
                         int64_t maxeval, int64_t capacity, int nthreads,
                         int key, or_result* out);

/* FP64 rounding bound of a result over n contributing regions of a rule of
 * p points in d dimensions (the same formula the product library reports):
 *   g_K A (1 + 4 (K + n) u) + 2 u |I| + 2 n^2 u^2 / (1 - 2 n u) E,
 *   K = p + d + 2, g_K = K u / (1 - K u), u = 2^-53,
 * A = sum over the regions of volume * sum_k |w_k f_k| (abssum), E = sum
 * |est|.  It bounds |I - sum_r V_r sum_k w_k f_k| for the FP64 integrand
 * values f_k the loop obtained: the rule sums (each term passes through at
 * most p roundings, g_p), the volume and the estimate's product (g_d), and
 * the TwoSum-carried sum over regions (u |T| + g_{2n}^2-type compensation
 * term; any summation tree of depth <= n).  The integrand's own rounding
 * is the functor's, not the engine's, and is not included. */
double or_fp64_bound(int p, int d, int64_t n, double A, double E, double I);

/* Rounding check of one breadth-first run, evaluated beside the FP64 loop
 * in __float128 (w_k * f_k of two doubles is exact in 113 bits):
 *   istar = sum over the contributing regions of V_r sum_k w_k fhat_k, the
 *           exact-arithmetic value of the loop's own FP64 integrand values;
 *   irule = the same sums with the nodes c + h g_k and the integrand in
 *           __float128 (families with a quad restatement; have_rule = 0
 *           otherwise and for key 9): the rule's value free of FP64.
 * Each as a double pair (hi + lo).  abs_total, aest_total, ncontrib are
 * the bound's A, E and n. */
typedef struct {
  double istar_hi, istar_lo, irule_hi, irule_lo;
  double abs_total, aest_total;
  int64_t ncontrib;
  int32_t have_rule, pad;
} or_check;

int or_integrate_bfs_check(const or_integrand* f, const double* lows,
                           const double* highs, double epsrel, double epsabs,
                           int64_t maxeval, int64_t capacity, int nthreads,
                           int key, int want_rule, or_result* out, or_check* chk);

/* The breadth-first loop (GM7/5) with the shared-allotment finish policy:
 * a region is active iff err > max(epsrel ((1 - allot) |est| + allot |I|
 * vrel), epsabs), I the iteration's estimate (finished so far plus the
 * iteration's regions), vrel the region's share of the box volume;
 * allot = 0 is the plain loop.  The Genz-Malik global test is unchanged. */
int or_integrate_bfs_allot(const or_integrand* f, const double* lows, const double* highs, double epsrel,
                           double epsabs, int64_t maxeval, int64_t capacity, int nthreads, double allot,
                           or_result* out);

/* The same loop with the normalized allotment: active iff err >
 * max(epsrel |I| ((1 - allot) |est| / E + allot vrel), epsabs), E = sum |est|
 * over the finished regions and the iteration's; the tolerances add up to
 * epsrel |I| whatever the signs (for f >= 0 and allot = 0 the plain test). */
int or_integrate_bfs_normalized(const or_integrand* f, const double* lows, const double* highs, double epsrel,
                                double epsabs, int64_t maxeval, int64_t capacity, int nthreads, double allot,
                                or_result* out);

/* One region with its __float128 shadow: out[0..5] = est, abssum,
 * istar (hi, lo), irule (hi, lo; NaN where no quad restatement). */
void or_region_check(int key, const or_integrand* f, const double* lower,
                     const double* width, double* out);

/* The integrand in __float128 at a double point (NaN where the family has
 * no quad restatement: the tabulated ones). */
double or_fq_residual(const or_integrand* f, const double* x, double* lo);

/* The reference's two-phase method (GPUquad.h:611-727): Phase 1 is the
 * breadth-first loop (at most 100 iterations) until the global test holds,
 * no region is active, or the split list reaches phase1_max regions
 * (GPUKernelquad.cu:507-668); Phase 2 then refines every region of that
 * list on its own (GPUQuadPhases.cu:422-536): a max-error heap, up to
 * pool_max bisections, until the region's own sum meets
 * max(epsrel |RESULT|, epsabs), children's errors refined from the parent
 * difference as Cuba does (GPUQuadPhases.cu:498-515).  Ties in the heap go
 * to the region created first.  fail = number of Phase-2 regions that did
 * not meet their test, 0 when the global test holds (GPUquad.h:713-714). */
int or_integrate_two_phase(const or_integrand* f, const double* lows, const double* highs,
                           double epsrel, double epsabs, int key, int64_t phase1_max,
                           int pool_max, int nthreads, or_result* out);

/* Phase 2 of one region (GPUQuadPhases.cu:422-536) as the two-phase method
 * runs it: out[0..4] = result, error, pool size, still active, zeroed. */
void or_phase2_region(int key, const or_integrand* f, const double* lower, const double* width, double epsrel,
                      double epsabs, int pool_max, double* out);

/* Cuhre-style global adaptive (heap of regions, always bisect the region
 * with the largest error), same GM7/5 rule: the CPU baseline flavour (i). */
int or_integrate_heap(const or_integrand* f, const double* lows,
                      const double* highs, double epsrel, double epsabs,
                      int64_t maxeval, or_result* out);

/* Exact integral over the box, evaluated in __float128, returned as double
 * (hi part) plus the residual (lo part). */
int or_exact(const or_integrand* f, const double* lows, const double* highs,
             double* hi, double* lo);

/* Splitting step on a host list (for compaction/bisection checks):
 * given n regions SoA (lower[m*n+r], width[m*n+r]) and flags/axes, writes
 * 2*A children (child 0 at 2j, child 1 at 2j+1, SoA with stride 2A).
 * Returns A. */
int64_t or_split(int dim, int64_t n, const double* lower, const double* width,
                 const int32_t* active, const int32_t* axis,
                 double* out_lower, double* out_width);

/* Initial grid: k^d regions, k = 4 (d<=4), 2 (5<=d<=10), 1 (d>10).
 * Writes SoA with stride n; returns n. lower/width may be NULL to query n. */
int64_t or_initial_grid(int dim, const double* lows, const double* highs,
                        double* lower, double* width);

#ifdef __cplusplus
}
#endif
#endif
