/*
 * cuhre_oracle.c -- CPU FP64 restatement of the adaptive cubature path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker).  The product library never links
 * this file; tests/, smoke() and bench.py's CPU-baseline leg load it.
 *
 * Every function below cites the reference lines whose semantics it follows.
 * Arithmetic that the GPU kernels also perform is written with explicit fma()
 * where the GPU contracts, so that per-region values agree to a few ulps.
 */
#include "cuhre_oracle.h"

#include <float.h>
#include <math.h>
#include <quadmath.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------------ */
/* Genz-Malik degree-7 / degree-5 rule (Genz & Malik 1980).                  */
/* Point order: centre; for each axis i: +l2 e_i, -l2 e_i, +l3 e_i, -l3 e_i; */
/* for each pair i<j: (+,+), (+,-), (-,+), (-,-) at l4; then the 2^d corners */
/* with coordinate m negative iff bit m of the corner index is set.          */
/* ------------------------------------------------------------------------ */

int or_gm_npoints(int d) { return (1 << d) + 2 * d * d + 2 * d + 1; }

int or_rule_init(or_rule* r, int d)
{
  if (d < 2 || d > OR_MAXDIM) return -1;
  int p = or_gm_npoints(d);
  r->dim = d;
  r->npts = p;
  r->g = (double*)calloc((size_t)p * d, sizeof(double));
  r->w7 = (double*)calloc((size_t)p, sizeof(double));
  r->w5 = (double*)calloc((size_t)p, sizeof(double));
  if (!r->g || !r->w7 || !r->w5) return -3;
  const double l2 = sqrt(9.0 / 70.0), l3 = sqrt(9.0 / 10.0),
               l4 = sqrt(9.0 / 10.0), l5 = sqrt(9.0 / 19.0);
  const double dd = (double)d;
  /* weights on [-1,1]^d divided by the volume 2^d */
  const double w1 = (12824.0 - 9120.0 * dd + 400.0 * dd * dd) / 19683.0;
  const double w2 = 980.0 / 6561.0;
  const double w3 = (1820.0 - 400.0 * dd) / 19683.0;
  const double w4 = 200.0 / 19683.0;
  const double w5 = 6859.0 / 19683.0 / (double)(1 << d);
  const double v1 = (729.0 - 950.0 * dd + 50.0 * dd * dd) / 729.0;
  const double v2 = 245.0 / 486.0;
  const double v3 = (265.0 - 100.0 * dd) / 1458.0;
  const double v4 = 25.0 / 729.0;
  int k = 0;
  r->w7[k] = w1; r->w5[k] = v1; k++;
  for (int i = 0; i < d; ++i) {
    const double gs[4] = {l2, -l2, l3, -l3};
    for (int s = 0; s < 4; ++s) {
      r->g[(size_t)k * d + i] = gs[s];
      r->w7[k] = s < 2 ? w2 : w3;
      r->w5[k] = s < 2 ? v2 : v3;
      k++;
    }
  }
  for (int i = 0; i < d; ++i)
    for (int j = i + 1; j < d; ++j)
      for (int s = 0; s < 4; ++s) {
        r->g[(size_t)k * d + i] = (s < 2) ? l4 : -l4;
        r->g[(size_t)k * d + j] = (s & 1) ? -l4 : l4;
        r->w7[k] = w4;
        r->w5[k] = v4;
        k++;
      }
  for (int c = 0; c < (1 << d); ++c) {
    for (int m = 0; m < d; ++m) r->g[(size_t)k * d + m] = ((c >> m) & 1) ? -l5 : l5;
    r->w7[k] = w5;
    r->w5[k] = 0.0;
    k++;
  }
  return k == p ? 0 : -1;
}

void or_rule_free(or_rule* r)
{
  free(r->g); free(r->w7); free(r->w5);
  r->g = r->w7 = r->w5 = NULL;
}

/* ------------------------------------------------------------------------ */
/* Integrand families.  Derived constants (1/a^2, 2 pi u) are computed here  */
/* by the same expressions the product library uses on its host side.        */
/* The reference's integrand concept is operator()(double x[], int dim)      */
/* (cudaCuhre.cu:11-22); families follow function.cu:111-156 and the         */
/* workload definitions.                                                     */
/* ------------------------------------------------------------------------ */

static double ipow(double b, int e)
{
  double r = 1.0;
  while (e > 0) {
    if (e & 1) r *= b;
    b *= b;
    e >>= 1;
  }
  return r;
}

double or_f(const or_integrand* f, const double* x)
{
  const int d = f->dim;
  switch (f->family) {
  case OR_GAUSSIAN: {
    double s = 0.0;
    for (int i = 0; i < d; ++i) {
      double t = f->a[i] * (x[i] - f->w[i]);
      s = fma(t, t, s);
    }
    return exp(-s);
  }
  case OR_PRODUCT_PEAK: {
    double p = 1.0;
    for (int i = 0; i < d; ++i) {
      double b = 1.0 / (f->a[i] * f->a[i]);
      double t = x[i] - f->w[i];
      p *= fma(t, t, b);
    }
    return 1.0 / p;
  }
  case OR_OSCILLATORY: {
    double s = 2.0 * M_PI * f->u;
    for (int i = 0; i < d; ++i) s = fma(f->a[i], x[i], s);
    return cos(s);
  }
  case OR_CORNER_PEAK: {
    double s = 1.0;
    for (int i = 0; i < d; ++i) s = fma(f->a[i], x[i], s);
    return ipow(1.0 / s, f->exponent);
  }
  case OR_CONTINUOUS: {
    double s = 0.0;
    for (int i = 0; i < d; ++i) s = fma(f->a[i], fabs(x[i] - f->w[i]), s);
    return exp(-s);
  }
  case OR_POLYNOMIAL: {
    double s = 0.0;
    for (int t = 0; t < f->nterms; ++t) {
      double p = f->coef[t];
      for (int i = 0; i < d; ++i) p *= ipow(x[i], f->powers[t][i]);
      s += p;
    }
    return s;
  }
  case OR_COSINE_SUM: {
    /* the reference driver's test functor, cudaCuhre.cu:11-22 */
    double s = 0.0;
    for (int i = 0; i < d; ++i) s = s - cos(10.0 * x[i]) / 0.054402111088937;
    return s / 2.0;
  }
  case OR_TABLE6: {
    const int nx = f->table_nx, ny = f->table_ny;
    double fx = x[0] * (double)(nx - 1), fy = x[1] * (double)(ny - 1);
    int ix = (int)floor(fx), iy = (int)floor(fy);
    if (ix < 0) ix = 0;
    if (ix > nx - 2) ix = nx - 2;
    if (iy < 0) iy = 0;
    if (iy > ny - 2) iy = ny - 2;
    double tx = fx - (double)ix, ty = fy - (double)iy;
    const double* T = f->table;
    double t00 = T[(size_t)ix * ny + iy], t01 = T[(size_t)ix * ny + iy + 1];
    double t10 = T[(size_t)(ix + 1) * ny + iy], t11 = T[(size_t)(ix + 1) * ny + iy + 1];
    double b0 = fma(tx, t10 - t00, t00);
    double b1 = fma(tx, t11 - t01, t01);
    double B = fma(ty, b1 - b0, b0);
    double x3 = x[2];
    double p3 = 1.0 / (x3 * sqrt(x3));
    double t4 = x[3] - 0.5;
    double g4 = exp(-(t4 * t4) / 0.02);
    double q = fma(x[4], x[5], 1.0);
    return B * p3 * g4 * q;
  }
  default:
    return NAN;
  }
}

/* ------------------------------------------------------------------------ */
/* One region: GPUQuadSample.cu:99-220 semantics with the GM7/5 pair.         */
/*   est = vol * sum w7 f ;  err = vol * |sum w7 f - sum w5 f|                */
/*   D_i = |(f(+l2 e_i)+f(-l2 e_i)-2f0) - (l2/l3)^2 (f(+l3 e_i)+f(-l3 e_i)-2f0)| */
/*   axis = first strict argmax of D_i; if every D_i is 0 the first widest   */
/*   axis (GPUQuadSample.cu:113-125, 150-172).                               */
/* ------------------------------------------------------------------------ */

void or_region_eval(const or_rule* r, const or_integrand* f,
                    const double* lower, const double* width, double* est,
                    double* err, int* axis, double* abssum, double* i5,
                    double* dvec)
{
  const int d = r->dim, p = r->npts;
  double c[OR_MAXDIM] = {0}, h[OR_MAXDIM], x[OR_MAXDIM];
  double vol = 1.0, maxw = 0.0;
  int widest = 0;
  for (int m = 0; m < d; ++m) {
    c[m] = fma(0.5, width[m], lower[m]);
    h[m] = 0.5 * width[m];
    vol *= width[m];
    if (width[m] > maxw) { maxw = width[m]; widest = m; }
  }
  const double ratio = 1.0 / 7.0; /* (l2/l3)^2 = (9/70)/(9/10) */
  double s7 = 0.0, s5 = 0.0, sa = 0.0;
  double f0 = or_f(f, c);
  s7 = fma(r->w7[0], f0, s7);
  s5 = fma(r->w5[0], f0, s5);
  sa = fma(fabs(r->w7[0]), fabs(f0), sa);
  double maxd = 0.0;
  int ax = widest;
  int k = 1;
  for (int i = 0; i < d; ++i) {
    double fv[4];
    for (int s = 0; s < 4; ++s, ++k) {
      for (int m = 0; m < d; ++m) x[m] = fma(h[m], r->g[(size_t)k * d + m], c[m]);
      fv[s] = or_f(f, x);
      s7 = fma(r->w7[k], fv[s], s7);
      s5 = fma(r->w5[k], fv[s], s5);
      sa = fma(fabs(r->w7[k]), fabs(fv[s]), sa);
    }
    double t2 = fma(-2.0, f0, fv[0] + fv[1]);
    double t3 = fma(-2.0, f0, fv[2] + fv[3]);
    double di = fabs(fma(-ratio, t3, t2));
    if (dvec) dvec[i] = di;
    if (di > maxd) { maxd = di; ax = i; }
  }
  const int kc = 1 + 2 * d * d + 2 * d; /* first corner point */
  for (; k < kc; ++k) {
    for (int m = 0; m < d; ++m) x[m] = fma(h[m], r->g[(size_t)k * d + m], c[m]);
    double v = or_f(f, x);
    s7 = fma(r->w7[k], v, s7);
    s5 = fma(r->w5[k], v, s5);
    sa = fma(fabs(r->w7[k]), fabs(v), sa);
  }
  for (; k < p; ++k) {
    for (int m = 0; m < d; ++m) x[m] = fma(h[m], r->g[(size_t)k * d + m], c[m]);
    double v = or_f(f, x);
    s7 = fma(r->w7[k], v, s7);
    sa = fma(fabs(r->w7[k]), fabs(v), sa);
  }
  *est = vol * s7;
  *err = vol * fabs(s7 - s5);
  if (i5) *i5 = vol * s5;
  *abssum = vol * sa;
  *axis = ax;
}

/* ------------------------------------------------------------------------ */
/* Initial grid (GPUKernelquad.cu:23-61, 300-316) and bisection              */
/* (GPUKernelquad.cu:97-149; children interleaved 2j, 2j+1).                 */
/* ------------------------------------------------------------------------ */

static int grid_k(int d) { return d <= 4 ? 4 : (d <= 10 ? 2 : 1); }

int64_t or_initial_grid(int d, const double* lows, const double* highs,
                        double* lower, double* width)
{
  int k = grid_k(d);
  int64_t n = 1;
  for (int m = 0; m < d; ++m) n *= k;
  if (!lower) return n;
  for (int64_t r = 0; r < n; ++r) {
    int64_t q = r;
    for (int m = 0; m < d; ++m) {
      int64_t id = q % k;
      q /= k;
      double w = (highs[m] - lows[m]) / (double)k;
      lower[(size_t)m * n + r] = lows[m] + (double)id * w;
      width[(size_t)m * n + r] = w;
    }
  }
  return n;
}

int64_t or_split(int d, int64_t n, const double* lower, const double* width,
                 const int32_t* active, const int32_t* axis, double* out_lower,
                 double* out_width)
{
  int64_t A = 0;
  for (int64_t r = 0; r < n; ++r) A += active[r] ? 1 : 0;
  int64_t n2 = 2 * A, j = 0;
  for (int64_t r = 0; r < n; ++r) {
    if (!active[r]) continue;
    int a = axis[r];
    for (int m = 0; m < d; ++m) {
      double lo = lower[(size_t)m * n + r], w = width[(size_t)m * n + r];
      if (m == a) {
        double hw = 0.5 * w;
        out_lower[(size_t)m * n2 + 2 * j] = lo;
        out_width[(size_t)m * n2 + 2 * j] = hw;
        out_lower[(size_t)m * n2 + 2 * j + 1] = lo + hw;
        out_width[(size_t)m * n2 + 2 * j + 1] = hw;
      } else {
        out_lower[(size_t)m * n2 + 2 * j] = lo;
        out_width[(size_t)m * n2 + 2 * j] = w;
        out_lower[(size_t)m * n2 + 2 * j + 1] = lo;
        out_width[(size_t)m * n2 + 2 * j + 1] = w;
      }
    }
    ++j;
  }
  return A;
}

/* Neumaier-compensated running sum (the GPU reduces with TwoSum-carried   */
/* pairs, so both sides are accurate to about one ulp of the total).       */
typedef struct { double s, c; } ksum;
static void kadd(ksum* k, double x)
{
  double t = k->s + x;
  if (fabs(k->s) >= fabs(x)) k->c += (k->s - t) + x;
  else k->c += (x - t) + k->s;
  k->s = t;
}
static double kval(const ksum* k) { return k->s + k->c; }

/* ------------------------------------------------------------------------ */
/* Breadth-first loop: GPUKernelquad.cu:507-668 with the global test of      */
/* :583-598 and the per-region classification of GPUQuadPhases.cu:60-78,     */
/* run to convergence (no 2^14 Phase-1 cap), bounded by maxeval/capacity.    */
/* ------------------------------------------------------------------------ */

int or_integrate_bfs(const or_integrand* f, const double* lows,
                     const double* highs, double epsrel, double epsabs,
                     int64_t maxeval, int64_t capacity, int nthreads,
                     or_result* out)
{
  memset(out, 0, sizeof(*out));
  const int d = f->dim;
  or_rule rule;
  if (or_rule_init(&rule, d) != 0) return -1;
  const int p = rule.npts;
  int64_t n = or_initial_grid(d, lows, highs, NULL, NULL);
  if (capacity < n) capacity = n;
  int64_t cap = capacity;
  double* lo = (double*)malloc(sizeof(double) * (size_t)cap * d);
  double* wd = (double*)malloc(sizeof(double) * (size_t)cap * d);
  double* lo2 = (double*)malloc(sizeof(double) * (size_t)cap * d);
  double* wd2 = (double*)malloc(sizeof(double) * (size_t)cap * d);
  double* est = (double*)malloc(sizeof(double) * (size_t)cap);
  double* err = (double*)malloc(sizeof(double) * (size_t)cap);
  double* ab = (double*)malloc(sizeof(double) * (size_t)cap);
  int32_t* ax = (int32_t*)malloc(sizeof(int32_t) * (size_t)cap);
  int32_t* act = (int32_t*)malloc(sizeof(int32_t) * (size_t)cap);
  int rc = 0;
  if (!lo || !wd || !lo2 || !wd2 || !est || !err || !ab || !ax || !act) { rc = -3; goto done; }
  or_initial_grid(d, lows, highs, lo, wd);
  ksum ifin = {0, 0}, efin = {0, 0}, afin = {0, 0}, absfin = {0, 0};
  int64_t ncontrib_fin = 0;
  const double u = DBL_EPSILON / 2.0;
  for (int it = 0;; ++it) {
    int nonfinite = 0;
#ifdef _OPENMP
#pragma omp parallel for num_threads(nthreads > 0 ? nthreads : 1) schedule(static) reduction(| : nonfinite)
#endif
    for (int64_t r = 0; r < n; ++r) {
      double l[OR_MAXDIM], w[OR_MAXDIM];
      for (int m = 0; m < d; ++m) { l[m] = lo[(size_t)m * n + r]; w[m] = wd[(size_t)m * n + r]; }
      int a;
      or_region_eval(&rule, f, l, w, &est[r], &err[r], &a, &ab[r], NULL, NULL);
      ax[r] = a;
      if (!isfinite(est[r]) || !isfinite(err[r])) nonfinite = 1;
      double tol = fmax(epsrel * fabs(est[r]), epsabs);
      act[r] = err[r] > tol ? 1 : 0;
    }
    out->iterations = it + 1;
    out->nregions += n;
    out->neval += n * (int64_t)p;
    if (n > out->peak_regions) out->peak_regions = n;
    if (nonfinite) { rc = -2; break; }
    ksum sall = ifin, eall = efin, aall = afin, absall = absfin;
    int64_t A = 0;
    for (int64_t r = 0; r < n; ++r) {
      kadd(&sall, est[r]);
      kadd(&eall, err[r]);
      kadd(&aall, ab[r]);
      kadd(&absall, fabs(est[r]));
      if (!act[r]) {
        kadd(&ifin, est[r]);
        kadd(&efin, err[r]);
        kadd(&afin, ab[r]);
        kadd(&absfin, fabs(est[r]));
        ncontrib_fin++;
        out->nfinished++;
      } else {
        A++;
      }
    }
    double rG = kval(&sall), eG = kval(&eall);
    out->integral = rG;
    out->error = eG;
    out->bound = (double)(p + 4) * u * kval(&aall) + 2.0 * u * fabs(rG) +
                 (double)(ncontrib_fin + n) * u * u * kval(&absall);
    if (eG <= fmax(epsrel * fabs(rG), epsabs)) { rc = 0; break; }
    if (A == 0) { rc = 1; break; }
    if (out->neval + 2 * A * (int64_t)p > maxeval || 2 * A > cap || it + 1 >= 1000) { rc = 1; break; }
    or_split(d, n, lo, wd, act, ax, lo2, wd2);
    double* t;
    t = lo; lo = lo2; lo2 = t;
    t = wd; wd = wd2; wd2 = t;
    n = 2 * A;
  }
  out->fail = rc > 0 ? rc : 0;
done:
  free(lo); free(wd); free(lo2); free(wd2); free(est); free(err); free(ab); free(ax); free(act);
  or_rule_free(&rule);
  return rc;
}

/* ------------------------------------------------------------------------ */
/* Global adaptive (Cuhre-style) CPU baseline: max-heap on err.              */
/* ------------------------------------------------------------------------ */

typedef struct { double err, est; int axis; int64_t slot; } hnode;

static void hpush(hnode* h, int64_t* n, hnode v)
{
  int64_t i = (*n)++;
  while (i > 0) {
    int64_t pp = (i - 1) / 2;
    if (h[pp].err >= v.err) break;
    h[i] = h[pp];
    i = pp;
  }
  h[i] = v;
}

static hnode hpop(hnode* h, int64_t* n)
{
  hnode top = h[0];
  hnode last = h[--(*n)];
  int64_t i = 0;
  for (;;) {
    int64_t l = 2 * i + 1, rr = l + 1, b = i;
    double be = last.err;
    if (l < *n && h[l].err > be) { b = l; be = h[l].err; }
    if (rr < *n && h[rr].err > be) { b = rr; }
    if (b == i) break;
    h[i] = h[b];
    i = b;
  }
  h[i] = last;
  return top;
}

int or_integrate_heap(const or_integrand* f, const double* lows,
                      const double* highs, double epsrel, double epsabs,
                      int64_t maxeval, or_result* out)
{
  memset(out, 0, sizeof(*out));
  const int d = f->dim;
  or_rule rule;
  if (or_rule_init(&rule, d) != 0) return -1;
  const int p = rule.npts;
  int64_t n0 = or_initial_grid(d, lows, highs, NULL, NULL);
  int64_t maxreg = maxeval / p + n0 + 2;
  double* glo = (double*)malloc(sizeof(double) * (size_t)n0 * d);
  double* gwd = (double*)malloc(sizeof(double) * (size_t)n0 * d);
  double* lo = (double*)malloc(sizeof(double) * (size_t)maxreg * d);
  double* wd = (double*)malloc(sizeof(double) * (size_t)maxreg * d);
  hnode* heap = (hnode*)malloc(sizeof(hnode) * (size_t)maxreg);
  int rc = 0;
  if (!glo || !gwd || !lo || !wd || !heap) { rc = -3; goto done; }
  or_initial_grid(d, lows, highs, glo, gwd);
  int64_t nh = 0, nslot = 0;
  double tot = 0.0, toterr = 0.0;
  for (int64_t r = 0; r < n0; ++r) {
    double* l = lo + (size_t)nslot * d;
    double* w = wd + (size_t)nslot * d;
    for (int m = 0; m < d; ++m) { l[m] = glo[(size_t)m * n0 + r]; w[m] = gwd[(size_t)m * n0 + r]; }
    hnode v; double ab;
    or_region_eval(&rule, f, l, w, &v.est, &v.err, &v.axis, &ab, NULL, NULL);
    if (!isfinite(v.est) || !isfinite(v.err)) { rc = -2; goto done; }
    v.slot = nslot++;
    tot += v.est; toterr += v.err;
    hpush(heap, &nh, v);
  }
  out->nregions = n0;
  out->neval = n0 * (int64_t)p;
  int64_t since = 0;
  for (;;) {
    if (toterr <= fmax(epsrel * fabs(tot), epsabs)) { rc = 0; break; }
    if (out->neval + 2 * (int64_t)p > maxeval || nslot + 1 >= maxreg) { rc = 1; break; }
    hnode v = hpop(heap, &nh);
    double* l = lo + (size_t)v.slot * d;
    double* w = wd + (size_t)v.slot * d;
    double* l2 = lo + (size_t)nslot * d;
    double* w2 = wd + (size_t)nslot * d;
    memcpy(l2, l, sizeof(double) * d);
    memcpy(w2, w, sizeof(double) * d);
    double hw = 0.5 * w[v.axis];
    w[v.axis] = hw;
    w2[v.axis] = hw;
    l2[v.axis] = l[v.axis] + hw;
    hnode c0, c1; double ab;
    or_region_eval(&rule, f, l, w, &c0.est, &c0.err, &c0.axis, &ab, NULL, NULL);
    or_region_eval(&rule, f, l2, w2, &c1.est, &c1.err, &c1.axis, &ab, NULL, NULL);
    if (!isfinite(c0.est + c0.err + c1.est + c1.err)) { rc = -2; break; }
    c0.slot = v.slot;
    c1.slot = nslot++;
    tot += c0.est + c1.est - v.est;
    toterr += c0.err + c1.err - v.err;
    hpush(heap, &nh, c0);
    hpush(heap, &nh, c1);
    out->nregions += 2;
    out->neval += 2 * (int64_t)p;
    if (++since >= 4096) { /* refresh the running sums against drift */
      ksum a = {0, 0}, e = {0, 0};
      for (int64_t i = 0; i < nh; ++i) { kadd(&a, heap[i].est); kadd(&e, heap[i].err); }
      tot = kval(&a); toterr = kval(&e);
      since = 0;
    }
  }
  {
    ksum a = {0, 0}, e = {0, 0};
    for (int64_t i = 0; i < nh; ++i) { kadd(&a, heap[i].est); kadd(&e, heap[i].err); }
    out->integral = kval(&a);
    out->error = kval(&e);
  }
  out->peak_regions = nh;
  out->fail = rc > 0 ? rc : 0;
done:
  free(glo); free(gwd); free(lo); free(wd); free(heap);
  or_rule_free(&rule);
  return rc;
}

/* ------------------------------------------------------------------------ */
/* Closed forms in __float128.                                               */
/* ------------------------------------------------------------------------ */

static int qexact(const or_integrand* f, const double* lows, const double* highs,
                  __float128* res)
{
  const int d = f->dim;
  switch (f->family) {
  case OR_GAUSSIAN: {
    __float128 r = 1.0Q;
    for (int i = 0; i < d; ++i) {
      __float128 a = f->a[i], w = f->w[i];
      r *= sqrtq(M_PIq) / (2.0Q * a) * (erfq(a * ((__float128)highs[i] - w)) - erfq(a * ((__float128)lows[i] - w)));
    }
    *res = r;
    return 0;
  }
  case OR_PRODUCT_PEAK: {
    __float128 r = 1.0Q;
    for (int i = 0; i < d; ++i) {
      __float128 a = f->a[i], w = f->w[i];
      r *= a * (atanq(a * ((__float128)highs[i] - w)) - atanq(a * ((__float128)lows[i] - w)));
    }
    *res = r;
    return 0;
  }
  case OR_OSCILLATORY: {
    /* Re( e^{i 2 pi u} prod (e^{i a hi} - e^{i a lo})/(i a) ) */
    __float128 re = cosq(2.0Q * M_PIq * (__float128)f->u), im = sinq(2.0Q * M_PIq * (__float128)f->u);
    for (int i = 0; i < d; ++i) {
      __float128 a = f->a[i];
      __float128 zr, zi;
      if (a == 0.0Q) { zr = (__float128)highs[i] - (__float128)lows[i]; zi = 0.0Q; }
      else {
        __float128 dr = cosq(a * highs[i]) - cosq(a * lows[i]);
        __float128 di = sinq(a * highs[i]) - sinq(a * lows[i]);
        /* (dr + i di)/(i a) = (di - i dr)/a */
        zr = di / a; zi = -dr / a;
      }
      __float128 nr = re * zr - im * zi, ni = re * zi + im * zr;
      re = nr; im = ni;
    }
    *res = re;
    return 0;
  }
  case OR_CORNER_PEAK: {
    /* repeated integration of (1 + a.x)^-k over the box: inclusion-exclusion */
    const int k = f->exponent;
    __float128 den = 1.0Q;
    for (int i = 0; i < d; ++i) den *= (__float128)f->a[i];
    for (int j = 1; j <= d; ++j) {
      if (j == k) return -1;
      den *= (__float128)(j - k);
    }
    __float128 s = 0.0Q;
    for (int v = 0; v < (1 << d); ++v) {
      __float128 t = 1.0Q;
      int nlo = 0;
      for (int i = 0; i < d; ++i) {
        if ((v >> i) & 1) t += (__float128)f->a[i] * highs[i];
        else { t += (__float128)f->a[i] * lows[i]; nlo++; }
      }
      __float128 term = powq(t, (__float128)(d - k));
      s += (nlo & 1) ? -term : term;
    }
    *res = s / den;
    return 0;
  }
  case OR_CONTINUOUS: {
    __float128 r = 1.0Q;
    for (int i = 0; i < d; ++i) {
      __float128 a = f->a[i], w = f->w[i], lo = lows[i], hi = highs[i];
      __float128 v;
      if (w <= lo) v = (expq(-a * (lo - w)) - expq(-a * (hi - w))) / a;
      else if (w >= hi) v = (expq(-a * (w - hi)) - expq(-a * (w - lo))) / a;
      else v = (2.0Q - expq(-a * (w - lo)) - expq(-a * (hi - w))) / a;
      r *= v;
    }
    *res = r;
    return 0;
  }
  case OR_POLYNOMIAL: {
    __float128 s = 0.0Q;
    for (int t = 0; t < f->nterms; ++t) {
      __float128 p = f->coef[t];
      for (int i = 0; i < d; ++i) {
        int e = f->powers[t][i];
        p *= (powq((__float128)highs[i], e + 1) - powq((__float128)lows[i], e + 1)) / (__float128)(e + 1);
      }
      s += p;
    }
    *res = s;
    return 0;
  }
  case OR_COSINE_SUM: {
    __float128 vol = 1.0Q;
    for (int i = 0; i < d; ++i) vol *= (__float128)highs[i] - (__float128)lows[i];
    __float128 s = 0.0Q;
    for (int i = 0; i < d; ++i) {
      __float128 wi = (__float128)highs[i] - (__float128)lows[i];
      __float128 ci = (sinq(10.0Q * highs[i]) - sinq(10.0Q * lows[i])) / 10.0Q;
      s -= ci * (vol / wi);
    }
    *res = s / (2.0Q * (__float128)0.054402111088937);
    return 0;
  }
  case OR_TABLE6: {
    if (lows[0] != 0.0 || highs[0] != 1.0 || lows[1] != 0.0 || highs[1] != 1.0) return -1;
    const int nx = f->table_nx, ny = f->table_ny;
    /* a bilinear interpolant integrates exactly to the trapezoid sum */
    __float128 s = 0.0Q;
    for (int i = 0; i < nx; ++i)
      for (int j = 0; j < ny; ++j) {
        __float128 wgt = ((i == 0 || i == nx - 1) ? 0.5Q : 1.0Q) * ((j == 0 || j == ny - 1) ? 0.5Q : 1.0Q);
        s += wgt * (__float128)f->table[(size_t)i * ny + j];
      }
    s /= (__float128)(nx - 1) * (__float128)(ny - 1);
    __float128 l3 = lows[2], h3 = highs[2];
    __float128 i3 = 2.0Q * (1.0Q / sqrtq(l3) - 1.0Q / sqrtq(h3));
    __float128 sg = sqrtq(0.02Q);
    __float128 i4 = sqrtq(M_PIq) * sg / 2.0Q * (erfq(((__float128)highs[3] - 0.5Q) / sg) - erfq(((__float128)lows[3] - 0.5Q) / sg));
    __float128 l5 = lows[4], h5 = highs[4], l6 = lows[5], h6 = highs[5];
    __float128 i56 = (h5 - l5) * (h6 - l6) + (h5 * h5 - l5 * l5) * (h6 * h6 - l6 * l6) / 4.0Q;
    *res = s * i3 * i4 * i56;
    return 0;
  }
  default:
    return -1;
  }
}

int or_exact(const or_integrand* f, const double* lows, const double* highs,
             double* hi, double* lo)
{
  __float128 r;
  int rc = qexact(f, lows, highs, &r);
  if (rc) return rc;
  *hi = (double)r;
  *lo = (double)(r - (__float128)(*hi));
  return 0;
}
