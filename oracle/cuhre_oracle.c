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

/* The reference's LC_LT_t (lc_lt_t.h:20-50).  Interp2D (GPUQuadInterp2D.cu:
 * 49-100) on float tables: SmallNearestN picks the last node below the
 * coordinate (index 0 below the grid; the interval is kept inside the grid
 * above the last node, where the reference reads past the end), the nodes
 * and values stay float, the weights are double, the result is rounded to
 * float.  The closed form keeps the reference's float sub-expressions
 * (tau * sigma * sigma, mu + ..., fmsk * fprj) in float: this file is built
 * with -ffp-contract=off and FLT_EVAL_METHOD 0, so C evaluates them as the
 * reference's CUDA does. */
static int lclt_interval(const double* g, int n, double t)
{
  int k = 0;
  for (int i = 1; i < n; ++i) k += (double)(float)g[i] < t ? 1 : 0;
  return k > n - 2 ? n - 2 : k;
}

static float lclt_field(const double* F, int nx, int ix, int iy, double wy1, double wy2, double t1, double t2)
{
  const float q11 = (float)F[ix * nx + iy], q12 = (float)F[ix * nx + iy + 1];
  const float q21 = (float)F[(ix + 1) * nx + iy], q22 = (float)F[(ix + 1) * nx + iy + 1];
  return (float)((q11 * wy2 + q12 * wy1) * t1 + (q21 * wy2 + q22 * wy1) * t2);
}

static double lclt_gaussian(double x, double mu, double sigma)
{
  const double z = (x - mu) / sigma;
  return exp(-z * z / 2.) * 0.3989422804014327 / sigma;
}

static double lclt_eval(const or_integrand* f, const double* x)
{
  const double* T = f->table;
  const int nx = f->table_nx, ny = f->table_ny;
  const double lt = x[0], zt = x[1], lc = x[2];
  const int iy = lclt_interval(T, nx, lt);
  const int ix = lclt_interval(T + nx, ny, zt);
  const float y1 = (float)T[iy], y2 = (float)T[iy + 1];
  const float x1 = (float)T[nx + ix], x2 = (float)T[nx + ix + 1];
  const float den = (x2 - x1) * (y2 - y1);
  const double t1 = (x2 - zt) / den, t2 = (zt - x1) / den;
  const double wy2 = y2 - lt, wy1 = lt - y1;
  const double* F = T + nx + ny;
  const int fs = nx * ny;
  const float tau = lclt_field(F, nx, ix, iy, wy1, wy2, t1, t2);
  const float mu = lclt_field(F + fs, nx, ix, iy, wy1, wy2, t1, t2);
  const float sigma = lclt_field(F + 2 * fs, nx, ix, iy, wy1, wy2, t1, t2);
  const float fmsk = lclt_field(F + 3 * fs, nx, ix, iy, wy1, wy2, t1, t2);
  const float fprj = lclt_field(F + 4 * fs, nx, ix, iy, wy1, wy2, t1, t2);
  const double exptau = exp(tau * (2.0 * mu + tau * sigma * sigma - 2.0 * lc) / 2.0);
  const double root_two_sigma = sqrt(2.0) * sigma;
  const double mu_tau_sig_sqr = mu + tau * sigma * sigma;
  return (1.0 - fmsk) * (1.0 - fprj) * lclt_gaussian(lc, mu, sigma) +
         0.5 * ((1.0 - fmsk) * fprj * tau + fmsk * fprj / lt) * exptau *
             erfc((mu_tau_sig_sqr - lc) / root_two_sigma) +
         0.5 * fmsk / lt * (erfc((lc - mu) / root_two_sigma) - erfc((lc + lt - mu) / root_two_sigma)) -
         0.5 * fmsk * fprj / lt * (exp(-tau * lt) * exptau * erfc((mu_tau_sig_sqr - (lc + lt)) / root_two_sigma));
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
  case OR_LC_LT_T:
    return lclt_eval(f, x);
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
/* __float128 restatement of the families (rounding check only): the same   */
/* formulas with every operation in 113-bit arithmetic.  NaN for the        */
/* tabulated families.                                                      */
/* ------------------------------------------------------------------------ */

static __float128 ipowq(__float128 b, int e)
{
  __float128 r = 1.0Q;
  while (e > 0) {
    if (e & 1) r *= b;
    b *= b;
    e >>= 1;
  }
  return r;
}

static __float128 or_fq(const or_integrand* f, const __float128* x)
{
  const int d = f->dim;
  switch (f->family) {
  case OR_GAUSSIAN: {
    __float128 s = 0.0Q;
    for (int i = 0; i < d; ++i) {
      const __float128 t = (__float128)f->a[i] * (x[i] - (__float128)f->w[i]);
      s += t * t;
    }
    return expq(-s);
  }
  case OR_PRODUCT_PEAK: {
    __float128 p = 1.0Q;
    for (int i = 0; i < d; ++i) {
      const __float128 a = f->a[i], t = x[i] - (__float128)f->w[i];
      p *= 1.0Q / (a * a) + t * t;
    }
    return 1.0Q / p;
  }
  case OR_OSCILLATORY: {
    __float128 s = 2.0Q * M_PIq * (__float128)f->u;
    for (int i = 0; i < d; ++i) s += (__float128)f->a[i] * x[i];
    return cosq(s);
  }
  case OR_CORNER_PEAK: {
    __float128 s = 1.0Q;
    for (int i = 0; i < d; ++i) s += (__float128)f->a[i] * x[i];
    return ipowq(1.0Q / s, f->exponent);
  }
  case OR_CONTINUOUS: {
    __float128 s = 0.0Q;
    for (int i = 0; i < d; ++i) s += (__float128)f->a[i] * fabsq(x[i] - (__float128)f->w[i]);
    return expq(-s);
  }
  case OR_POLYNOMIAL: {
    __float128 s = 0.0Q;
    for (int t = 0; t < f->nterms; ++t) {
      __float128 p = f->coef[t];
      for (int i = 0; i < d; ++i) p *= ipowq(x[i], f->powers[t][i]);
      s += p;
    }
    return s;
  }
  case OR_COSINE_SUM: {
    __float128 s = 0.0Q;
    for (int i = 0; i < d; ++i) s -= cosq(10.0Q * x[i]) / (__float128)0.054402111088937;
    return s / 2.0Q;
  }
  default:
    return nanq("");
  }
}

double or_fq_residual(const or_integrand* f, const double* x, double* lo)
{
  __float128 xq[OR_MAXDIM];
  for (int i = 0; i < f->dim; ++i) xq[i] = x[i];
  const __float128 v = or_fq(f, xq);
  const double hi = (double)v;
  *lo = (double)(v - (__float128)hi);
  return hi;
}

/* Shadow of one region's basic-rule sum: s_star adds w_k * fhat_k (exact
 * product of two doubles) in 113 bits; with want_rule, s_rule adds
 * w_k * f(c + h g_k) with node and integrand in 113 bits. */
typedef struct {
  __float128 s_star, s_rule;
  int want_rule;
  __float128 cq[OR_MAXDIM], hq[OR_MAXDIM];
} qshadow;

static void qs_begin(qshadow* q, int d, const double* lower, const double* width)
{
  q->s_star = 0.0Q;
  q->s_rule = 0.0Q;
  for (int m = 0; m < d; ++m) {
    q->hq[m] = 0.5Q * (__float128)width[m];
    q->cq[m] = (__float128)lower[m] + q->hq[m];
  }
}

/* point with generator offsets g (half-width units) */
static void qs_point(qshadow* q, const or_integrand* f, const double* g, double w, double fhat)
{
  q->s_star += (__float128)w * (__float128)fhat;
  if (q->want_rule) {
    __float128 x[OR_MAXDIM];
    for (int m = 0; m < f->dim; ++m) x[m] = q->cq[m] + q->hq[m] * (__float128)g[m];
    q->s_rule += (__float128)w * or_fq(f, x);
  }
}

static __float128 qs_volume(int d, const double* width)
{
  __float128 v = 1.0Q;
  for (int m = 0; m < d; ++m) v *= (__float128)width[m];
  return v;
}

/* ------------------------------------------------------------------------ */
/* One region: GPUQuadSample.cu:99-220 semantics with the GM7/5 pair.         */
/*   est = vol * sum w7 f ;  err = vol * |sum w7 f - sum w5 f|                */
/*   D_i = |(f(+l2 e_i)+f(-l2 e_i)-2f0) - (l2/l3)^2 (f(+l3 e_i)+f(-l3 e_i)-2f0)| */
/*   axis = first strict argmax of D_i; if every D_i is 0 the first widest   */
/*   axis (GPUQuadSample.cu:113-125, 150-172).                               */
/* ------------------------------------------------------------------------ */

static void gm_region(const or_rule* r, const or_integrand* f, const double* lower,
                      const double* width, double* est, double* err, int* axis,
                      double* abssum, double* i5, double* dvec, qshadow* q)
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
  if (q) qs_point(q, f, r->g, r->w7[0], f0);
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
      if (q) qs_point(q, f, r->g + (size_t)k * d, r->w7[k], fv[s]);
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
    if (q) qs_point(q, f, r->g + (size_t)k * d, r->w7[k], v);
  }
  for (; k < p; ++k) {
    for (int m = 0; m < d; ++m) x[m] = fma(h[m], r->g[(size_t)k * d + m], c[m]);
    double v = or_f(f, x);
    s7 = fma(r->w7[k], v, s7);
    sa = fma(fabs(r->w7[k]), fabs(v), sa);
    if (q) qs_point(q, f, r->g + (size_t)k * d, r->w7[k], v);
  }
  *est = vol * s7;
  *err = vol * fabs(s7 - s5);
  if (i5) *i5 = vol * s5;
  *abssum = vol * sa;
  *axis = ax;
}

void or_region_eval(const or_rule* r, const or_integrand* f,
                    const double* lower, const double* width, double* est,
                    double* err, int* axis, double* abssum, double* i5,
                    double* dvec)
{
  gm_region(r, f, lower, width, est, err, axis, abssum, i5, dvec, NULL);
}

/* ------------------------------------------------------------------------ */
/* Degree-9 rule (KEY 9), quad/GPUquad/GPUQuadRule.cu:30-226.  The weight    */
/* table and the radii are the reference's (they define the rule); the       */
/* point enumeration is set by set; scale/norm as GPUQuadRule.cu:209-225;    */
/* error and axis as GPUQuadSample.cu:113-172, 195-216.                      */
/* ------------------------------------------------------------------------ */

int or_r9_npoints(int d)
{
  return 1 + 8 * d + 2 * d * (d - 1) + 4 * d * (d - 1) + 4 * d * (d - 1) * (d - 2) / 3 + (1 << d);
}

int or_rule9_init(or_rule9* r, int d)
{
  if (d < 2 || d > OR_MAXDIM) return -1;
  const double n = (double)d;
  /* radii as fractions of the region's width (GPUQuadRule.cu:172-176) */
  const double a[5] = {.47795365790226950619, .20302858736911986780, .44762735462617812882, .125,
                       .34303789878087814570};
  for (int k = 0; k < 5; ++k) r->t[k] = 2.0 * a[k]; /* half-width units */
  /* weights (GPUQuadRule.cu:33-116): recurring polynomials in d */
  const double A = n * (n * (n * (-.0023611709677855117884) + .11415390023857325268) + (-.63833920076702389094)) +
                   .74849988504685208004;
  const double B = n * (n * (n * (-.0014324017033399125142) + .057471507864489725949) + (-.14225104571434243234)) -
                   (-.062875028738286979989);
  const double C = n * (n * .0035417564516782676826 + (-.072609367395893679605)) + .10557491625218991012;
  const double D = n * (n * .0021486025550098687713 + (-.032268563892953949998)) + .010636783990231217481;
  const double E = n * (-.04508628929435784076) + .21415883524352793401;
  const double F = n * (-.027351546526545644722) + .054941067048711234101;
  const double G = .015532417276607053264 - n * .0035417564516782676826;
  const double H = .0035328099607090870236 - n * .0021486025550098687713;
  const double c = (double)(1 << d);
  const double w[OR_R9_SETS][OR_R9_RULES] = {
      {A, B, n * .254591133248959089 - B, n * (n * (-1.207328566678236261) + .89567365764160676508) - 1 + A,
       n * (-.36479356986049146661) + 1 - A},
      {C, D, .014689102496143490175 - D, n * .51134708346467591431 + .45976448120806344646 + C,
       .18239678493024573331 - C},
      {E, F, .11937596202570775297 - F, n * .65089519391920250593 + .14744939829434460168, -E},
      {.057693384490973483573, .034999626602143583822, -.057693384490973483573, -1.3868627719278281436,
       -.057693384490973483573},
      {0, 0, -.2386668732575008879, 0, 0},
      {G, H, -H, .09231719987444221619 + G, -G},
      {.02254314464717892038, .013675773263272822361, -.013675773263272822361, -.32544759695960125297,
       -.02254314464717892038},
      {.0017708782258391338413, .0010743012775049343856, -.0010743012775049343856, .0017708782258391338413,
       -.0017708782258391338413},
      {.25150011495314791996 / c, -.062875028738286979989 / c, .062875028738286979989 / c,
       .25150011495314791996 / c, -.25150011495314791996 / c}};
  const int cnt[OR_R9_SETS] = {1, 2 * d, 2 * d, 2 * d, 2 * d, 2 * d * (d - 1), 4 * d * (d - 1),
                               4 * d * (d - 1) * (d - 2) / 3, 1 << d};
  r->dim = d;
  r->npts = or_r9_npoints(d);
  for (int s = 0; s < OR_R9_SETS; ++s) {
    r->count[s] = cnt[s];
    for (int q = 0; q < OR_R9_RULES; ++q) {
      r->w[s][q] = w[s][q];
      r->scale[s][q] = 0.0;
      r->norm[s][q] = 0.0;
    }
  }
  /* null-rule combinations: for rules 1..3 and each set, the scale that
   * cancels that set's weight in w[r+1] + scale*w[r], and the inverse
   * 1-norm of the combined rule (GPUQuadRule.cu:209-225) */
  for (int s = 0; s < OR_R9_SETS; ++s)
    for (int q = 1; q < OR_R9_RULES - 1; ++q) {
      const double sc = w[s][q] == 0.0 ? 100.0 : -w[s][q + 1] / w[s][q];
      double sum = 0.0;
      for (int x = 0; x < OR_R9_SETS; ++x) sum += (double)cnt[x] * fabs(w[x][q + 1] + sc * w[x][q]);
      r->scale[s][q] = sc;
      r->norm[s][q] = 1.0 / sum;
    }
  return 0;
}

/* Error of the degree-9 rule from its five rule sums (GPUQuadSample.cu:
 * 195-216): for each pair of successive null rules (q, q+1), q = 1..3, the
 * largest over the sets t of the normalised combination |N_{q+1} +
 * scale[t][q] N_q| norm[t][q]; then with errcoeff {5, 1, 5}: e1 when the
 * three fall off geometrically (5 e1 <= e2 and 5 e2 <= e3), else 5 max. */
double or_r9_error(const or_rule9* r, const double* sums, double* e_out)
{
  double e[OR_R9_RULES] = {0, 0, 0, 0, 0};
  for (int q = 1; q < OR_R9_RULES - 1; ++q) {
    double m = 0.0;
    for (int s = 0; s < OR_R9_SETS; ++s) m = fmax(m, fabs(sums[q + 1] + r->scale[s][q] * sums[q]) * r->norm[s][q]);
    e[q] = m;
  }
  if (e_out)
    for (int q = 1; q < 4; ++q) e_out[q - 1] = e[q];
  const double e1 = e[1], e2 = e[2], e3 = e[3];
  return (5.0 * e1 <= e2 && 5.0 * e2 <= e3) ? e1 : 5.0 * fmax(fmax(e1, e2), e3);
}

/* accumulates one point of set s */
static void r9_add(const or_rule9* r, int s, double v, double* sums, double* sa, qshadow* qs)
{
  for (int q = 0; q < OR_R9_RULES; ++q) sums[q] = fma(r->w[s][q], v, sums[q]);
  *sa = fma(fabs(r->w[s][0]), fabs(v), *sa);
  if (qs) qs->s_star += (__float128)r->w[s][0] * (__float128)v;
}

static void r9_region(const or_rule9* r, const or_integrand* f, const double* lower,
                      const double* width, double* est, double* err, int* axis, double* abssum,
                      double* dvec, double* sums_out, qshadow* qs)
{
  const int d = r->dim;
  double c[OR_MAXDIM] = {0}, h[OR_MAXDIM], x[OR_MAXDIM];
  double vol = 1.0, maxw = 0.0;
  int widest = 0;
  for (int m = 0; m < d; ++m) {
    c[m] = fma(0.5, width[m], lower[m]);
    h[m] = 0.5 * width[m];
    vol *= width[m];
    if (width[m] > maxw) { maxw = width[m]; widest = m; }
  }
  double sums[OR_R9_RULES] = {0, 0, 0, 0, 0}, sa = 0.0;
  const double f0 = or_f(f, c);
  r9_add(r, 0, f0, sums, &sa, qs);
  /* axis sets 1..4, fourth difference from sets 1 and 2 */
  const double ratio = (r->t[1] / r->t[0]) * (r->t[1] / r->t[0]);
  double maxd = 0.0;
  int ax = widest;
  for (int i = 0; i < d; ++i) {
    double fv[4][2];
    for (int k = 0; k < 4; ++k)
      for (int sg = 0; sg < 2; ++sg) {
        for (int m = 0; m < d; ++m) x[m] = c[m];
        x[i] = fma(h[i], sg ? -r->t[k] : r->t[k], c[i]);
        fv[k][sg] = or_f(f, x);
        r9_add(r, 1 + k, fv[k][sg], sums, &sa, qs);
      }
    const double di = fabs(2.0 * (1.0 - ratio) * f0 + ratio * (fv[0][0] + fv[0][1]) - (fv[1][0] + fv[1][1]));
    if (dvec) dvec[i] = di;
    if (di > maxd) { maxd = di; ax = i; }
  }
  /* pair sets 5 (t1, t1) and 6 (t1, t2 in both orders) */
  for (int i = 0; i < d; ++i)
    for (int j = i + 1; j < d; ++j) {
      for (int sg = 0; sg < 4; ++sg) {
        for (int m = 0; m < d; ++m) x[m] = c[m];
        x[i] = fma(h[i], (sg & 1) ? -r->t[0] : r->t[0], c[i]);
        x[j] = fma(h[j], (sg & 2) ? -r->t[0] : r->t[0], c[j]);
        r9_add(r, 5, or_f(f, x), sums, &sa, qs);
      }
      for (int o = 0; o < 2; ++o)
        for (int sg = 0; sg < 4; ++sg) {
          for (int m = 0; m < d; ++m) x[m] = c[m];
          const double ti = o ? r->t[1] : r->t[0], tj = o ? r->t[0] : r->t[1];
          x[i] = fma(h[i], (sg & 1) ? -ti : ti, c[i]);
          x[j] = fma(h[j], (sg & 2) ? -tj : tj, c[j]);
          r9_add(r, 6, or_f(f, x), sums, &sa, qs);
        }
    }
  /* triple set 7 */
  for (int i = 0; i < d; ++i)
    for (int j = i + 1; j < d; ++j)
      for (int k = j + 1; k < d; ++k)
        for (int sg = 0; sg < 8; ++sg) {
          for (int m = 0; m < d; ++m) x[m] = c[m];
          x[i] = fma(h[i], (sg & 1) ? -r->t[0] : r->t[0], c[i]);
          x[j] = fma(h[j], (sg & 2) ? -r->t[0] : r->t[0], c[j]);
          x[k] = fma(h[k], (sg & 4) ? -r->t[0] : r->t[0], c[k]);
          r9_add(r, 7, or_f(f, x), sums, &sa, qs);
        }
  /* corner set 8 */
  for (int q = 0; q < (1 << d); ++q) {
    for (int m = 0; m < d; ++m) x[m] = fma(h[m], ((q >> m) & 1) ? -r->t[4] : r->t[4], c[m]);
    r9_add(r, 8, or_f(f, x), sums, &sa, qs);
  }
  if (sums_out)
    for (int q = 0; q < OR_R9_RULES; ++q) sums_out[q] = sums[q];
  *est = vol * sums[0];
  *err = vol * or_r9_error(r, sums, NULL);
  *abssum = vol * sa;
  *axis = ax;
}

void or_region_eval9(const or_rule9* r, const or_integrand* f, const double* lower,
                     const double* width, double* est, double* err, int* axis, double* abssum,
                     double* dvec, double* sums_out)
{
  r9_region(r, f, lower, width, est, err, axis, abssum, dvec, sums_out, NULL);
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
  return or_integrate_bfs_key(f, lows, highs, epsrel, epsabs, maxeval, capacity, nthreads, 7, out);
}

/* The bound's derivation (standard model fl(a op b) = (a op b)(1 + e),
 * |e| <= u; an FMA rounds once):
 *   rule sum  s^ = sum_k w_k f_k with every term through at most p
 *             roundings (an FMA chain, or segment chains merged by adds):
 *             |s^ - s| <= g_p sum |w_k f_k|;
 *   estimate  est = fl(V^ s^), V^ = fl(prod of d widths):
 *             |est - V s| <= (g_p + g_d + g_p g_d) V sum |w f| <= g_{p+d} V sum |w f|;
 *   abssum    A_r = fl(V^ sum^ |w f|) >= V sum |w f| (1 - g_{p+d}), and the
 *             plain FP64 sum of the A_r is low by at most g_n: the factor
 *             (1 + 4 (K + n) u) covers both, K = p + d + 2;
 *   regions   T = sum est_r summed as (value, TwoSum error) pairs in any
 *             tree of depth <= n: |I - T| <= u |T| + (1 + u) |c^ - c|, the
 *             compensation's own rounding |c^ - c| <= g_{2n} sum |e_i|
 *             <= 2 n^2 u^2 / (1 - 2 n u) sum |est_r|; u |T| <= 2 u |I| for
 *             any |I - T| <= |I| / 2.
 * (Higham, Accuracy and Stability of Numerical Algorithms, 2nd ed., 3.1,
 * 4.2; Ogita, Rump and Oishi 2005, Sum2.) */
double or_fp64_bound(int p, int d, int64_t n, double A, double E, double I)
{
  const double u = DBL_EPSILON / 2.0;
  const double K = (double)p + (double)d + 2.0;
  const double gK = K * u / (1.0 - K * u);
  const double nn = (double)n;
  return gK * A * (1.0 + 4.0 * (K + nn) * u) + 2.0 * u * fabs(I) +
         2.0 * nn * nn * u * u / (1.0 - 2.0 * nn * u) * E;
}

static void q_split(__float128 v, double* hi, double* lo)
{
  *hi = (double)v;
  *lo = (double)(v - (__float128)(*hi));
}

void or_region_check(int key, const or_integrand* f, const double* lower, const double* width,
                     double* out)
{
  const int d = f->dim;
  qshadow q;
  q.want_rule = key != 9;
  qs_begin(&q, d, lower, width);
  double est, err, ab;
  int ax;
  if (key == 9) {
    or_rule9 r9;
    or_rule9_init(&r9, d);
    r9_region(&r9, f, lower, width, &est, &err, &ax, &ab, NULL, NULL, &q);
  } else {
    or_rule r7;
    or_rule_init(&r7, d);
    gm_region(&r7, f, lower, width, &est, &err, &ax, &ab, NULL, NULL, &q);
    or_rule_free(&r7);
  }
  const __float128 V = qs_volume(d, width);
  out[0] = est;
  out[1] = ab;
  q_split(V * q.s_star, &out[2], &out[3]);
  if (key == 9 || isnanq(q.s_rule)) {
    out[4] = out[5] = NAN;
  } else {
    q_split(V * q.s_rule, &out[4], &out[5]);
  }
}

int or_integrate_bfs_key(const or_integrand* f, const double* lows,
                         const double* highs, double epsrel, double epsabs,
                         int64_t maxeval, int64_t capacity, int nthreads,
                         int key, or_result* out)
{
  return or_integrate_bfs_check(f, lows, highs, epsrel, epsabs, maxeval, capacity, nthreads, key, 0, out,
                                NULL);
}

static int bfs_core(const or_integrand* f, const double* lows, const double* highs, double epsrel,
                    double epsabs, int64_t maxeval, int64_t capacity, int nthreads, int key, int want_rule,
                    double allot, int normalized, or_result* out, or_check* chk);

int or_integrate_bfs_check(const or_integrand* f, const double* lows,
                           const double* highs, double epsrel, double epsabs,
                           int64_t maxeval, int64_t capacity, int nthreads,
                           int key, int want_rule, or_result* out, or_check* chk)
{
  return bfs_core(f, lows, highs, epsrel, epsabs, maxeval, capacity, nthreads, key, want_rule, 0.0, 0, out, chk);
}

int or_integrate_bfs_allot(const or_integrand* f, const double* lows, const double* highs, double epsrel,
                           double epsabs, int64_t maxeval, int64_t capacity, int nthreads, double allot,
                           or_result* out)
{
  if (!(allot >= 0.0) || allot > 1.0) return -1;
  return bfs_core(f, lows, highs, epsrel, epsabs, maxeval, capacity, nthreads, 7, 0, allot, 0, out, NULL);
}

int or_integrate_bfs_normalized(const or_integrand* f, const double* lows, const double* highs, double epsrel,
                                double epsabs, int64_t maxeval, int64_t capacity, int nthreads, double allot,
                                or_result* out)
{
  if (!(allot >= 0.0) || allot > 1.0) return -1;
  return bfs_core(f, lows, highs, epsrel, epsabs, maxeval, capacity, nthreads, 7, 0, allot, 1, out, NULL);
}

/* The loop.  allot = b > 0 replaces the per-region test by the shared
 * allotment: active iff err > max(epsrel ((1 - b) |est| + b |I| vrel),
 * epsabs), with I the iteration's estimate (finished so far plus every
 * region of the iteration) and vrel the region's share of the box volume;
 * b = 0 is the reference's test (GPUQuadPhases.cu:63-73).  normalized: the
 * tolerance is epsrel |I| ((1 - b) |est| / E + b vrel) instead, E = sum
 * |est| over the finished regions and the iteration's, so the tolerances of
 * all regions add up to the global epsrel |I| (any b, cancellation or not). */
static int bfs_core(const or_integrand* f, const double* lows, const double* highs, double epsrel,
                    double epsabs, int64_t maxeval, int64_t capacity, int nthreads, int key, int want_rule,
                    double allot, int normalized, or_result* out, or_check* chk)
{
  memset(out, 0, sizeof(*out));
  if (chk) memset(chk, 0, sizeof(*chk));
  const int d = f->dim;
  if (key != 0 && key != 7 && key != 9) return -1;
  or_rule rule;
  or_rule9 rule9;
  if (or_rule_init(&rule, d) != 0) return -1;
  if (key == 9 && or_rule9_init(&rule9, d) != 0) { or_rule_free(&rule); return -1; }
  const int p = key == 9 ? rule9.npts : rule.npts;
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
  /* __float128 shadows of every region's estimate (check runs only) */
  __float128* qstar = chk ? (__float128*)malloc(sizeof(__float128) * (size_t)cap) : NULL;
  __float128* qrule = chk ? (__float128*)malloc(sizeof(__float128) * (size_t)cap) : NULL;
  const int rule_q = chk && want_rule && key != 9;
  __float128 qfin_star = 0.0Q, qfin_rule = 0.0Q;
  int rc = 0;
  if (!lo || !wd || !lo2 || !wd2 || !est || !err || !ab || !ax || !act || (chk && (!qstar || !qrule))) {
    rc = -3;
    goto done;
  }
  or_initial_grid(d, lows, highs, lo, wd);
  ksum ifin = {0, 0}, efin = {0, 0}, afin = {0, 0}, absfin = {0, 0};
  int64_t ncontrib_fin = 0;
  for (int it = 0;; ++it) {
    int nonfinite = 0;
#ifdef _OPENMP
#pragma omp parallel for num_threads(nthreads > 0 ? nthreads : 1) schedule(static) reduction(| : nonfinite)
#endif
    for (int64_t r = 0; r < n; ++r) {
      double l[OR_MAXDIM], w[OR_MAXDIM];
      for (int m = 0; m < d; ++m) { l[m] = lo[(size_t)m * n + r]; w[m] = wd[(size_t)m * n + r]; }
      int a;
      qshadow qs, *qp = NULL;
      if (chk) {
        qs.want_rule = rule_q;
        qs_begin(&qs, d, l, w);
        qp = &qs;
      }
      if (key == 9)
        r9_region(&rule9, f, l, w, &est[r], &err[r], &a, &ab[r], NULL, NULL, qp);
      else
        gm_region(&rule, f, l, w, &est[r], &err[r], &a, &ab[r], NULL, NULL, qp);
      if (chk) {
        const __float128 V = qs_volume(d, w);
        qstar[r] = V * qs.s_star;
        qrule[r] = V * qs.s_rule;
      }
      ax[r] = a;
      if (!isfinite(est[r]) || !isfinite(err[r])) nonfinite = 1;
      double tol = fmax(epsrel * fabs(est[r]), epsabs);
      act[r] = err[r] > tol ? 1 : 0;
    }
    if (allot > 0.0 || normalized) {
      /* the shared allotment needs the iteration's estimate first */
      ksum all = ifin, aall = absfin;
      for (int64_t r = 0; r < n; ++r) {
        kadd(&all, est[r]);
        kadd(&aall, fabs(est[r]));
      }
      const double absI = fabs(kval(&all)), E = kval(&aall);
      double boxv = 1.0;
      for (int m = 0; m < d; ++m) boxv *= highs[m] - lows[m];
      for (int64_t r = 0; r < n; ++r) {
        double vrel = 1.0;
        for (int m = 0; m < d; ++m) vrel *= wd[(size_t)m * n + r] / (highs[m] - lows[m]);
        const double a = normalized ? epsrel * absI * ((1.0 - allot) * fabs(est[r]) / E + allot * vrel)
                                    : epsrel * ((1.0 - allot) * fabs(est[r]) + allot * absI * vrel);
        act[r] = isfinite(est[r]) && isfinite(err[r]) && err[r] > fmax(a, epsabs) ? 1 : 0;
      }
      (void)boxv;
    }
    out->iterations = it + 1;
    out->nregions += n;
    out->neval += n * (int64_t)p;
    if (n > out->peak_regions) out->peak_regions = n;
    if (nonfinite) { rc = -2; break; }
    ksum sall = ifin, eall = efin, aall = afin, absall = absfin;
    __float128 qall_star = qfin_star, qall_rule = qfin_rule;
    int64_t A = 0;
    for (int64_t r = 0; r < n; ++r) {
      kadd(&sall, est[r]);
      kadd(&eall, err[r]);
      kadd(&aall, ab[r]);
      kadd(&absall, fabs(est[r]));
      if (chk) {
        qall_star += qstar[r];
        qall_rule += qrule[r];
      }
      if (!act[r]) {
        kadd(&ifin, est[r]);
        kadd(&efin, err[r]);
        kadd(&afin, ab[r]);
        kadd(&absfin, fabs(est[r]));
        if (chk) {
          qfin_star += qstar[r];
          qfin_rule += qrule[r];
        }
        ncontrib_fin++;
        out->nfinished++;
      } else {
        A++;
      }
    }
    double rG = kval(&sall), eG = kval(&eall);
    out->integral = rG;
    out->error = eG;
    out->bound = or_fp64_bound(p, d, ncontrib_fin + n, kval(&aall), kval(&absall), rG);
    if (chk) {
      q_split(qall_star, &chk->istar_hi, &chk->istar_lo);
      if (rule_q && !isnanq(qall_rule)) {
        q_split(qall_rule, &chk->irule_hi, &chk->irule_lo);
        chk->have_rule = 1;
      } else {
        chk->irule_hi = chk->irule_lo = NAN;
        chk->have_rule = 0;
      }
      chk->abs_total = kval(&aall);
      chk->aest_total = kval(&absall);
      chk->ncontrib = ncontrib_fin + n;
    }
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
  free(qstar); free(qrule);
  or_rule_free(&rule);
  return rc;
}

/* ------------------------------------------------------------------------ */
/* Two-phase method of the reference (Phase 1 + per-region Phase 2).         */
/* ------------------------------------------------------------------------ */

typedef struct {
  int key, d;
  or_rule r7;
  or_rule9 r9;
} anyrule;

static void any_eval(const anyrule* R, const or_integrand* f, const double* l, const double* w,
                     double* est, double* err, int* ax, double* ab)
{
  if (R->key == 9)
    or_region_eval9(&R->r9, f, l, w, est, err, ax, ab, NULL, NULL);
  else
    or_region_eval(&R->r7, f, l, w, est, err, ax, ab, NULL, NULL);
}

/* heap of slots ordered by (err desc, seq asc) */
static int p2_before(const double* err, const int64_t* seq, int a, int b)
{
  return err[a] > err[b] || (err[a] == err[b] && seq[a] < seq[b]);
}

static void p2_push(int* heap, int* hn, int slot, const double* err, const int64_t* seq)
{
  int i = (*hn)++;
  heap[i] = slot;
  while (i > 0) {
    int pa = (i - 1) / 2;
    if (!p2_before(err, seq, heap[i], heap[pa])) break;
    int t = heap[i]; heap[i] = heap[pa]; heap[pa] = t;
    i = pa;
  }
}

static int p2_pop(int* heap, int* hn, const double* err, const int64_t* seq)
{
  int top = heap[0];
  heap[0] = heap[--(*hn)];
  int i = 0;
  for (;;) {
    int l = 2 * i + 1, r = l + 1, b = i;
    if (l < *hn && p2_before(err, seq, heap[l], heap[b])) b = l;
    if (r < *hn && p2_before(err, seq, heap[r], heap[b])) b = r;
    if (b == i) break;
    int t = heap[i]; heap[i] = heap[b]; heap[b] = t;
    i = b;
  }
  return top;
}

/* Phase 2 of one region (GPUQuadPhases.cu:422-536). */
/* budget[0] receives the sum of the abssums of every evaluated region,
 * budget[1] the RES chain's sum of 2 (|child 0| + |child 1| + |parent|) + |RES|
 * (the library's Phase-2 rounding budget); *zeroed_out 1 when the region's
 * error passed 1e10 and it was zeroed. */
static void phase2_region(const anyrule* R, const or_integrand* f, const double* lo0, const double* w0,
                          double epsrel, double epsabs, int pool_max, double* res_out, double* err_out,
                          int* nreg_out, int* active_out, double* budget, int* zeroed_out)
{
  const int d = R->d, cap = pool_max + 2;
  double* lo = (double*)malloc(sizeof(double) * (size_t)cap * d);
  double* wd = (double*)malloc(sizeof(double) * (size_t)cap * d);
  double* est = (double*)malloc(sizeof(double) * cap);
  double* er = (double*)malloc(sizeof(double) * cap);
  int* ax = (int*)malloc(sizeof(int) * cap);
  int64_t* seq = (int64_t*)malloc(sizeof(int64_t) * cap);
  int* heap = (int*)malloc(sizeof(int) * cap);
  int hn = 0;
  int64_t next = 0;
  double ab;
  for (int m = 0; m < d; ++m) { lo[m] = lo0[m]; wd[m] = w0[m]; }
  any_eval(R, f, lo, wd, &est[0], &er[0], &ax[0], &ab);
  double rab = ab, rres = 0.0;
  seq[0] = next++;
  p2_push(heap, &hn, 0, er, seq);
  double RES = est[0], ERR = er[0];
  int count = 1;
  while (count <= pool_max && (count == 1 || ERR > fmax(epsrel * fabs(RES), epsabs))) {
    const int a = p2_pop(heap, &hn, er, seq);
    const int b = count;
    const double pe = est[a], pr = er[a];
    const int k = ax[a];
    for (int m = 0; m < d; ++m) { lo[(size_t)b * d + m] = lo[(size_t)a * d + m]; wd[(size_t)b * d + m] = wd[(size_t)a * d + m]; }
    const double hw = 0.5 * wd[(size_t)a * d + k];
    wd[(size_t)a * d + k] = hw;
    wd[(size_t)b * d + k] = hw;
    lo[(size_t)b * d + k] = lo[(size_t)a * d + k] + hw;
    any_eval(R, f, &lo[(size_t)a * d], &wd[(size_t)a * d], &est[a], &er[a], &ax[a], &ab);
    rab += ab;
    any_eval(R, f, &lo[(size_t)b * d], &wd[(size_t)b * d], &est[b], &er[b], &ax[b], &ab);
    rab += ab;
    /* Cuba's refinement from the parent difference (GPUQuadPhases.cu:498-515) */
    const double diff = fabs(0.25 * (est[a] + est[b] - pe));
    const double e = er[a] + er[b];
    if (e > 0) {
      const double c = 1 + 2 * diff / e;
      er[a] *= c;
      er[b] *= c;
    }
    er[a] += diff;
    er[b] += diff;
    ERR += er[a] + er[b] - pr;
    RES += est[a] + est[b] - pe;
    rres += 2.0 * (fabs(est[a]) + fabs(est[b]) + fabs(pe)) + fabs(RES);
    seq[a] = next++;
    seq[b] = next++;
    p2_push(heap, &hn, a, er, seq);
    p2_push(heap, &hn, b, er, seq);
    ++count;
  }
  int active = ERR > fmax(epsrel * fabs(RES), epsabs);
  *zeroed_out = 0;
  if (ERR > 1e10) { /* the reference zeroes a diverging region and flags it */
    RES = 0.0;
    ERR = 0.0;
    active = 1;
    *zeroed_out = 1;
  }
  budget[0] = rab;
  budget[1] = rres;
  *res_out = RES;
  *err_out = ERR;
  *nreg_out = count;
  *active_out = active;
  free(lo); free(wd); free(est); free(er); free(ax); free(seq); free(heap);
}

void or_phase2_region(int key, const or_integrand* f, const double* lower, const double* width, double epsrel,
                      double epsabs, int pool_max, double* out)
{
  anyrule R;
  R.key = key == 9 ? 9 : 7;
  R.d = f->dim;
  or_rule_init(&R.r7, f->dim);
  if (R.key == 9) or_rule9_init(&R.r9, f->dim);
  double res, err, budget[2];
  int cnt, act, zero;
  phase2_region(&R, f, lower, width, epsrel, epsabs, pool_max, &res, &err, &cnt, &act, budget, &zero);
  out[0] = res;
  out[1] = err;
  out[2] = (double)cnt;
  out[3] = (double)act;
  out[4] = (double)zero;
  or_rule_free(&R.r7);
}

int or_integrate_two_phase(const or_integrand* f, const double* lows, const double* highs,
                           double epsrel, double epsabs, int key, int64_t phase1_max,
                           int pool_max, int nthreads, or_result* out)
{
  memset(out, 0, sizeof(*out));
  const int d = f->dim;
  if (key != 0 && key != 7 && key != 9) return -1;
  if (phase1_max < 1 || pool_max < 1) return -1;
  anyrule R;
  R.key = key == 9 ? 9 : 7;
  R.d = d;
  if (or_rule_init(&R.r7, d) != 0) return -1;
  if (R.key == 9 && or_rule9_init(&R.r9, d) != 0) { or_rule_free(&R.r7); return -1; }
  const int p = R.key == 9 ? R.r9.npts : R.r7.npts;
  int64_t n = or_initial_grid(d, lows, highs, NULL, NULL);
  int64_t cap = 2 * (phase1_max > n ? phase1_max : n) + 2;
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
  if (!lo || !wd || !lo2 || !wd2 || !est || !err || !ab || !ax || !act) { rc = -3; goto done2; }
  or_initial_grid(d, lows, highs, lo, wd);
  {
    ksum ifin = {0, 0}, efin = {0, 0};
    double fin_abs = 0.0, fin_aest = 0.0;
    int64_t nfin = 0;
    int phase2 = 0;
    for (int it = 0; it < 100; ++it) {
      int nonfinite = 0;
#ifdef _OPENMP
#pragma omp parallel for num_threads(nthreads > 0 ? nthreads : 1) schedule(static) reduction(| : nonfinite)
#endif
      for (int64_t r = 0; r < n; ++r) {
        double l[OR_MAXDIM], w[OR_MAXDIM];
        for (int m = 0; m < d; ++m) { l[m] = lo[(size_t)m * n + r]; w[m] = wd[(size_t)m * n + r]; }
        int a;
        any_eval(&R, f, l, w, &est[r], &err[r], &a, &ab[r]);
        ax[r] = a;
        if (!isfinite(est[r]) || !isfinite(err[r])) nonfinite = 1;
        act[r] = err[r] > fmax(epsrel * fabs(est[r]), epsabs) ? 1 : 0;
      }
      out->iterations = it + 1;
      out->nregions += n;
      out->neval += n * (int64_t)p;
      if (n > out->peak_regions) out->peak_regions = n;
      if (nonfinite) { rc = -2; goto done2; }
      ksum sall = ifin, eall = efin;
      double all_abs = fin_abs, all_aest = fin_aest;
      int64_t A = 0;
      for (int64_t r = 0; r < n; ++r) {
        kadd(&sall, est[r]);
        kadd(&eall, err[r]);
        all_abs += ab[r];
        all_aest += fabs(est[r]);
        if (!act[r]) {
          kadd(&ifin, est[r]);
          kadd(&efin, err[r]);
          fin_abs += ab[r];
          fin_aest += fabs(est[r]);
          ++nfin;
          out->nfinished++;
        } else {
          ++A;
        }
      }
      const double rG = kval(&sall), eG = kval(&eall);
      if (eG <= fmax(epsrel * fabs(rG), epsabs)) { /* converged in Phase 1 */
        out->integral = rG;
        out->error = eG;
        /* contributing: the regions finished before this iteration and all n of it */
        out->bound = or_fp64_bound(p, d, nfin + A, all_abs, all_aest, rG);
        out->fail = 0;
        rc = 0;
        goto done2;
      }
      out->integral = kval(&ifin);
      out->error = kval(&efin);
      if (A == 0) { phase2 = 0; break; } /* nothing left for Phase 2 */
      or_split(d, n, lo, wd, act, ax, lo2, wd2);
      double* t;
      t = lo; lo = lo2; lo2 = t;
      t = wd; wd = wd2; wd2 = t;
      n = 2 * A;
      phase2 = 1;
      if (n >= phase1_max) break;
    }
    if (phase2) {
      /* Phase 2 on the n (not yet evaluated) regions of the split list */
      double* pres = (double*)malloc(sizeof(double) * (size_t)n);
      double* perr = (double*)malloc(sizeof(double) * (size_t)n);
      int* pcnt = (int*)malloc(sizeof(int) * (size_t)n);
      int* pact = (int*)malloc(sizeof(int) * (size_t)n);
      int* pzero = (int*)malloc(sizeof(int) * (size_t)n);
      double* pbud = (double*)malloc(sizeof(double) * 2 * (size_t)n);
#ifdef _OPENMP
#pragma omp parallel for num_threads(nthreads > 0 ? nthreads : 1) schedule(dynamic, 4)
#endif
      for (int64_t r = 0; r < n; ++r) {
        double l[OR_MAXDIM], w[OR_MAXDIM];
        for (int m = 0; m < d; ++m) { l[m] = lo[(size_t)m * n + r]; w[m] = wd[(size_t)m * n + r]; }
        phase2_region(&R, f, l, w, epsrel, epsabs, pool_max, &pres[r], &perr[r], &pcnt[r], &pact[r],
                      &pbud[2 * r], &pzero[r]);
      }
      ksum I = ifin, E = efin;
      int64_t failed = 0, zeroed = 0;
      double ab_tot = fin_abs, res_round = 0.0, aest = fin_aest;
      for (int64_t r = 0; r < n; ++r) {
        kadd(&I, pres[r]);
        kadd(&E, perr[r]);
        out->nregions += pcnt[r];
        out->neval += (int64_t)(pcnt[r] - 1) * 2 * p + p;
        failed += pact[r];
        zeroed += pzero[r];
        ab_tot += pbud[2 * r];
        res_round += pbud[2 * r + 1];
        aest += fabs(pres[r]);
      }
      out->integral = kval(&I);
      out->error = kval(&E);
      {
        const double u = DBL_EPSILON / 2.0;
        out->bound = or_fp64_bound(p, d, nfin + n, ab_tot, aest, out->integral) +
                     u * res_round * (1.0 + 4.0 * ((double)pool_max + (double)n + 2.0) * u);
      }
      free(pres); free(perr); free(pcnt); free(pact); free(pzero); free(pbud);
      /* a zeroed region's contribution is dropped: the result is not certified */
      const int ok = out->error <= fmax(epsrel * fabs(out->integral), epsabs) && zeroed == 0;
      rc = ok ? 0 : 1;
      out->fail = (int32_t)(rc == 0 ? 0 : (failed > 0x7fffffff ? 0x7fffffff : (failed > 0 ? failed : 1)));
    } else {
      rc = out->error <= fmax(epsrel * fabs(out->integral), epsabs) ? 0 : 1;
      out->fail = rc;
    }
  }
done2:
  free(lo); free(wd); free(lo2); free(wd2); free(est); free(err); free(ab); free(ax); free(act);
  or_rule_free(&R.r7);
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
