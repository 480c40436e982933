// host_logic.cpp -- the parts of the engine that need no device: error
// reporting, the rule tables, the flop model, the initial grid and static
// workspace arithmetic, argument checks and the C-ABI queries built on them.
// Built with the host compiler into the library, and with AddressSanitizer
// and UndefinedBehaviorSanitizer into the host check (tests/sanitize_host.cpp).
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>

#include "context.hpp"

namespace cuhre {

// ------------------------------------------------------------------ errors

static thread_local std::string g_err;

int fail(int code, const char* fmt, ...)
{
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

int set_error(int code, const char* msg) { return fail(code, "%s", msg); }


// ------------------------------------------------------------------ rule

int gm_points(int d) { return (1 << d) + 2 * d * d + 2 * d + 1; }

// Genz-Malik (1980) degree-7 rule and embedded degree-5 rule; weights on
// [-1,1]^d divided by the volume 2^d.
int r9_points(int d)
{
  return 1 + 8 * d + 2 * d * (d - 1) + 4 * d * (d - 1) + 4 * d * (d - 1) * (d - 2) / 3 + (1 << d);
}

// The reference's degree-9 rule (KEY 9): radii and weights as
// quad/GPUquad/GPUQuadRule.cu:33-176 defines them (they define the rule),
// null-rule scale/norm as :209-225.
static Rule9Const make_rule9(int d)
{
  Rule9Const R;
  const double n = (double)d;
  const double a[5] = {.47795365790226950619, .20302858736911986780, .44762735462617812882, .125,
                       .34303789878087814570};
  for (int k = 0; k < 5; ++k) R.t[k] = 2.0 * a[k];  // fractions of the width -> half-width units
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
  const double w[9][5] = {
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
  const int cnt[9] = {1, 2 * d, 2 * d, 2 * d, 2 * d, 2 * d * (d - 1), 4 * d * (d - 1),
                      4 * d * (d - 1) * (d - 2) / 3, 1 << d};
  for (int s = 0; s < 9; ++s)
    for (int q = 0; q < 5; ++q) {
      R.w[s][q] = w[s][q];
      R.scale[s][q] = 0.0;
      R.norm[s][q] = 0.0;
    }
  for (int s = 0; s < 9; ++s)
    for (int q = 1; q < 4; ++q) {
      const double sc = w[s][q] == 0.0 ? 100.0 : -w[s][q + 1] / w[s][q];
      double sum = 0.0;
      for (int x = 0; x < 9; ++x) sum += (double)cnt[x] * std::fabs(w[x][q + 1] + sc * w[x][q]);
      R.scale[s][q] = sc;
      R.norm[s][q] = 1.0 / sum;
    }
  R.ratio = (R.t[1] / R.t[0]) * (R.t[1] / R.t[0]);
  return R;
}

RuleConst make_rule(int d, int key)
{
  RuleConst R;
  std::memset(&R, 0, sizeof R);
  const double dd = (double)d;
  R.l2 = std::sqrt(9.0 / 70.0);
  R.l3 = std::sqrt(9.0 / 10.0);
  R.l4 = std::sqrt(9.0 / 10.0);
  R.l5 = std::sqrt(9.0 / 19.0);
  R.w7c = (12824.0 - 9120.0 * dd + 400.0 * dd * dd) / 19683.0;
  R.w7a2 = 980.0 / 6561.0;
  R.w7a3 = (1820.0 - 400.0 * dd) / 19683.0;
  R.w7p = 200.0 / 19683.0;
  R.w7v = 6859.0 / 19683.0 / (double)(1 << d);
  R.w5c = (729.0 - 950.0 * dd + 50.0 * dd * dd) / 729.0;
  R.w5a2 = 245.0 / 486.0;
  R.w5a3 = (265.0 - 100.0 * dd) / 1458.0;
  R.w5p = 25.0 / 729.0;
  R.ratio = 1.0 / 7.0;
  R.key = key == 9 ? 9 : 7;
  R.npts = key == 9 ? r9_points(d) : gm_points(d);
  if (key == 9) R.r9 = make_rule9(d);
  return R;
}

// Algorithmic FP64 flops of one region's evaluation for a family, by the
// counting convention add/sub/mul = 1, fma = 2, and div / exp / cos / sqrt
// charged the FP64 VALU flops of their gfx950 instruction sequences
// (measured with the SQ_INSTS_VALU_*_F64 counters on a one-op kernel).
// FP64 flops per call (add/sub/mul 1, fma 2, transcendental estimate 1) of
// the device math each family uses, counted with the SQ_INSTS_VALU_*_F64
// counters on one-operation kernels (flopcost.cpp); recip is recip_dev.
struct FlopCosts { double div = 14, exp = 27, cos = 57, sqrt = 17, recip = 9, erfc = 115; };
static FlopCosts g_costs;

// multiplies of b^e: ipow_const for Genz's exponent d + 1, else the loop
double ipow_flops(int e, int d)
{
  if (e <= 0) return 0.0;
  int sq = 0, bits = 0;
  for (int t = e; t > 1; t >>= 1) ++sq;
  for (int t = e; t; t >>= 1) bits += t & 1;
  return e == d + 1 ? (double)(sq + bits - 1) : (double)(sq + bits);
}

double family_flops(const cuhre_integrand* f, int d)
{
  switch (f->family) {
  case CUHRE_GAUSSIAN: return 4.0 * d + g_costs.exp;
  case CUHRE_PRODUCT_PEAK: return 4.0 * d - 1 + g_costs.recip;
  case CUHRE_OSCILLATORY: return 2.0 * d + 1 + g_costs.cos;
  case CUHRE_CORNER_PEAK: return 2.0 * d + ipow_flops(f->exponent, d) + g_costs.recip;
  case CUHRE_CONTINUOUS: return 4.0 * d + g_costs.exp;
  case CUHRE_COSINE_SUM: return d * (2.0 + g_costs.cos + g_costs.div) + 1;
  case CUHRE_TABLE6: return 20.0 + g_costs.exp + g_costs.sqrt + 2 * g_costs.div;
  // five bilinear fields (8 flops each, 2 interpolation weights), the closed
  // form's ~40 flops, 3 exp, 4 erfc and 11 divisions
  case CUHRE_LC_LT_T: return 84.0 + 3 * g_costs.exp + 4 * g_costs.erfc + 11 * g_costs.div;
  default: return 4.0 * d;
  }
}

// Per region: every point costs its integrand plus d coordinate FMAs.
// GM7/5: the sums take three FMAs (I7, I5, sum |w f|) at centre/axis/pair
// points and two at the corners (no degree-5 weight).  Degree 9: two adds
// per point (f and |f| into the set group), six FMAs per set group, the
// null-rule combination.  Plus the per-region setup.
double region_flops(const cuhre_integrand* f, int d, int key)
{
  if (key == 9) {
    const double p = r9_points(d);
    const double groups = 1 + 4.0 * d + d * (d - 1) + d * (d - 1) * (d - 2) / 6.0 + 1;
    return p * (family_flops(f, d) + 2.0 * d + 2.0) + 12.0 * groups + 3 * 9 * 4.0 + (8.0 * d + 12.0);
  }
  const double p = gm_points(d), corners = (double)(1 << d);
  return p * (family_flops(f, d) + 2.0 * d) + 6.0 * (p - corners) + 4.0 * corners + (8.0 * d + 12.0);
}

int grid_k(int d) { return d <= 4 ? 4 : (d <= 10 ? 2 : 1); }

// Initial uniform k^d grid over the box (GPUKernelquad.cu:23-61, 300-316),
// region r has coordinate index (r / k^m) % k along axis m.
int64_t initial_grid_count(int d)
{
  const int k = grid_k(d);
  int64_t n = 1;
  for (int m = 0; m < d; ++m) n *= k;
  return n;
}


size_t al256(size_t b) { return (b + 255) & ~(size_t)255; }

// Bytes of the list buffers (two ping-pong lists, est/err/abssum/axis, block
// partials and offsets with their group tails) for cap regions of dim axes.
size_t scratch_list_bytes(int64_t cap, int dim)
{
  const int64_t nb = block_capacity(cap), ng = group_capacity(nb);
  return 4 * al256(sizeof(double) * (size_t)cap * dim) + 3 * al256(sizeof(double) * (size_t)cap) +
         al256((size_t)cap) + al256(sizeof(Partial) * (size_t)(nb + ng)) + al256(sizeof(int64_t) * (size_t)(nb + ng));
}

// Regions per list the workspace holds at dim (a multiple of 256).
int64_t scratch_capacity(int dim)
{
  const size_t avail = kScratchBytes - kScratchTableBytes;
  int64_t lo = 0, hi = (int64_t)(avail / (4 * sizeof(double) * (size_t)dim)) / 256 + 1;
  while (hi - lo > 1) {
    const int64_t mid = (lo + hi) / 2;
    if (scratch_list_bytes(mid * 256, dim) <= avail) lo = mid;
    else hi = mid;
  }
  return lo * 256;
}


// Integrand, dimension and rule key of a call (the checks that need no device).
int check_integrand_args(const cuhre_integrand* f, int dim, int key)
{
  if (!f) return fail(CUHRE_ERR_ARG, "null integrand");
  if (dim < 2 || dim > 10) return fail(CUHRE_ERR_UNSUPPORTED, "dimension %d outside 2..10", dim);
  if (f->dim != dim) return fail(CUHRE_ERR_ARG, "integrand dim %d != dim %d", f->dim, dim);
  if (key != 0 && key != 7 && key != 9)
    return fail(CUHRE_ERR_UNSUPPORTED, "rule key %d not available (0 or 7: Genz-Malik 7/5; 9: degree 9)", key);
  return 0;
}

// Box and tolerances of a call.
int check_bounds(int dim, double epsrel, double epsabs, const double* lows, const double* highs)
{
  if (!lows || !highs) return fail(CUHRE_ERR_ARG, "null bounds");
  if (!(epsrel >= 0.0) || !(epsabs >= 0.0)) return fail(CUHRE_ERR_ARG, "tolerances must be >= 0");
  for (int m = 0; m < dim; ++m)
    if (!(highs[m] > lows[m]) || !std::isfinite(lows[m]) || !std::isfinite(highs[m]))
      return fail(CUHRE_ERR_ARG, "empty or non-finite box along axis %d", m);
  return 0;
}

}  // namespace cuhre

using namespace cuhre;

extern "C" {

const char* cuhre_last_error(void) { return g_err.c_str(); }

int64_t cuhre_points_per_region(int dim, int key)
{
  if (dim < 2 || dim > kMaxDim || (key != 0 && key != 7 && key != 9)) return -1;
  return key == 9 ? r9_points(dim) : gm_points(dim);
}


int64_t cuhre_noalloc_capacity(int dim)
{
  if (dim < 2 || dim > 10) return -1;
  return scratch_capacity(dim);
}


}  // extern "C"
