// gpucuhre.hpp -- C++ drop-in facade over the C-ABI (include/cuhre.h).
//
// Mirrors the reference's host interface so existing callers keep their
// shape:
//   quad::GPUcuhre<T>(argc, argv, dim, key, verbose, numDevices)
//       quad/GPUquad/GPUquad.h:46-63
//   int integrate(epsrel, epsabs, T& integral, T& error, size_t& nregions,
//                 size_t& neval, integT* f, Volume* vol = NULL)
//       quad/GPUquad/GPUquad.h:611-727
//   quad::Volume(float* lows, float* highs, int dim)
//       quad/GPUquad/GPUQuadVolume.cu:4-25
//   IntegralGPU::Cuhre typedef, CuhreInteface.h:9-12
//
// Semantics kept from the reference: integral, error, nregions and neval
// are accumulated into (+=), so callers zero them between calls
// (cudaCuhre.cu:67-70); the return value is 0 when the global tolerance is
// met.  Differences: no CUDA device reset, no MPI (numDevices > 1 is
// rejected), a Cuba-style `maxeval` member bounds the work (default 1e9),
// negative returns are error codes (cuhre_last_error() has the text), and
// the integrand is described by a compiled-in family: integT must provide
// `cuhre_integrand descriptor() const` (the functors below do).
#pragma once

#include <cstddef>
#include <cstdlib>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "cuhre.h"

namespace quad {

// Integration box.  The reference stores float bounds in textures; here the
// bounds are kept in double (float inputs are widened exactly).
struct Volume {
  double lows[CUHRE_MAXDIM];
  double highs[CUHRE_MAXDIM];
  int dim;
  Volume(const float* l, const float* h, int d) : dim(d)
  {
    for (int i = 0; i < d && i < CUHRE_MAXDIM; ++i) {
      lows[i] = l[i];
      highs[i] = h[i];
    }
  }
  Volume(const double* l, const double* h, int d) : dim(d)
  {
    for (int i = 0; i < d && i < CUHRE_MAXDIM; ++i) {
      lows[i] = l[i];
      highs[i] = h[i];
    }
  }
};

template <typename T>
class GPUcuhre;

template <>
class GPUcuhre<double> {
 public:
  long long maxeval = 1000000000LL;
  // The reference's method end to end (GPUquad.h:611-727): with key 9 and
  // two_phase set, integrate() runs its breadth-first Phase 1 (at most 100
  // iterations, until phase1_regions regions; 0: 2^14) and its per-region
  // Phase 2 (at most pool_max bisections per region; 0: 2048), and returns
  // its flag (the number of Phase-2 regions that missed their own test when
  // the global test fails).  maxeval does not apply to this method.
  bool two_phase = false;
  long long phase1_regions = 0;
  int pool_max = 0;

  // argv flags as in the reference's CommandLineArgs (quad/util/cudaUtil.h:
  // 37-56): --device=N selects the GPU, --v turns on verbose output.
  GPUcuhre(int argc, char** argv, int dim, int key = 0, int verbose = 0, int numDevices = 1)
      : dim_(dim), key_(key), verbose_(verbose)
  {
    if (numDevices != 1)
      throw std::invalid_argument("GPUcuhre: one device per engine (numDevices must be 1)");
    int device = 0;
    for (int i = 1; i < argc; ++i) {
      if (std::strncmp(argv[i], "--device=", 9) == 0) device = std::atoi(argv[i] + 9);
      if (std::strcmp(argv[i], "--v") == 0) verbose_ = 1;
    }
    if (cuhre_context_create(&ctx_, device, 0) != 0)
      throw std::runtime_error(std::string("GPUcuhre: ") + cuhre_last_error());
  }
  ~GPUcuhre() { cuhre_context_destroy(ctx_); }
  GPUcuhre(const GPUcuhre&) = delete;
  GPUcuhre& operator=(const GPUcuhre&) = delete;

  template <typename integT>
  int integrate(double epsrel, double epsabs, double& integral, double& error, size_t& nregions,
                size_t& neval, integT* f, Volume* vol = nullptr)
  {
    const cuhre_integrand d = f->descriptor();
    double lo[CUHRE_MAXDIM], hi[CUHRE_MAXDIM];
    for (int i = 0; i < dim_; ++i) {
      lo[i] = vol ? vol->lows[i] : 0.0;
      hi[i] = vol ? vol->highs[i] : 1.0;
    }
    cuhre_result r;
    const int rc = two_phase ? cuhre_integrate_two_phase(ctx_, &d, dim_, epsrel, epsabs, key_, phase1_regions,
                                                         pool_max, lo, hi, &r)
                             : cuhre_integrate(ctx_, &d, dim_, epsrel, epsabs, maxeval, key_, lo, hi, &r);
    if (rc < 0) return rc;
    integral += r.integral;
    error += r.error;
    nregions += (size_t)r.nregions;
    neval += (size_t)r.neval;
    last_ = r;
    return r.fail;
  }

  // Everything the engine reported about the last call (rounding bound,
  // iterations, timings).
  const cuhre_result& last_result() const { return last_; }

 private:
  cuhre_context* ctx_ = nullptr;
  int dim_, key_, verbose_;
  cuhre_result last_{};
};

// Integrand functors for the compiled-in families.
namespace integrands {

inline cuhre_integrand blank(int family, int dim)
{
  cuhre_integrand d;
  std::memset(&d, 0, sizeof d);
  d.family = family;
  d.dim = dim;
  return d;
}

struct GaussianPeak {
  cuhre_integrand d;
  GaussianPeak(int dim, const double* a, const double* w) : d(blank(CUHRE_GAUSSIAN, dim))
  {
    for (int i = 0; i < dim; ++i) { d.a[i] = a[i]; d.w[i] = w[i]; }
  }
  cuhre_integrand descriptor() const { return d; }
};

struct ProductPeak {
  cuhre_integrand d;
  ProductPeak(int dim, const double* a, const double* w) : d(blank(CUHRE_PRODUCT_PEAK, dim))
  {
    for (int i = 0; i < dim; ++i) { d.a[i] = a[i]; d.w[i] = w[i]; }
  }
  cuhre_integrand descriptor() const { return d; }
};

struct Oscillatory {
  cuhre_integrand d;
  Oscillatory(int dim, const double* a, double u) : d(blank(CUHRE_OSCILLATORY, dim))
  {
    for (int i = 0; i < dim; ++i) d.a[i] = a[i];
    d.u = u;
  }
  cuhre_integrand descriptor() const { return d; }
};

struct CornerPeak {
  cuhre_integrand d;
  CornerPeak(int dim, const double* a, int exponent) : d(blank(CUHRE_CORNER_PEAK, dim))
  {
    for (int i = 0; i < dim; ++i) d.a[i] = a[i];
    d.exponent = exponent;
  }
  cuhre_integrand descriptor() const { return d; }
};

// The reference's LC_LT_t<IntegralGPU> (lc_lt_t.h:7-259): five fields
// tau, mu, sigma, fmsk, fprj tabulated on float nodes in lt (rows) and zt
// (columns), each array laid out as the reference's (value at (lt_i, zt_j)
// at index j*nrows + i); x = (lt, zt, lc).  The constructor takes the
// arrays the reference's constructor hands to Interp2D::Initialize.
struct LcLtT {
  std::vector<double> packed;
  cuhre_integrand d;
  LcLtT(const float* lt, int rows, const float* zt, int cols, const float* tau, const float* mu,
        const float* sigma, const float* fmsk, const float* fprj)
      : d(blank(CUHRE_LC_LT_T, 3))
  {
    packed.assign(lt, lt + rows);
    packed.insert(packed.end(), zt, zt + cols);
    for (const float* f : {tau, mu, sigma, fmsk, fprj}) packed.insert(packed.end(), f, f + (size_t)rows * cols);
    d.table = packed.data();
    d.table_nx = rows;
    d.table_ny = cols;
  }
  LcLtT(const LcLtT&) = delete;
  LcLtT& operator=(const LcLtT&) = delete;
  cuhre_integrand descriptor() const { return d; }
};

// The reference driver's test functor: -(1/2) sum cos(10 x_i)/0.054402111088937
// (cudaCuhre.cu:9-23).
struct CosineSum {
  cuhre_integrand d;
  explicit CosineSum(int dim) : d(blank(CUHRE_COSINE_SUM, dim)) {}
  cuhre_integrand descriptor() const { return d; }
};

}  // namespace integrands
}  // namespace quad

// Facade typedef of the reference (CuhreInteface.h:9-12).
struct IntegralGPU {
  typedef quad::GPUcuhre<double> Cuhre;
};
