// Caller of the C++ facade written the way the reference's driver
// (cudaCuhre.cu:25-87) calls GPUcuhre: construct once, integrate the same
// integrand over a Volume at two tolerances, zeroing the accumulated outputs
// in between, and print one tab-separated line per call.  With
// --two-phase it runs the reference's own method (degree-9 rule, Phase 1
// then Phase 2).
#include <cstdio>
#include <cstring>

#include "gpucuhre.hpp"

int main(int argc, char** argv)
{
  double epsrel = 1e-7, epsabs = 1e-12;
  double integral = 0, error = 0;
  size_t nregions = 0, neval = 0;
  int dim = 3;
  bool two_phase = false;
  for (int i = 1; i < argc; ++i)
    if (std::strcmp(argv[i], "--two-phase") == 0) two_phase = true;
  IntegralGPU::Cuhre cuhre(argc, argv, dim, two_phase ? 9 : 0, 0, 1);
  cuhre.two_phase = two_phase;
  quad::integrands::CosineSum t(dim);
  float highs[3] = {2, .3f, 1};
  float lows[3] = {1, .1f, 0};
  quad::Volume vol(lows, highs, 3);
  int errorFlag = cuhre.integrate(epsrel, epsabs, integral, error, nregions, neval, &t, &vol);
  std::printf("%d\t%e\t%.17g\t%.17g\t%zu\t%zu\t%d\n", dim, epsrel, integral, error, nregions, neval,
              errorFlag);
  integral = 0;
  error = 0;
  nregions = 0;
  neval = 0;
  errorFlag = cuhre.integrate(1e-10, epsabs, integral, error, nregions, neval, &t, &vol);
  std::printf("%d\t%e\t%.17g\t%.17g\t%zu\t%zu\t%d\n", dim, 1e-10, integral, error, nregions, neval,
              errorFlag);
  // accumulation semantics: a second call without zeroing adds up
  double acc = 0, acc_err = 0;
  size_t acc_r = 0, acc_n = 0;
  cuhre.integrate(1e-7, epsabs, acc, acc_err, acc_r, acc_n, &t, &vol);
  cuhre.integrate(1e-7, epsabs, acc, acc_err, acc_r, acc_n, &t, &vol);
  std::printf("acc\t%.17g\t%zu\n", acc, acc_r);
  return 0;
}
