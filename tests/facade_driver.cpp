// Caller of the C++ facade written the way the reference's driver
// (cudaCuhre.cu:25-87) calls GPUcuhre: construct once, integrate the same
// integrand over a Volume at two tolerances, zeroing the accumulated outputs
// in between, and print one tab-separated line per call.  With
// --two-phase it runs the reference's own method (degree-9 rule, Phase 1
// then Phase 2).  With --lclt FILE it integrates the reference's LC_LT_t
// (tables read from FILE: int32 rows, int32 cols, then float lt[rows],
// zt[cols], tau, mu, sigma, fmsk, fprj [rows*cols]) at the driver's
// tolerances 1e-12 and 1e-7 instead of the cosine sum.
#include <cstdio>
#include <cstring>
#include <vector>

#include "gpucuhre.hpp"

// The reference driver's LC_LT_t run (cudaCuhre.cu:43-79).
static int run_lclt(IntegralGPU::Cuhre& cuhre, const char* path)
{
  FILE* fh = std::fopen(path, "rb");
  if (!fh) return 2;
  int32_t rows = 0, cols = 0;
  if (std::fread(&rows, 4, 1, fh) != 1 || std::fread(&cols, 4, 1, fh) != 1 || rows < 2 || cols < 2) return 2;
  std::vector<float> buf((size_t)rows + cols + 5 * (size_t)rows * cols);
  if (std::fread(buf.data(), sizeof(float), buf.size(), fh) != buf.size()) return 2;
  std::fclose(fh);
  const float* lt = buf.data();
  const float* zt = lt + rows;
  const float* f0 = zt + cols;
  const size_t fs = (size_t)rows * cols;
  quad::integrands::LcLtT t(lt, rows, zt, cols, f0, f0 + fs, f0 + 2 * fs, f0 + 3 * fs, f0 + 4 * fs);
  float highs[3] = {2, .3f, 1};
  float lows[3] = {1, .1f, 0};
  quad::Volume vol(lows, highs, 3);
  for (double epsrel : {1e-12, 1e-7}) {
    double integral = 0, error = 0;
    size_t nregions = 0, neval = 0;
    const int flag = cuhre.integrate(epsrel, 1e-12, integral, error, nregions, neval, &t, &vol);
    std::printf("%d\t%e\t%.17g\t%.17g\t%zu\t%zu\t%d\n", 3, epsrel, integral, error, nregions, neval, flag);
  }
  return 0;
}

int main(int argc, char** argv)
{
  double epsrel = 1e-7, epsabs = 1e-12;
  double integral = 0, error = 0;
  size_t nregions = 0, neval = 0;
  int dim = 3;
  bool two_phase = false;
  const char* lclt_file = nullptr;
  for (int i = 1; i < argc; ++i) {
    if (std::strcmp(argv[i], "--two-phase") == 0) two_phase = true;
    if (std::strcmp(argv[i], "--lclt") == 0 && i + 1 < argc) lclt_file = argv[i + 1];
  }
  IntegralGPU::Cuhre cuhre(argc, argv, dim, two_phase ? 9 : 0, 0, 1);
  cuhre.two_phase = two_phase;
  if (lclt_file) return run_lclt(cuhre, lclt_file);
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
