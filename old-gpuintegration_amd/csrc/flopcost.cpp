// flopcost -- one-operation FP64 kernels whose VALU instruction counts give
// the flop charge of div, exp, cos, sqrt, erfc, the library's reciprocal and
// its integer power (exponent 8) in the roofline convention.
// Run under rocprofv3 --pmc with the SQ_INSTS_VALU_{ADD,MUL,FMA,TRANS}_F64
// counters; per element flops = 64*(ADD+MUL+TRANS+2*FMA)/N minus the copy
// kernel's count.  Not part of the library.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <vector>

#include "integrands.hpp"

#define CHECK(x)                                                             \
  do {                                                                       \
    hipError_t e = (x);                                                      \
    if (e != hipSuccess) {                                                   \
      std::fprintf(stderr, "%s: %s\n", #x, hipGetErrorString(e));            \
      std::exit(1);                                                          \
    }                                                                        \
  } while (0)

__global__ void k_copy(const double* x, double* y, int n) { int i = blockIdx.x * blockDim.x + threadIdx.x; if (i < n) y[i] = x[i]; }
__global__ void k_div(const double* x, double* y, int n) { int i = blockIdx.x * blockDim.x + threadIdx.x; if (i < n) y[i] = 1.0 / x[i]; }
__global__ void k_exp(const double* x, double* y, int n) { int i = blockIdx.x * blockDim.x + threadIdx.x; if (i < n) y[i] = exp(-x[i]); }
__global__ void k_cos(const double* x, double* y, int n) { int i = blockIdx.x * blockDim.x + threadIdx.x; if (i < n) y[i] = cos(x[i]); }
__global__ void k_sqrt(const double* x, double* y, int n) { int i = blockIdx.x * blockDim.x + threadIdx.x; if (i < n) y[i] = sqrt(x[i]); }
__global__ void k_erfc(const double* x, double* y, int n) { int i = blockIdx.x * blockDim.x + threadIdx.x; if (i < n) y[i] = erfc(x[i] - 2.0); }
__global__ void k_recip(const double* x, double* y, int n) { int i = blockIdx.x * blockDim.x + threadIdx.x; if (i < n) y[i] = cuhre::recip_dev(x[i]); }
__global__ void k_ipow(const double* x, double* y, int n, int e) { int i = blockIdx.x * blockDim.x + threadIdx.x; if (i < n) y[i] = cuhre::ipow_dev(x[i], e); }

int main()
{
  const int n = 1 << 22;  // multiple of 256: every lane active
  std::vector<double> h(n);
  for (int i = 0; i < n; ++i) h[i] = 0.5 + 3.0 * (double)i / n;  // args in [0.5, 3.5)
  double *x, *y;
  CHECK(hipMalloc(&x, n * sizeof(double)));
  CHECK(hipMalloc(&y, n * sizeof(double)));
  CHECK(hipMemcpy(x, h.data(), n * sizeof(double), hipMemcpyHostToDevice));
  dim3 g(n / 256), b(256);
  hipLaunchKernelGGL(k_copy, g, b, 0, 0, x, y, n);
  hipLaunchKernelGGL(k_div, g, b, 0, 0, x, y, n);
  hipLaunchKernelGGL(k_exp, g, b, 0, 0, x, y, n);
  hipLaunchKernelGGL(k_cos, g, b, 0, 0, x, y, n);
  hipLaunchKernelGGL(k_sqrt, g, b, 0, 0, x, y, n);
  hipLaunchKernelGGL(k_recip, g, b, 0, 0, x, y, n);
  hipLaunchKernelGGL(k_ipow, g, b, 0, 0, x, y, n, 8);
  hipLaunchKernelGGL(k_erfc, g, b, 0, 0, x, y, n);  // erfc over [-1.5, 1.5)
  CHECK(hipDeviceSynchronize());
  std::printf("flopcost n=%d done\n", n);
  CHECK(hipFree(x));
  CHECK(hipFree(y));
  return 0;
}
