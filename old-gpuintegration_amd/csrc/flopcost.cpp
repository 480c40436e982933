// flopcost -- one-operation FP64 kernels whose VALU instruction counts give
// the flop charge of div, exp, cos, sqrt, erfc, the library's reciprocal and
// its integer power (exponent 8) in the roofline convention.
// Run under rocprofv3 --pmc with the SQ_INSTS_VALU_{ADD,MUL,FMA,TRANS}_F64
// counters; per element flops = 64*(ADD+MUL+TRANS+2*FMA)/N minus the copy
// kernel's count.  It also times the issue cost of v_rcp_f64 against
// v_fma_f64 and v_mul_f64 (eight independent chains per lane, every CU
// full): the transcendental FP64 unit issues at a fraction of the FMA rate,
// which caps the fraction of the FP64 peak an integrand with a reciprocal
// per point can reach.  Not part of the library.
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

// OP 0: v_rcp_f64, 1: v_fma_f64, 2: v_mul_f64, each on 8 independent chains
template <int OP>
__global__ __launch_bounds__(256) void k_rate(double* out, double x0, int iters)
{
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  double a[8];
  for (int j = 0; j < 8; ++j) a[j] = x0 + 1e-9 * (t + j);
  for (int i = 0; i < iters; ++i) {
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      if constexpr (OP == 0)
        a[j] = __builtin_amdgcn_rcp(a[j]);
      else if constexpr (OP == 1)
        a[j] = fma(a[j], 0.999999, 1e-7);
      else
        a[j] = a[j] * 0.9999999;
    }
  }
  double s = 0;
  for (int j = 0; j < 8; ++j) s += a[j];
  out[t] = s;
}

// Times each k_rate kernel (second of two launches) and prints the wall
// time per wave instruction in CU cycles at the nominal 2.4 GHz.
static void issue_rates(double* out, int blocks)
{
  const int iters = 2048;
  hipEvent_t e0, e1;
  CHECK(hipEventCreate(&e0));
  CHECK(hipEventCreate(&e1));
  const char* names[3] = {"v_rcp_f64", "v_fma_f64", "v_mul_f64"};
  for (int op = 0; op < 3; ++op) {
    float ms = 0.f;
    for (int rep = 0; rep < 2; ++rep) {
      CHECK(hipEventRecord(e0));
      if (op == 0) hipLaunchKernelGGL(k_rate<0>, dim3(blocks), dim3(256), 0, 0, out, 1.3, iters);
      if (op == 1) hipLaunchKernelGGL(k_rate<1>, dim3(blocks), dim3(256), 0, 0, out, 1.3, iters);
      if (op == 2) hipLaunchKernelGGL(k_rate<2>, dim3(blocks), dim3(256), 0, 0, out, 1.3, iters);
      CHECK(hipEventRecord(e1));
      CHECK(hipEventSynchronize(e1));
      CHECK(hipEventElapsedTime(&ms, e0, e1));
    }
    const double wave_instr = (double)blocks * 256 / 64 * iters * 8;
    std::printf("%s: %.3f ms, %.2f CU-cycles per wave instruction (256 CUs, 2.4 GHz)\n", names[op], ms,
                ms * 1e-3 * 2.4e9 * 256 / wave_instr);
  }
  CHECK(hipEventDestroy(e0));
  CHECK(hipEventDestroy(e1));
}

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
  issue_rates(y, n / 1024);  // 4096 blocks of 256 threads: y holds one double per thread
  CHECK(hipFree(x));
  CHECK(hipFree(y));
  return 0;
}
