// learn.hip -- the learned split guide: network training and inference on
// the MFMA units, sample collection from plain runs, and the guided loop.
//
// No reference counterpart (the reference always splits on the largest
// fourth difference, quad/GPUquad/GPUQuadSample.cu:150-172, and finishes a
// region only by its local test, quad/GPUquad/GPUQuadPhases.cu:63-73).
//
// The guide sees 25 features of a region (d = 7): its position and depth per
// axis in the box, its fourth differences, its degree-7 estimate against the
// mean density, its error against the volume share of the tolerance, the
// relative I7 - I5 gap and its total depth.  It proposes
//   * a split axis (trained to imitate the look-ahead axis: the bisection
//     whose two children have the smallest summed Genz-Malik error), used
//     when its softmax confidence passes a threshold, else the fourth
//     difference decides as in the plain method;
//   * a finish flag (trained on regions whose error fits inside their volume
//     share of epsrel * |I|, I being the run's final estimate), honoured only
//     when the region's error really fits inside its volume share of
//     budget * tol (tol = max(epsrel * |current estimate|, epsabs)).
// The retirements of one iteration are taken all or none: only when the
// retired error of the run stays within budget * tol and the frozen error
// (finished plus retired) within (1 - budget) * tol, so the regions still
// being split keep budget * tol or more for themselves and a run whose every
// region is retired has met the global test.  The Genz-Malik error sum stays
// the only convergence test.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <vector>

#include "common.hpp"
#include "context.hpp"
#include "eval_kernels.hpp"
#include "mlp.hpp"
#include "reduce.hpp"

namespace cuhre {

int set_error(int code, const char* msg);

// ---------------------------------------------------------------- training

// Master FP32 parameters -> BF16 W ([K][N]) and W^T ([N][K]) copies and the
// training image (whose padding the caller has zeroed).
__global__ void k_mlp_prepare(const float* params, uint16_t* W, uint16_t* WT, uint16_t* img)
{
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= kMlpParams) return;
  if (p >= kMlpBOff[0]) {
    reinterpret_cast<float*>(reinterpret_cast<char*>(img) + kImgBiasByte)[p - kMlpBOff[0]] = params[p];
    return;
  }
  int l = 3;
  while (p < kMlpWOff[l]) --l;
  const int K = kMlpK[l], N = kMlpN[l];
  const int q = p - kMlpWOff[l];
  const int k = q / N, n = q - k * N;
  const uint16_t h = f2bf(params[p]);
  W[kMlpWOff[l] + k * N + n] = h;
  WT[kMlpWOff[l] + n * K + k] = h;
  img[img_wt(l, k, n)] = h;
  img[img_w(l, k, n)] = h;
}

// Inference on rows given as FP32 features (n x 32) -> logits (n x 16): one
// wave per 64 rows, the same network pass as the guide step's.
__global__ __launch_bounds__(64) void k_mlp_forward(MlpDev M, const float* feat, int64_t n, float* out)
{
  __shared__ __align__(16) uint16_t A[kMlpWaveBuf], B[kMlpWaveBuf];
  const int64_t r0 = (int64_t)blockIdx.x * kMlpRows;
  for (int e = threadIdx.x; e < kMlpRows * kMlpIn; e += kWave) {
    const int rr = e / kMlpIn, q = e % kMlpIn;
    const int64_t r = r0 + rr;
    A[rr * kLdIn + q] = r < n ? f2bf(feat[r * kMlpIn + q]) : (uint16_t)0;
  }
  __syncwarp();
  mlp_forward64_wave(M, A, B, 0xFu);
  const float* Y = reinterpret_cast<const float*>(A);
  for (int e = threadIdx.x; e < kMlpRows * kMlpOut; e += kWave) {
    const int64_t r = r0 + e / kMlpOut;
    if (r < n) out[r * kMlpOut + e % kMlpOut] = Y[e];
  }
}

// dH[16 x K] = dZ[16 x N] . W^T for this wave's rows, then dZprev = dH * (H > 0)
// stored as BF16.  W row-major [K][N] with row stride LDW (the training
// image in LDS) gives contiguous B operands; dZ, H and dZprev are LDS rows
// with padded strides.
template <int K, int N, int LDW>
__device__ __forceinline__ void mlp_layer_bwd(const uint16_t* dZs, const uint16_t* W, const uint16_t* Hs,
                                              uint16_t* dZprev)
{
  const int l = __lane_id(), row = l & 15, kq = (l >> 4) * 4;
  f32x4 acc[K / 16];
#pragma unroll
  for (int jt = 0; jt < K / 16; ++jt) acc[jt] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int kt = 0; kt < N / 16; ++kt) {
    const bf16x4 a = *reinterpret_cast<const bf16x4*>(dZs + row * (N + kPad) + kt * 16 + kq);
#pragma unroll
    for (int jt = 0; jt < K / 16; ++jt) {
      const bf16x4 b = *reinterpret_cast<const bf16x4*>(W + (jt * 16 + row) * LDW + kt * 16 + kq);
      acc[jt] = mfma16(a, b, acc[jt]);
    }
  }
#pragma unroll
  for (int jt = 0; jt < K / 16; ++jt)
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int idx = (kq + v) * (K + kPad) + jt * 16 + row;
      dZprev[idx] = Hs[idx] != 0 ? f2bf(acc[jt][v]) : (uint16_t)0;
    }
}

// One 16x16 tile of dW = H^T dZ over the block's 64 rows (4 MFMA steps);
// ldh, ldz: LDS row strides of H and dZ.
__device__ __forceinline__ f32x4 mlp_dw_tile(const uint16_t* Hs, int ldh, const uint16_t* dZs, int ldz, int mt, int nt)
{
  const int l = __lane_id(), row = l & 15, kq = (l >> 4) * 4;
  f32x4 acc = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int kt = 0; kt < kMlpRows / 16; ++kt) {
    bf16x4 a, b;
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int s = kt * 16 + kq + v;
      a[v] = (short)Hs[s * ldh + mt * 16 + row];
      b[v] = (short)dZs[s * ldz + nt * 16 + row];
    }
    acc = mfma16(a, b, acc);
  }
  return acc;
}

// SplitMix64 finaliser: sample hashes for logging and axis relabelling.
__host__ __device__ __forceinline__ uint64_t mix64(uint64_t z)
{
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// A training block's slot in the partial buffer: the gradient over its rows
// (kMlpParams), then its rows' summed loss, axis hits and finish hits.
constexpr int kPartStride = kMlpParams + 4;

struct TrainArgs {
  const uint4* img;      // training image of the weights (kImgBytes)
  const float* feat;     // samples x 32
  const int8_t* axis;    // samples
  const float* finish;   // samples
  const int32_t* order;  // batch rows -> sample index; null: row r is sample
                         // shuffle_index(shuf, first + r)
  Shuffle shuf;
  int64_t first;
  int64_t rows;          // rows in this batch
  int dim;
  float inv_batch;
  float* part;           // blocks x kPartStride
  uint64_t perm_seed;    // 0: samples as logged; else a random axis relabelling per row
};

__device__ __forceinline__ int64_t train_sample(const TrainArgs& T, int64_t r)
{
  return T.order != nullptr ? (int64_t)T.order[r] : (int64_t)shuffle_index(T.shuf, (uint64_t)(T.first + r));
}

// Forward + backward of 64 rows per block, 16 per wavefront; the block's
// weight gradient goes to its own slot (reduced in a fixed order by
// k_mlp_reduce_adam).  The weights are staged in LDS by one linear copy of
// the training image, issued together with the row gathers; the forward and
// backward passes are wave-local, and one block barrier orders the weight
// gradient after them.
__global__ __launch_bounds__(256) void k_mlp_train(TrainArgs T)
{
  constexpr int kChunks = kImgBytes / 16, kFull = kChunks / 256, kTail = kChunks % 256;
  __shared__ __align__(16) uint4 s_img[kChunks];
  __shared__ __align__(16) uint16_t X[kMlpRows * kLdIn];
  __shared__ __align__(16) uint16_t H1[kMlpRows * kLdHid], H2[kMlpRows * kLdHid], H3[kMlpRows * kLdHid];
  __shared__ __align__(16) float Y[kMlpRows * kMlpOut];
  __shared__ __align__(16) uint16_t dZ4[kMlpRows * kLdOut];
  __shared__ __align__(16) uint16_t dZ3[kMlpRows * kLdHid], dZ2[kMlpRows * kLdHid], dZ1[kMlpRows * kLdHid];
  __shared__ int8_t s_dst[kMlpRows][8];  // logged axis m of row rr sits at new axis s_dst[rr][m]
  __shared__ int s_label[kMlpRows];
  __shared__ float s_fin[kMlpRows];
  __shared__ float s_st[4][3];
  const int t = threadIdx.x, w = t >> 6, lane = t & 63;
  const int64_t r0 = (int64_t)blockIdx.x * kMlpRows;
  // thread t gathers features 8 (t & 3) .. 8 (t & 3) + 7 of row t >> 2 (a
  // wave gathers its own 16 rows)
  const int gr = t >> 2, gq = (t & 3) * 8;
  float4 fa = {0.f, 0.f, 0.f, 0.f}, fb = {0.f, 0.f, 0.f, 0.f};
  if (r0 + gr < T.rows) {
    const float4* src = reinterpret_cast<const float4*>(T.feat + train_sample(T, r0 + gr) * kMlpIn + gq);
    fa = src[0];
    fb = src[1];
  }
  uint4 wv[kFull], wt = {0u, 0u, 0u, 0u};
#pragma unroll
  for (int i = 0; i < kFull; ++i) wv[i] = T.img[t + i * 256];
  if (t < kTail) wt = T.img[kFull * 256 + t];
  if (t < kMlpRows) {
    // Axis relabelling (the rule is symmetric under permuting the axes, so
    // is the label): position, width and fourth-difference features of
    // logged axis m move to new axis dst[m], the label with them.
    // Fisher-Yates from a per-sample hash.
    const int64_t r = r0 + t;
    uint32_t pm = 0x76543210u;  // nibble j: the logged axis new axis j holds
    int lab = 0;
    float fin = 0.f;
    if (r < T.rows) {
      const int64_t smp = train_sample(T, r);
      lab = T.axis[smp];
      fin = T.finish[smp];
      if (T.perm_seed != 0) {
        uint64_t h = mix64(T.perm_seed ^ (uint64_t)smp * 0x9E3779B97F4A7C15ull);
        for (int j = T.dim - 1; j > 0; --j) {
          const int k = (int)(((h >> 32) * (uint64_t)(j + 1)) >> 32);  // uniform in [0, j]
          h = mix64(h);
          const uint32_t a = (pm >> (4 * j)) & 15u, b = (pm >> (4 * k)) & 15u;
          pm &= ~((15u << (4 * j)) | (15u << (4 * k)));
          pm |= (b << (4 * j)) | (a << (4 * k));
        }
      }
    }
    for (int j = 0; j < 8; ++j) s_dst[t][(pm >> (4 * j)) & 15u] = (int8_t)j;
    s_label[t] = s_dst[t][lab];
    s_fin[t] = fin;
  }
#pragma unroll
  for (int i = 0; i < kFull; ++i) s_img[t + i * 256] = wv[i];
  if (t < kTail) s_img[kFull * 256 + t] = wt;
  __syncthreads();
  const uint16_t* img = reinterpret_cast<const uint16_t*>(s_img);
  const float* bias = reinterpret_cast<const float*>(reinterpret_cast<const char*>(s_img) + kImgBiasByte);
  {
    const float v[8] = {fa.x, fa.y, fa.z, fa.w, fb.x, fb.y, fb.z, fb.w};
    uint16_t* xr = X + gr * kLdIn;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      const int q = gq + i;
      const int dst = (q < 24 && (q & 7) < T.dim) ? ((q & ~7) | s_dst[gr][q & 7]) : q;
      xr[dst] = f2bf(v[i]);
    }
  }
  __syncwarp();
  mlp_forward16_wave(img, bias, X + w * 16 * kLdIn, H1 + w * 16 * kLdHid, H2 + w * 16 * kLdHid,
                     H3 + w * 16 * kLdHid, Y + w * 16 * kMlpOut);
  // loss and dY of the wave's rows (softmax cross-entropy over the d axis
  // logits + BCE finish), one row per lane 0..15
  float st_l = 0.f, st_a = 0.f, st_f = 0.f;
  if (lane < 16) {
    const int rr = w * 16 + lane;
    const int64_t r = r0 + rr;
    uint16_t* dz = dZ4 + rr * kLdOut;
    for (int j = 0; j < kMlpOut; ++j) dz[j] = 0;
    if (r < T.rows) {
      const float* y = Y + rr * kMlpOut;
      const int a = s_label[rr];
      const float f = s_fin[rr];
      float m = y[0];
      int am = 0;
      for (int j = 1; j < T.dim; ++j)
        if (y[j] > m) {
          m = y[j];
          am = j;
        }
      float se = 0.f;
      for (int j = 0; j < T.dim; ++j) se += expf(y[j] - m);
      const float lse = m + logf(se);
      const float z = y[kMlpFinish];
      const float q = 1.f / (1.f + expf(-z));
      const float bce = fmaxf(z, 0.f) - z * f + log1pf(expf(-fabsf(z)));
      st_l = (lse - y[a]) + bce;
      st_a = am == a ? 1.f : 0.f;
      st_f = (q >= 0.5f) == (f >= 0.5f) ? 1.f : 0.f;
      for (int j = 0; j < T.dim; ++j) {
        const float pj = expf(y[j] - lse);
        dz[j] = f2bf((pj - (j == a ? 1.f : 0.f)) * T.inv_batch);
      }
      dz[kMlpFinish] = f2bf((q - f) * T.inv_batch);
    }
  }
#pragma unroll
  for (int o = 8; o > 0; o >>= 1) {
    st_l += __shfl_xor(st_l, o);
    st_a += __shfl_xor(st_a, o);
    st_f += __shfl_xor(st_f, o);
  }
  if (lane == 0) {
    s_st[w][0] = st_l;
    s_st[w][1] = st_a;
    s_st[w][2] = st_f;
  }
  __syncwarp();
  mlp_layer_bwd<64, 16, 16 + kPad>(dZ4 + w * 16 * kLdOut, img + kImgW[3], H3 + w * 16 * kLdHid,
                                   dZ3 + w * 16 * kLdHid);
  __syncwarp();
  mlp_layer_bwd<64, 64, 64 + kPad>(dZ3 + w * 16 * kLdHid, img + kImgW[2], H2 + w * 16 * kLdHid,
                                   dZ2 + w * 16 * kLdHid);
  __syncwarp();
  mlp_layer_bwd<64, 64, 64 + kPad>(dZ2 + w * 16 * kLdHid, img + kImgW[1], H1 + w * 16 * kLdHid,
                                   dZ1 + w * 16 * kLdHid);
  __syncthreads();
  // weight gradients: 44 tiles of 16x16 shared out over the 4 waves
  float* part = T.part + (int64_t)blockIdx.x * kPartStride;
  const int row = lane & 15, kq = (lane >> 4) * 4;
  for (int tl = w; tl < 44; tl += 4) {
    int layer, mt, nt;
    if (tl < 8) { layer = 0; mt = tl / 4; nt = tl % 4; }
    else if (tl < 24) { layer = 1; mt = (tl - 8) / 4; nt = (tl - 8) % 4; }
    else if (tl < 40) { layer = 2; mt = (tl - 24) / 4; nt = (tl - 24) % 4; }
    else { layer = 3; mt = tl - 40; nt = 0; }
    const uint16_t* Hs = layer == 0 ? X : (layer == 1 ? H1 : (layer == 2 ? H2 : H3));
    const uint16_t* dZs = layer == 0 ? dZ1 : (layer == 1 ? dZ2 : (layer == 2 ? dZ3 : dZ4));
    const int K = kMlpK[layer], N = kMlpN[layer];
    const f32x4 acc = mlp_dw_tile(Hs, K + kPad, dZs, N + kPad, mt, nt);
#pragma unroll
    for (int v = 0; v < 4; ++v) part[kMlpWOff[layer] + (mt * 16 + kq + v) * N + nt * 16 + row] = acc[v];
  }
  // bias gradients: column sums of the BF16 upstream gradients
  for (int c = t; c < kMlpParams - kMlpBOff[0]; c += blockDim.x) {
    int layer = 3;
    while (c < kMlpBOff[layer] - kMlpBOff[0]) --layer;
    const int N = kMlpN[layer], col = c - (kMlpBOff[layer] - kMlpBOff[0]);
    const uint16_t* dZs = layer == 0 ? dZ1 : (layer == 1 ? dZ2 : (layer == 2 ? dZ3 : dZ4));
    float sacc = 0.f;
    for (int s = 0; s < kMlpRows; ++s) sacc += bf2f(dZs[s * (N + kPad) + col]);
    part[kMlpBOff[0] + c] = sacc;
  }
  if (t < 3) part[kMlpParams + t] = ((s_st[0][t] + s_st[1][t]) + s_st[2][t]) + s_st[3][t];
}

// grad[p] = sum over the training blocks' partials in a fixed order (each of
// kRedWaves wavefronts sums every kRedWaves-th block for 64 parameters, then
// wavefront 0 adds the kRedWaves sums in order); optional Adam step on the
// FP32 master parameters followed by the BF16 copies (W, W^T and the
// training image).  With `acc` set, one extra block adds the blocks' loss
// and hit counts, in a fixed order, into acc[0..2].
constexpr int kRedWaves = 8;

struct AdamArgs {
  const float* part;
  int nblocks;
  float* grad;
  float* params;
  float* m;
  float* v;
  int step;
  float lr;
  int do_update;
  uint16_t* W;
  uint16_t* WT;
  uint16_t* img;
  double* acc;
};

__global__ __launch_bounds__(kRedWaves * 64) void k_mlp_reduce_adam(AdamArgs A)
{
  __shared__ float s_g[kRedWaves][64];
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  if (A.acc != nullptr && blockIdx.x == gridDim.x - 1) {
    if (w != 0) return;
    double s[3] = {0.0, 0.0, 0.0};
    for (int b = lane; b < A.nblocks; b += 64)
      for (int j = 0; j < 3; ++j) s[j] += A.part[(int64_t)b * kPartStride + kMlpParams + j];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1)
      for (int j = 0; j < 3; ++j) s[j] += shfl_xor_d(s[j], o);
    if (lane < 3) A.acc[lane] += s[lane];
    return;
  }
  const int p = blockIdx.x * 64 + lane;
  float g = 0.f;
  if (p < kMlpParams) {
    const float* q = A.part + p;
    int b = w;
    // the same left-to-right sum as one block at a time, its loads batched
    for (; b + 3 * kRedWaves < A.nblocks; b += 4 * kRedWaves) {
      const float x0 = q[(int64_t)b * kPartStride], x1 = q[(int64_t)(b + kRedWaves) * kPartStride];
      const float x2 = q[(int64_t)(b + 2 * kRedWaves) * kPartStride];
      const float x3 = q[(int64_t)(b + 3 * kRedWaves) * kPartStride];
      g += x0;
      g += x1;
      g += x2;
      g += x3;
    }
    for (; b < A.nblocks; b += kRedWaves) g += q[(int64_t)b * kPartStride];
  }
  s_g[w][lane] = g;
  __syncthreads();
  if (w != 0 || p >= kMlpParams) return;
  g = s_g[0][lane];
  for (int j = 1; j < kRedWaves; ++j) g += s_g[j][lane];
  A.grad[p] = g;
  if (!A.do_update) return;
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  const float mm = b1 * A.m[p] + (1.f - b1) * g;
  const float vv = b2 * A.v[p] + (1.f - b2) * g * g;
  A.m[p] = mm;
  A.v[p] = vv;
  const float mh = mm / (1.f - powf(b1, (float)A.step));
  const float vh = vv / (1.f - powf(b2, (float)A.step));
  const float np = A.params[p] - A.lr * mh / (sqrtf(vh) + eps);
  A.params[p] = np;
  if (p < kMlpBOff[0]) {
    int l = 3;
    while (p < kMlpWOff[l]) --l;
    const int K = kMlpK[l], N = kMlpN[l];
    const int qq = p - kMlpWOff[l];
    const int k = qq / N, n = qq - k * N;
    const uint16_t h = f2bf(np);
    A.W[kMlpWOff[l] + k * N + n] = h;
    A.WT[kMlpWOff[l] + n * K + k] = h;
    A.img[img_wt(l, k, n)] = h;
    A.img[img_w(l, k, n)] = h;
  } else {
    reinterpret_cast<float*>(reinterpret_cast<char*>(A.img) + kImgBiasByte)[p - kMlpBOff[0]] = np;
  }
}


// ---------------------------------------------------------------- features

struct GuideArgs {
  const double* lower;
  const double* width;
  int64_t ld;
  int64_t n;
  const double* est;
  const double* err;
  const double* i5;
  const double* dvec;
  const double* abssum;
  int8_t* axis;    // in/out: the fourth-difference axis, replaced by a confident guide
  int8_t* retire;  // out: 1 = retired as finished, not split
  int dim;
  double lows[kMaxDim], highs[kMaxDim];
  double absI;  // |current estimate of the integral|
  double absE;  // sum |est| over the finished regions and the iteration's (normalized allotment)
  double fin_aest;  // sum |est| of the regions finished before this iteration
  double epsrel, epsabs;
  double tol;   // max(epsrel * absI, epsabs)
  const IterState* state;  // when set: absI and tol come from this iteration's state,
                           // and a stopped iteration (stop != 0) skips the guide step
  double loc_epsrel, loc_epsabs;  // the per-region test (0, 0 under the shared allotment)
  int axis_mode;    // 0: fourth difference, 1: guide when confident
  int finish_mode;  // 0: none, 1: guide proposal + volume-share guard, 2: guard alone,
                    // 3: shared value / volume allotment replaces the per-region test,
                    // 4: the same normalised so the tolerances add up to epsrel |I|
  int use_mlp;
  float axis_conf, finish_thr;
  double budget;  // fraction of epsrel*|I| the retired regions may spend
  MlpDev M;
  int64_t* tile_cnt;     // regions split per 64-region tile
  int64_t* tile_logcnt;  // regions logged per tile (collection)
  double* tile_ret;      // per tile: retired est (s,c), err (s,c), abssum, |est|, count
  uint64_t log_stride;   // 0: no collection
  uint64_t log_salt;
};

constexpr int kRetStride = 8;

struct GuideSums {
  int64_t n_split, n_retired, n_logged, n_overridden;
  double est_s, est_c, err_s, err_c, abs_r, aest_r;
};

// What one guided iteration hands the host through coherent host memory:
// the iteration's state and the guide step's sums, valid once the sequence
// word next to it holds the iteration's number.
struct GuidePub {
  IterState st;
  GuideSums sums;
};

__device__ __forceinline__ void publish_guided(GuidePub* pub, uint32_t* pub_seq, uint32_t seq, const IterState& st,
                                               const GuideSums& g)
{
  pub->st = st;
  pub->sums = g;
  __hip_atomic_store(pub_seq, seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
}

__device__ __forceinline__ bool guide_selected(const GuideArgs& G, int64_t r, bool active)
{
  return active && G.log_stride > 0 && (mix64((uint64_t)r ^ G.log_salt) % G.log_stride) == 0;
}

// log2 of a positive double to float precision: its binary exponent plus
// the float log2 of its mantissa in [0.5, 1).
__device__ __forceinline__ float log2_mant(double x)
{
  int ex;
  const double m = frexp(x, &ex);
  return (float)ex + log2f((float)m);
}

// The 32 input features of region r (25 used at d = 7) and its relative
// volume.  The axis loops run over the guide's largest dimension, 8, with
// the unused axes masked, so the loads of all axes go out together.
__device__ __forceinline__ double guide_features(const GuideArgs& G, double absI, int64_t r, float* f)
{
  constexpr int kAx = 8;
  constexpr float kLog10Of2 = 0.30102999566398120f;
  const int d = G.dim;
  double lo[kAx], wd[kAx], dv[kAx];
#pragma unroll
  for (int m = 0; m < kAx; ++m)
    if (m < d) {
      lo[m] = G.lower[m * G.ld + r];
      wd[m] = G.width[m * G.ld + r];
      dv[m] = G.dvec[m * G.ld + r];
    }
  double vrel = 1.0, maxd = 0.0;
#pragma unroll
  for (int j = 0; j < 32; ++j) f[j] = 0.f;
#pragma unroll
  for (int m = 0; m < kAx; ++m)
    if (m < d) {
      const double span = G.highs[m] - G.lows[m];
      const double wr = wd[m] / span;
      vrel *= wr;
      f[m] = (float)((lo[m] - G.lows[m]) / span);
      f[8 + m] = -log2_mant(wr) / 16.f;
      maxd = fmax(maxd, dv[m]);
    }
  const double inv = maxd > 0.0 ? 1.0 / maxd : 0.0;
#pragma unroll
  for (int m = 0; m < kAx; ++m)
    if (m < d) f[16 + m] = (float)(dv[m] * inv);
  const double e = G.est[r], er = G.err[r], i5 = G.i5[r];
  const double scale = absI * vrel;
  f[24] = log2_mant(fabs(e) / scale + 1e-30) * (kLog10Of2 / 8.f);
  f[25] = log2_mant(er / (G.epsrel * scale) + 1e-30) * (kLog10Of2 / 8.f);
  f[26] = (float)((e - i5) / (fabs(e) + fabs(i5) + 1e-300));
  f[27] = -log2_mant(vrel) / 64.f;
  return vrel;
}

// Guide step of one iteration, one 64-region tile per wave: features (a
// region per lane), network, axis / finish decisions, per-tile counts and
// retired sums.  With MLP a block is one wave, which runs the network on its
// tile in LDS of its own with no block barrier (16-row groups whose regions
// are all inactive skip the arithmetic); without, a block holds four tiles.
template <bool MLP>
constexpr int guide_tiles()
{
  return MLP ? 1 : 4;
}

template <bool MLP>
__global__ __launch_bounds__(64 * guide_tiles<MLP>()) void k_guide_apply(GuideArgs G)
{
  constexpr int kL = MLP ? 1 : 0;  // network buffers only when the network runs
  __shared__ __align__(16) uint16_t BufA[kMlpWaveBuf * kL + 1], BufB[kMlpWaveBuf * kL + 1];
  if (G.state != nullptr && G.state->stop != 0) return;  // block-uniform
  double absI = G.absI, tol = G.tol, absE = G.absE;
  if (G.state != nullptr) {
    absI = fabs(G.state->rG_s + G.state->rG_c);
    tol = fmax(G.epsrel * absI, G.epsabs);
    absE = G.fin_aest + G.state->tot.aest_all;
  }
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int64_t ntiles = (G.n + kMlpRows - 1) / kMlpRows;
  const int64_t tile = (int64_t)blockIdx.x * guide_tiles<MLP>() + w;
  const int64_t r = tile * kMlpRows + lane;
  bool active = false;
  double vrel = 0.0;
  int am = 0;                 // network's axis
  float conf = 0.f, q = 0.f;  // its softmax confidence, finish probability
  if (MLP) {
    float f[32];
    for (int j = 0; j < 32; ++j) f[j] = 0.f;
    if (r < G.n) {
      const double e = G.est[r], er = G.err[r];
      active = isfinite(e) && isfinite(er) && region_active(e, er, G.loc_epsrel, G.loc_epsabs);
      vrel = guide_features(G, absI, r, f);
    }
    uint16_t* xr = BufA + lane * kLdIn;
    for (int j = 0; j < 32; ++j) xr[j] = f2bf(f[j]);
    const uint64_t act = __ballot(active);
    unsigned work = 0;
    for (int s = 0; s < kMlpRows / 16; ++s)
      if ((act >> (16 * s)) & 0xFFFFull) work |= 1u << s;
    __syncwarp();
    mlp_forward64_wave(G.M, BufA, BufB, work);
    if (active) {
      const float* y = reinterpret_cast<const float*>(BufA) + lane * kMlpOut;
      float m = y[0];
      for (int j = 1; j < G.dim; ++j)
        if (y[j] > m) {
          m = y[j];
          am = j;
        }
      float se = 0.f;
      for (int j = 0; j < G.dim; ++j) se += expf(y[j] - m);
      conf = 1.f / se;  // softmax probability of the argmax
      q = 1.f / (1.f + expf(-y[kMlpFinish]));
    }
  } else if (r < G.n) {
    const double e = G.est[r], er = G.err[r];
    active = isfinite(e) && isfinite(er) && region_active(e, er, G.loc_epsrel, G.loc_epsabs);
    if (G.finish_mode >= 2) {
      vrel = 1.0;
      for (int m = 0; m < G.dim; ++m) vrel *= G.width[m * G.ld + r] / (G.highs[m] - G.lows[m]);
    }
  }
  int retire = 0, ovr = 0;
  double e = 0.0, er = 0.0, ab = 0.0;
  if (active) {
    e = G.est[r];
    er = G.err[r];
    const bool guard = er <= G.budget * tol * vrel;
    if (MLP) {
      if (G.axis_mode == 1 && conf >= G.axis_conf && G.axis[r] != am) {
        G.axis[r] = (int8_t)am;
        ovr = 1;
      }
      if (G.finish_mode == 1 && q >= G.finish_thr && guard) retire = 1;
    }
    if (G.finish_mode == 2 && guard) retire = 1;
    if (G.finish_mode == 3) {
      // shared allotment: a (1 - budget) share of tol by value, a budget share by volume
      const double a = G.epsrel * ((1.0 - G.budget) * fabs(e) + G.budget * absI * vrel);
      if (er <= fmax(a, G.epsabs)) retire = 1;
    }
    if (G.finish_mode == 4) {
      // normalised allotment: value shares |est| / E and volume shares of
      // epsrel |I|, so the tolerances add up to the global one even when the
      // integrand's signs cancel (|I| << E); for f >= 0 and budget 0 it is
      // the plain test
      const double a = G.epsrel * absI * ((1.0 - G.budget) * fabs(e) / absE + G.budget * vrel);
      if (er <= fmax(a, G.epsabs)) retire = 1;
    }
    ab = G.abssum[r];
  }
  if (r < G.n) G.retire[r] = (int8_t)retire;
  // the tile's counts and retired sums (a butterfly, the same order on every run)
  const uint64_t b_split = __ballot(active && !retire), b_log = __ballot(guide_selected(G, r, active));
  const uint64_t b_ovr = __ballot(ovr != 0), b_ret = __ballot(retire != 0);
  double es = retire ? e : 0.0, ec = 0.0, rs = retire ? er : 0.0, rc = 0.0;
  double abr = retire ? ab : 0.0, ae = retire ? fabs(e) : 0.0;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    dd_add(es, ec, shfl_xor_d(es, o), shfl_xor_d(ec, o));
    dd_add(rs, rc, shfl_xor_d(rs, o), shfl_xor_d(rc, o));
    abr += shfl_xor_d(abr, o);
    ae += shfl_xor_d(ae, o);
  }
  if (lane == 0 && tile < ntiles) {
    G.tile_cnt[tile] = __popcll(b_split);
    G.tile_logcnt[tile] = __popcll(b_log);
    double* tr = G.tile_ret + tile * kRetStride;
    tr[0] = es;
    tr[1] = ec;
    tr[2] = rs;
    tr[3] = rc;
    tr[4] = abr;
    tr[5] = ae;
    tr[6] = (double)__popcll(b_ret);
    tr[7] = (double)__popcll(b_ovr);
  }
}

// Retirement room of one iteration: computed on the device from the
// iteration's state when `state` is set, else `room` as given.
struct RoomArgs {
  const IterState* state;
  double room;
  double frozen_prev;  // finished + retired error before this iteration
  double ret_prev;     // retired error before this iteration
  double budget, epsrel, epsabs;
  int mode;            // finish mode
};

// Scan of the per-tile counts and retired sums, in three launches over
// chunks of kScanChunk tiles (every sum in a fixed order, so the result does
// not depend on scheduling):
//   k_gscan_part  per chunk: tile counts and retired sums;
//   k_gscan_top   one block: the retirement decision (honoured, *retire_on
//                 = 1, only when the retired error fits in the room; else
//                 every active region is split and the retired sums are
//                 zero), chunk offsets and the iteration's totals;
//   k_gscan_tiles per chunk: tile offsets = chunk offset + scan in chunk.
constexpr int kScanPer = 4;                   // tiles per thread
constexpr int kScanChunk = 256 * kScanPer;    // tiles per chunk

struct ChunkPart {
  int64_t cnt, all, log, pad;  // split with / without retirements, logged
  double es, ec, rs, rc, ab, ae, nr, no;
};

__device__ __forceinline__ void cp_zero(ChunkPart& q)
{
  q.cnt = q.all = q.log = q.pad = 0;
  q.es = q.ec = q.rs = q.rc = q.ab = q.ae = q.nr = q.no = 0.0;
}

__device__ __forceinline__ void cp_add(ChunkPart& a, const ChunkPart& b)
{
  a.cnt += b.cnt;
  a.all += b.all;
  a.log += b.log;
  dd_add(a.es, a.ec, b.es, b.ec);
  dd_add(a.rs, a.rc, b.rs, b.rc);
  a.ab += b.ab;
  a.ae += b.ae;
  a.nr += b.nr;
  a.no += b.no;
}

// the partner lane's ChunkPart (xor butterfly)
__device__ __forceinline__ ChunkPart cp_shfl_xor(const ChunkPart& q, int mask)
{
  ChunkPart r;
  r.cnt = __shfl_xor((long long)q.cnt, mask, kWave);
  r.all = __shfl_xor((long long)q.all, mask, kWave);
  r.log = __shfl_xor((long long)q.log, mask, kWave);
  r.pad = 0;
  r.es = shfl_xor_d(q.es, mask);
  r.ec = shfl_xor_d(q.ec, mask);
  r.rs = shfl_xor_d(q.rs, mask);
  r.rc = shfl_xor_d(q.rc, mask);
  r.ab = shfl_xor_d(q.ab, mask);
  r.ae = shfl_xor_d(q.ae, mask);
  r.nr = shfl_xor_d(q.nr, mask);
  r.no = shfl_xor_d(q.no, mask);
  return r;
}

__device__ __forceinline__ void cp_tile(ChunkPart& q, const int64_t* cnt, const int64_t* logcnt, const double* ret,
                                        int64_t t)
{
  const double* tr = ret + t * kRetStride;
  q.cnt += cnt[t];
  q.all += cnt[t] + (int64_t)tr[6];
  q.log += logcnt[t];
  dd_add(q.es, q.ec, tr[0], tr[1]);
  dd_add(q.rs, q.rc, tr[2], tr[3]);
  q.ab += tr[4];
  q.ae += tr[5];
  q.nr += tr[6];
  q.no += tr[7];
}

__device__ __forceinline__ bool scan_stopped(const IterState* st) { return st != nullptr && st->stop != 0; }

__global__ __launch_bounds__(256) void k_gscan_part(const int64_t* cnt, const int64_t* logcnt, const double* ret,
                                                    int64_t ntiles, const IterState* state, ChunkPart* part)
{
  if (scan_stopped(state)) return;
  __shared__ ChunkPart sp[256];
  ChunkPart q;
  cp_zero(q);
  const int64_t t0 = (int64_t)blockIdx.x * kScanChunk + (int64_t)threadIdx.x * kScanPer;
  for (int k = 0; k < kScanPer; ++k)
    if (t0 + k < ntiles) cp_tile(q, cnt, logcnt, ret, t0 + k);
  sp[threadIdx.x] = q;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) cp_add(sp[threadIdx.x], sp[threadIdx.x + o]);
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = sp[0];
}

// With `pub` set (and RA.state), it also publishes the state and the sums
// to the host under sequence number `seq`, stopped iterations included.
__global__ __launch_bounds__(256) void k_gscan_top(const ChunkPart* part, int64_t nchunks, RoomArgs RA,
                                                   int64_t* choff, int64_t* chlog, GuideSums* out, int* retire_on,
                                                   GuidePub* pub, uint32_t* pub_seq, uint32_t seq)
{
  if (scan_stopped(RA.state)) {
    if (pub != nullptr && threadIdx.x == 0) {
      GuideSums z = {};
      publish_guided(pub, pub_seq, seq, *RA.state, z);
    }
    return;
  }
  __shared__ ChunkPart sp[kMaxWaves];
  __shared__ int64_t sa[kMaxWaves], sb[kMaxWaves], sc[kMaxWaves];
  __shared__ int s_allow;
  double room = RA.room;
  if (RA.state != nullptr) {
    const IterState& st = *RA.state;
    const double tol = fmax(RA.epsrel * fabs(st.rG_s + st.rG_c), RA.epsabs);
    const double frozen = RA.frozen_prev + (st.tot.ferr_s + st.tot.ferr_c);
    room = RA.mode >= 3   ? 1e300
           : RA.mode != 0 ? fmin(RA.budget * tol - RA.ret_prev, (1.0 - RA.budget) * tol - frozen)
                          : -1.0;
  }
  const int T = blockDim.x, W = T / kWave;
  const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
  const int64_t per = (nchunks + T - 1) / T;
  const int64_t c0 = threadIdx.x * per;
  const int64_t c1 = c0 + per < nchunks ? c0 + per : nchunks;
  ChunkPart q;
  cp_zero(q);
  for (int64_t c = c0; c < c1; ++c) cp_add(q, part[c]);
  // inclusive scans of the three counts within the wave, the wave's
  // ChunkPart total by a butterfly (a fixed order), then across the waves
  long long ia = q.cnt, ib = q.all, ic = q.log;
#pragma unroll
  for (int off = 1; off < kWave; off <<= 1) {
    const long long va = __shfl_up(ia, off, kWave), vb = __shfl_up(ib, off, kWave), vc = __shfl_up(ic, off, kWave);
    if (lane >= off) {
      ia += va;
      ib += vb;
      ic += vc;
    }
  }
  ChunkPart t = q;
#pragma unroll
  for (int mask = 1; mask < kWave; mask <<= 1) cp_add(t, cp_shfl_xor(t, mask));
  if (lane == kWave - 1) {
    sa[wave] = ia;
    sb[wave] = ib;
    sc[wave] = ic;
  }
  if (lane == 0) sp[wave] = t;
  __syncthreads();
  int64_t pa = ia - q.cnt, pb = ib - q.all, pc = ic - q.log;  // exclusive, within the wave
  for (int w = 0; w < wave; ++w) {
    pa += sa[w];
    pb += sb[w];
    pc += sc[w];
  }
  if (threadIdx.x == 0) {
    ChunkPart tot = sp[0];
    for (int w = 1; w < W; ++w) cp_add(tot, sp[w]);
    const int allow = tot.nr > 0.0 && tot.rs + tot.rc <= room;
    GuideSums g;
    g.n_logged = tot.log;
    g.n_overridden = (int64_t)tot.no;
    if (allow) {
      g.n_split = tot.cnt;
      g.est_s = tot.es;
      g.est_c = tot.ec;
      g.err_s = tot.rs;
      g.err_c = tot.rc;
      g.abs_r = tot.ab;
      g.aest_r = tot.ae;
      g.n_retired = (int64_t)tot.nr;
    } else {
      g.n_split = tot.all;
      g.est_s = g.est_c = g.err_s = g.err_c = g.abs_r = g.aest_r = 0.0;
      g.n_retired = 0;
    }
    *out = g;
    *retire_on = allow;
    s_allow = allow;
    if (pub != nullptr) publish_guided(pub, pub_seq, seq, *RA.state, g);
  }
  __syncthreads();
  const int allow = s_allow;
  int64_t ra = allow ? pa : pb;
  int64_t rb = pc;
  for (int64_t c = c0; c < c1; ++c) {
    choff[c] = ra;
    chlog[c] = rb;
    ra += allow ? part[c].cnt : part[c].all;
    rb += part[c].log;
  }
}

__global__ __launch_bounds__(256) void k_gscan_tiles(const int64_t* cnt, const int64_t* logcnt, const double* ret,
                                                     int64_t ntiles, const IterState* state, const int64_t* choff,
                                                     const int64_t* chlog, const int* retire_on, int64_t* off,
                                                     int64_t* logoff)
{
  if (scan_stopped(state)) return;
  __shared__ int64_t sa[256], sb[256];
  const int allow = *retire_on;
  const int64_t t0 = (int64_t)blockIdx.x * kScanChunk + (int64_t)threadIdx.x * kScanPer;
  int64_t ma = 0, mb = 0;
  for (int k = 0; k < kScanPer; ++k) {
    const int64_t t = t0 + k;
    if (t < ntiles) {
      ma += allow ? cnt[t] : cnt[t] + (int64_t)ret[t * kRetStride + 6];
      mb += logcnt[t];
    }
  }
  sa[threadIdx.x] = ma;
  sb[threadIdx.x] = mb;
  __syncthreads();
  for (int o = 1; o < 256; o <<= 1) {
    const int64_t va = (int)threadIdx.x >= o ? sa[threadIdx.x - o] : 0;
    const int64_t vb = (int)threadIdx.x >= o ? sb[threadIdx.x - o] : 0;
    __syncthreads();
    sa[threadIdx.x] += va;
    sb[threadIdx.x] += vb;
    __syncthreads();
  }
  int64_t ra = choff[blockIdx.x] + sa[threadIdx.x] - ma;
  int64_t rb = chlog[blockIdx.x] + sb[threadIdx.x] - mb;
  for (int k = 0; k < kScanPer; ++k) {
    const int64_t t = t0 + k;
    if (t < ntiles) {
      off[t] = ra;
      logoff[t] = rb;
      ra += allow ? cnt[t] : cnt[t] + (int64_t)ret[t * kRetStride + 6];
      rb += logcnt[t];
    }
  }
}

struct CollectArgs {
  GuideArgs G;
  const int64_t* logoff;
  int64_t base;     // dataset rows before this iteration
  float* feat;      // dataset features
  double* s_err;    // dataset: region error
  double* s_vrel;   // dataset: relative volume
  double* la_lo;    // look-ahead children, SoA with stride la_ld
  double* la_wd;
  int64_t la_ld;
};

// Writes the logged regions' features and their 2d look-ahead children
// (each axis bisected) in deterministic order (tile offset + rank).
__global__ __launch_bounds__(64) void k_collect_write(CollectArgs C)
{
  const GuideArgs& G = C.G;
  const int64_t r = (int64_t)blockIdx.x * kMlpRows + threadIdx.x;
  bool active = false;
  if (r < G.n) {
    const double e = G.est[r], er = G.err[r];
    active = isfinite(e) && isfinite(er) && region_active(e, er, G.loc_epsrel, G.loc_epsabs);
  }
  const bool sel = guide_selected(G, r, active);
  const uint64_t mask = __ballot(sel);
  const int lane = threadIdx.x;
  const uint64_t below = lane ? (~0ull >> (kWave - lane)) : 0ull;
  if (!sel) return;
  const int64_t j = C.logoff[blockIdx.x] + __popcll(mask & below);
  float f[32];
  const double vrel = guide_features(G, G.absI, r, f);
  const int64_t s = C.base + j;
  for (int q = 0; q < 32; ++q) C.feat[s * 32 + q] = f[q];
  C.s_err[s] = G.err[r];
  C.s_vrel[s] = vrel;
  const int d = G.dim;
  for (int a = 0; a < d; ++a)
    for (int c = 0; c < 2; ++c) {
      const int64_t k = j * 2 * d + 2 * a + c;
      for (int m = 0; m < d; ++m) {
        const double lo = G.lower[m * G.ld + r], w = G.width[m * G.ld + r];
        const double hw = 0.5 * w;
        C.la_lo[m * C.la_ld + k] = m == a ? (c ? lo + hw : lo) : lo;
        C.la_wd[m * C.la_ld + k] = m == a ? hw : w;
      }
    }
}

// Axis label of each logged region: the bisection whose two children have the
// smallest summed error (first minimum).
__global__ void k_label_axis(const double* la_err, int64_t nlog, int d, int64_t base, int8_t* s_axis)
{
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= nlog) return;
  double best = 0.0;
  int ba = 0;
  for (int a = 0; a < d; ++a) {
    const double e = la_err[j * 2 * d + 2 * a] + la_err[j * 2 * d + 2 * a + 1];
    if (a == 0 || e < best) {
      best = e;
      ba = a;
    }
  }
  s_axis[base + j] = (int8_t)ba;
}

// Finish label of a run's samples: the region's error fits in its volume
// share of epsrel * |I| (I = the run's final estimate).
__global__ void k_label_finish(const double* s_err, const double* s_vrel, int64_t b0, int64_t b1, double tol,
                               float* s_fin)
{
  const int64_t s = b0 + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= b1) return;
  s_fin[s] = s_err[s] <= tol * s_vrel[s] ? 1.f : 0.f;
}

}  // namespace cuhre

using namespace cuhre;

// ------------------------------------------------------------------ host

struct cuhre_guide {
  cuhre_context* ctx = nullptr;
  float* params = nullptr;  // FP32 master
  float* m = nullptr;
  float* v = nullptr;
  uint16_t* W = nullptr;
  uint16_t* WT = nullptr;
  uint16_t* img = nullptr;  // training image (kImgBytes)
  float* grad = nullptr;
  float* part = nullptr;
  int part_blocks = 0;
  int adam_step = 0;
  // dataset
  int64_t ds_cap = 0, ds_n = 0;
  float* ds_feat = nullptr;
  int8_t* ds_axis = nullptr;
  float* ds_fin = nullptr;
  double* ds_err = nullptr;
  double* ds_vrel = nullptr;
  // look-ahead buffers
  int64_t la_cap = 0;
  double *la_lo = nullptr, *la_wd = nullptr, *la_est = nullptr, *la_err = nullptr, *la_ab = nullptr;
  int8_t* la_axis = nullptr;
  // per-iteration buffers of the guided loop
  int64_t reg_cap = 0;
  int8_t* retire = nullptr;
  int64_t *tile_cnt = nullptr, *tile_off = nullptr, *tile_logcnt = nullptr, *tile_logoff = nullptr;
  double* tile_ret = nullptr;
  GuideSums* d_sums = nullptr;
  int* d_retire_on = nullptr;  // the scan's verdict on this iteration's retirements
  ChunkPart* chunk_part = nullptr;  // scan chunks: partial sums and offsets
  int64_t *chunk_off = nullptr, *chunk_log = nullptr;
  GuideSums* h_sums = nullptr;
  // the guided loop's published iteration (coherent host memory)
  GuidePub* pub = nullptr;
  uint32_t* pub_seq = nullptr;
  GuidePub* pub_dev = nullptr;
  uint32_t* pub_seq_dev = nullptr;
  uint32_t seq = 0;
  // training scratch
  int64_t train_cap = 0;
  int32_t* order = nullptr;  // explicit row order (cuhre_guide_gradient)
  double* d_stats = nullptr;  // per epoch: loss, axis hits, finish hits, pad
  int stats_cap = 0;          // epochs d_stats holds
};

namespace {

void gfree(void* p)
{
  if (p) (void)hipFree(p);
}

int guide_reserve_regions(cuhre_guide* g, int64_t cap)
{
  if (cap <= g->reg_cap) return 0;
  cap = std::max(cap, 2 * g->reg_cap);
  gfree(g->retire);
  gfree(g->tile_cnt);
  gfree(g->tile_off);
  gfree(g->tile_logcnt);
  gfree(g->tile_logoff);
  gfree(g->tile_ret);
  gfree(g->chunk_part);
  gfree(g->chunk_off);
  gfree(g->chunk_log);
  const int64_t nt = cap / kMlpRows + 1;
  const int64_t nc = nt / kScanChunk + 1;
  if (hipMalloc(&g->retire, cap) != hipSuccess || hipMalloc(&g->tile_cnt, sizeof(int64_t) * nt) != hipSuccess ||
      hipMalloc(&g->tile_off, sizeof(int64_t) * nt) != hipSuccess ||
      hipMalloc(&g->tile_logcnt, sizeof(int64_t) * nt) != hipSuccess ||
      hipMalloc(&g->tile_logoff, sizeof(int64_t) * nt) != hipSuccess ||
      hipMalloc(&g->tile_ret, sizeof(double) * kRetStride * nt) != hipSuccess ||
      hipMalloc(&g->chunk_part, sizeof(ChunkPart) * nc) != hipSuccess ||
      hipMalloc(&g->chunk_off, sizeof(int64_t) * nc) != hipSuccess ||
      hipMalloc(&g->chunk_log, sizeof(int64_t) * nc) != hipSuccess) {
    (void)hipGetLastError();
    g->reg_cap = 0;
    return set_error(CUHRE_ERR_ALLOC, "cannot allocate guide buffers");
  }
  g->reg_cap = cap;
  return 0;
}

int guide_reserve_dataset(cuhre_guide* g, int64_t need)
{
  if (need <= g->ds_cap) return 0;
  const int64_t cap = std::max(need, 2 * g->ds_cap);
  float *feat, *fin;
  int8_t* ax;
  double *er, *vr;
  if (hipMalloc(&feat, sizeof(float) * 32 * cap) != hipSuccess || hipMalloc(&fin, sizeof(float) * cap) != hipSuccess ||
      hipMalloc(&ax, cap) != hipSuccess || hipMalloc(&er, sizeof(double) * cap) != hipSuccess ||
      hipMalloc(&vr, sizeof(double) * cap) != hipSuccess) {
    (void)hipGetLastError();
    return set_error(CUHRE_ERR_ALLOC, "cannot allocate the guide's sample store");
  }
  if (g->ds_n > 0) {
    const int64_t n = g->ds_n;
    if (hipMemcpy(feat, g->ds_feat, sizeof(float) * 32 * n, hipMemcpyDeviceToDevice) != hipSuccess ||
        hipMemcpy(fin, g->ds_fin, sizeof(float) * n, hipMemcpyDeviceToDevice) != hipSuccess ||
        hipMemcpy(ax, g->ds_axis, n, hipMemcpyDeviceToDevice) != hipSuccess ||
        hipMemcpy(er, g->ds_err, sizeof(double) * n, hipMemcpyDeviceToDevice) != hipSuccess ||
        hipMemcpy(vr, g->ds_vrel, sizeof(double) * n, hipMemcpyDeviceToDevice) != hipSuccess)
      return set_error(CUHRE_ERR_DEVICE, "sample store copy failed");
  }
  gfree(g->ds_feat);
  gfree(g->ds_fin);
  gfree(g->ds_axis);
  gfree(g->ds_err);
  gfree(g->ds_vrel);
  g->ds_feat = feat;
  g->ds_fin = fin;
  g->ds_axis = ax;
  g->ds_err = er;
  g->ds_vrel = vr;
  g->ds_cap = cap;
  return 0;
}

int guide_reserve_lookahead(cuhre_guide* g, int64_t need, int dim)
{
  if (need <= g->la_cap) return 0;
  const int64_t cap = (std::max(need, 2 * g->la_cap) + 255) & ~255ll;
  for (void* p : {(void*)g->la_lo, (void*)g->la_wd, (void*)g->la_est, (void*)g->la_err, (void*)g->la_ab,
                  (void*)g->la_axis})
    gfree(p);
  if (hipMalloc(&g->la_lo, sizeof(double) * cap * kMaxDim) != hipSuccess ||
      hipMalloc(&g->la_wd, sizeof(double) * cap * kMaxDim) != hipSuccess ||
      hipMalloc(&g->la_est, sizeof(double) * cap) != hipSuccess ||
      hipMalloc(&g->la_err, sizeof(double) * cap) != hipSuccess ||
      hipMalloc(&g->la_ab, sizeof(double) * cap) != hipSuccess || hipMalloc(&g->la_axis, cap) != hipSuccess) {
    (void)hipGetLastError();
    g->la_cap = 0;
    return set_error(CUHRE_ERR_ALLOC, "cannot allocate look-ahead buffers");
  }
  (void)dim;
  g->la_cap = cap;
  return 0;
}

MlpDev mlp_dev(const cuhre_guide* g)
{
  MlpDev M;
  M.W = g->W;
  M.WT = g->WT;
  M.b = g->params + kMlpBOff[0];
  return M;
}

// He-initialised weights from SplitMix64 + Box-Muller, zero biases.
void init_params(std::vector<float>& p, uint64_t seed)
{
  p.assign(kMlpParams, 0.f);
  uint64_t st = seed;
  auto u01 = [&]() {
    st += 0x9E3779B97F4A7C15ull;
    uint64_t z = st;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return (double)(z >> 11) * (1.0 / 9007199254740992.0);
  };
  for (int l = 0; l < 4; ++l) {
    const double sd = std::sqrt(2.0 / kMlpK[l]);
    for (int q = 0; q < kMlpK[l] * kMlpN[l]; ++q) {
      const double a = 1.0 - u01(), b = u01();
      p[kMlpWOff[l] + q] = (float)(sd * std::sqrt(-2.0 * std::log(a)) * std::cos(2.0 * M_PI * b));
    }
  }
}

int upload_params(cuhre_guide* g, const float* host)
{
  hipStream_t st = g->ctx->stream;
  if (hipMemcpyAsync(g->params, host, sizeof(float) * kMlpParams, hipMemcpyHostToDevice, st) != hipSuccess)
    return set_error(CUHRE_ERR_DEVICE, "parameter upload failed");
  (void)hipMemsetAsync(g->m, 0, sizeof(float) * kMlpParams, st);
  (void)hipMemsetAsync(g->v, 0, sizeof(float) * kMlpParams, st);
  g->adam_step = 0;
  hipLaunchKernelGGL(k_mlp_prepare, dim3((kMlpParams + 255) / 256), dim3(256), 0, st, g->params, g->W, g->WT,
                     g->img);
  if (hipGetLastError() != hipSuccess || hipStreamSynchronize(st) != hipSuccess)
    return set_error(CUHRE_ERR_DEVICE, "parameter preparation failed");
  return 0;
}

// Scratch for steps of `rows` rows, an explicit order of `order_len`
// entries and the statistics of `epochs` epochs.
int reserve_train(cuhre_guide* g, int64_t rows, int64_t order_len, int epochs)
{
  const int blocks = (int)((rows + kMlpRows - 1) / kMlpRows);
  if (blocks > g->part_blocks) {
    gfree(g->part);
    if (hipMalloc(&g->part, sizeof(float) * kPartStride * (size_t)blocks) != hipSuccess) {
      g->part_blocks = 0;
      return set_error(CUHRE_ERR_ALLOC, "cannot allocate gradient partials");
    }
    g->part_blocks = blocks;
  }
  if (order_len > g->train_cap) {
    gfree(g->order);
    if (hipMalloc(&g->order, sizeof(int32_t) * order_len) != hipSuccess) {
      g->train_cap = 0;
      return set_error(CUHRE_ERR_ALLOC, "cannot allocate the row order");
    }
    g->train_cap = order_len;
  }
  if (epochs > g->stats_cap) {
    gfree(g->d_stats);
    if (hipMalloc(&g->d_stats, sizeof(double) * 4 * epochs) != hipSuccess) {
      g->stats_cap = 0;
      return set_error(CUHRE_ERR_ALLOC, "cannot allocate training statistics");
    }
    g->stats_cap = epochs;
  }
  return 0;
}

// One optimisation step (or gradient only) over `rows` rows: row r is sample
// order[r], or with order null sample shuffle_index(shuf, first + r).  With
// `stats` set, the step's loss and hit counts are added into stats[0..2].
void train_step(cuhre_guide* g, const float* feat, const int8_t* ax, const float* fin, const int32_t* order,
                const Shuffle& shuf, int64_t first, int64_t rows, int dim, float lr, bool update, hipStream_t st,
                uint64_t perm_seed, double* stats)
{
  TrainArgs T;
  T.img = reinterpret_cast<const uint4*>(g->img);
  T.feat = feat;
  T.axis = ax;
  T.finish = fin;
  T.order = order;
  T.shuf = shuf;
  T.first = first;
  T.rows = rows;
  T.dim = dim;
  T.inv_batch = 1.0f / (float)rows;
  T.part = g->part;
  T.perm_seed = perm_seed;
  const int blocks = (int)((rows + kMlpRows - 1) / kMlpRows);
  hipLaunchKernelGGL(k_mlp_train, dim3(blocks), dim3(256), 0, st, T);
  if (update) ++g->adam_step;
  AdamArgs A;
  A.part = g->part;
  A.nblocks = blocks;
  A.grad = g->grad;
  A.params = g->params;
  A.m = g->m;
  A.v = g->v;
  A.step = std::max(g->adam_step, 1);
  A.lr = lr;
  A.do_update = update ? 1 : 0;
  A.W = g->W;
  A.WT = g->WT;
  A.img = g->img;
  A.acc = stats;
  const unsigned rblocks = (unsigned)((kMlpParams + 63) / 64) + (stats ? 1u : 0u);
  hipLaunchKernelGGL(k_mlp_reduce_adam, dim3(rblocks), dim3(kRedWaves * 64), 0, st, A);
}

struct GuidedRun {
  int axis_mode, finish_mode, use_mlp;
  float axis_conf, finish_thr;
  double budget;
  int64_t collect_max;  // collection: samples wanted from this run (0: none)
};

// Waits until the guide step's scan has published iteration g->seq (a
// stream query every ~2^18 polls guards against a launch that never does).
int wait_guided(cuhre_guide* g, hipStream_t st)
{
  for (uint64_t spin = 0;; ++spin) {
    if (__atomic_load_n(g->pub_seq, __ATOMIC_ACQUIRE) == g->seq) return 0;
    __builtin_ia32_pause();
    if ((spin & 262143) == 262143) {
      const hipError_t q = hipStreamQuery(st);
      if (q == hipSuccess && __atomic_load_n(g->pub_seq, __ATOMIC_ACQUIRE) != g->seq)
        return set_error(CUHRE_ERR_DEVICE, "guide step ended without publishing its state");
      if (q != hipSuccess && q != hipErrorNotReady) return set_error(CUHRE_ERR_DEVICE, hipGetErrorString(q));
    }
  }
}

// The plain loop of engine.hip plus the guide step between evaluation and
// split; with collect_max > 0 it also logs samples and their labels.
int guided_run(cuhre_context* c, cuhre_guide* g, const cuhre_integrand* f, int dim, double epsrel, double epsabs,
               int64_t maxeval, const double* lows, const double* highs, const GuidedRun& run, cuhre_result* out,
               cuhre_guide_stats* gs)
{
  const auto t0 = std::chrono::steady_clock::now();
  std::memset(out, 0, sizeof *out);
  if (gs) std::memset(gs, 0, sizeof *gs);
  int rc = check_common(c, f, dim, 0);
  if (rc) return rc;
  if (dim > 8) return set_error(CUHRE_ERR_UNSUPPORTED, "the guide handles dimensions up to 8");
  if (!lows || !highs) return set_error(CUHRE_ERR_ARG, "null bounds");
  DevIntegrand P;
  if ((rc = make_dev_integrand(c, f, dim, &P))) return rc;
  const RuleConst R = make_rule(dim);
  const int64_t p = R.npts;
  const int64_t n0 = initial_grid_count(dim);
  const int64_t capacity = list_capacity(c, dim);
  // the network and the sample log read I5 and the fourth differences
  const bool need_aux = run.use_mlp || run.collect_max > 0;
  // the error-budget finish replaces the per-region test
  const double loc_rel = run.finish_mode >= 3 ? 0.0 : epsrel, loc_abs = run.finish_mode >= 3 ? 0.0 : epsabs;
  if ((rc = ctx_reserve(c, std::min<int64_t>(capacity, std::max<int64_t>(4 * n0, 1 << 16)), dim))) return rc;
  if (need_aux && (rc = ctx_reserve_aux(c, c->cap, dim))) return rc;
  if ((rc = guide_reserve_regions(g, c->cap))) return rc;
  hipStream_t st = c->stream;
  int cur = 0;
  if (upload_initial_grid(c, cur, dim, lows, highs, n0) != 0)
    return set_error(CUHRE_ERR_DEVICE, "initial grid upload failed");
  const int64_t ds_base0 = g->ds_n;
  double fin_s = 0, fin_c = 0, fer_s = 0, fer_c = 0, fin_abs = 0, fin_aest = 0, ret_err = 0;
  int64_t nfin = 0, n = n0, logged = 0;
  int64_t bound = n0;  // 2 n of the previous list: picks the segments, as in the plain loop
  int result = CUHRE_NOT_CONVERGED;
  for (int it = 0; it < 1000; ++it) {
    const int S = pick_segments(c, bound, 7, f->family, dim);
    EvalArgs A;
    fill_eval_args(c, A, cur, n, dim, epsrel, epsabs);
    A.loc_epsrel = loc_rel;
    A.loc_epsabs = loc_abs;
    A.finalize = 1;
    A.i5 = need_aux ? c->aux : nullptr;
    A.dvec = need_aux ? c->aux + c->cap : nullptr;
    A.fin_est_s = fin_s;
    A.fin_est_c = fin_c;
    A.fin_err_s = fer_s;
    A.fin_err_c = fer_c;
    const int64_t left = maxeval - (out->neval + n * p);
    A.max_active = std::min<int64_t>(left >= 0 ? left / (2 * p) : -1, capacity / 2);
    A.alloc_active = c->cap / 2;
    find_eval_launcher(f->family, dim)(A, P, R, S, st);
    launch_finish(A, blocks_for(n, S), st);
    if (hipGetLastError() != hipSuccess) return set_error(CUHRE_ERR_DEVICE, "guided evaluation failed");

    // The guide step (network, finish rule, scan) and the split.  With
    // `dev` they read this iteration's estimate from the device state and
    // skip themselves when the evaluation stopped the run, so they follow
    // the evaluation without a host round trip; otherwise they take the
    // host's values (after a list growth, and in sampling runs, whose
    // buffers are sized on the host).
    const int64_t ntiles = (n + kMlpRows - 1) / kMlpRows;
    GuideArgs G;
    RoomArgs RA;
    auto guide_step = [&](bool dev) {
      std::memset(&G, 0, sizeof G);
      G.lower = c->lo[cur];
      G.width = c->wd[cur];
      G.ld = c->cap;
      G.n = n;
      G.est = c->est;
      G.err = c->err;
      G.i5 = c->aux;
      G.dvec = c->aux + c->cap;
      G.abssum = c->abssum;
      G.axis = c->axis;
      G.retire = g->retire;
      G.dim = dim;
      for (int m = 0; m < dim; ++m) {
        G.lows[m] = lows[m];
        G.highs[m] = highs[m];
      }
      G.absI = std::fabs(out->integral);
      G.fin_aest = fin_aest;
      G.absE = fin_aest + c->h_state->tot.aest_all;  // read only when !dev (synchronised)
      G.epsrel = epsrel;
      G.epsabs = epsabs;
      G.tol = std::max(epsrel * G.absI, epsabs);
      G.state = dev ? c->state : nullptr;
      G.loc_epsrel = loc_rel;
      G.loc_epsabs = loc_abs;
      G.axis_mode = run.axis_mode;
      G.finish_mode = run.finish_mode;
      G.use_mlp = run.use_mlp;
      G.axis_conf = run.axis_conf;
      G.finish_thr = run.finish_thr;
      G.budget = run.budget;
      G.M = mlp_dev(g);
      G.tile_cnt = g->tile_cnt;
      G.tile_logcnt = g->tile_logcnt;
      G.tile_ret = g->tile_ret;
      const int64_t want = run.collect_max - logged;
      const int64_t nact = c->h_state->tot.nactive;  // read only when !dev (synchronised)
      if (!dev && run.collect_max > 0 && want > 0 && nact > 0) {
        // about collect_max / 24 regions per iteration, so that the deep
        // iterations of a long run are represented as well as the early ones
        const int64_t target =
            std::max<int64_t>(1, std::min<int64_t>(want, std::max<int64_t>(run.collect_max / 24, 4096)));
        G.log_stride = (uint64_t)std::max<int64_t>(1, (nact + target - 1) / target);
        G.log_salt = 0x5DEECE66Dull * (uint64_t)(it + 1);
      }
      if (run.use_mlp)
        hipLaunchKernelGGL(k_guide_apply<true>, dim3((unsigned)ntiles), dim3(64), 0, st, G);
      else
        hipLaunchKernelGGL(k_guide_apply<false>, dim3((unsigned)((ntiles + 3) / 4)), dim3(256), 0, st, G);
      // retirement room: retired error within budget * tol, frozen error
      // (finished so far, finished now, retired) within (1 - budget) * tol
      RA.state = G.state;
      RA.frozen_prev = fer_s + fer_c;
      RA.ret_prev = ret_err;
      RA.budget = run.budget;
      RA.epsrel = epsrel;
      RA.epsabs = epsabs;
      RA.mode = run.finish_mode;
      RA.room = -1.0;
      if (!dev) {
        const IterState& hs = *c->h_state;
        const double frozen = RA.frozen_prev + (hs.tot.ferr_s + hs.tot.ferr_c);
        RA.room = run.finish_mode >= 3 ? 1e300
                  : run.finish_mode    ? std::min(run.budget * G.tol - ret_err, (1.0 - run.budget) * G.tol - frozen)
                                       : -1.0;
      }
      const int64_t nchunks = (ntiles + kScanChunk - 1) / kScanChunk;
      hipLaunchKernelGGL(k_gscan_part, dim3((unsigned)nchunks), dim3(256), 0, st, g->tile_cnt, g->tile_logcnt,
                         g->tile_ret, ntiles, RA.state, g->chunk_part);
      const bool publish = dev && run.collect_max == 0;
      if (publish) ++g->seq;
      hipLaunchKernelGGL(k_gscan_top, dim3(1), dim3(256), 0, st, g->chunk_part, nchunks, RA, g->chunk_off,
                         g->chunk_log, g->d_sums, g->d_retire_on, publish ? g->pub_dev : nullptr,
                         publish ? g->pub_seq_dev : nullptr, g->seq);
      hipLaunchKernelGGL(k_gscan_tiles, dim3((unsigned)nchunks), dim3(256), 0, st, g->tile_cnt, g->tile_logcnt,
                         g->tile_ret, ntiles, RA.state, g->chunk_off, g->chunk_log, g->d_retire_on, g->tile_off,
                         g->tile_logoff);
      SplitArgs SA;
      fill_split_args(c, SA, cur, n, dim, loc_rel, loc_abs, kMlpRows, dev);
      SA.offsets = g->tile_off;
      SA.retire = g->retire;
      SA.retire_on = g->d_retire_on;
      launch_split(SA, st);
    };
    const bool sampling = run.collect_max > 0;
    GuideSums gsum_pub = {};
    if (!sampling) {
      // the guide step publishes this iteration's state and sums; the host
      // queues the next iteration while the scan of tiles and the split run
      guide_step(true);
      if (hipGetLastError() != hipSuccess) return set_error(CUHRE_ERR_DEVICE, "guided iteration failed");
      if ((rc = wait_guided(g, st))) return rc;
      *c->h_state = g->pub->st;
      gsum_pub = g->pub->sums;
    } else if (hipMemcpyAsync(c->h_state, c->state, sizeof(IterState), hipMemcpyDeviceToHost, st) != hipSuccess ||
               hipStreamSynchronize(st) != hipSuccess) {
      return set_error(CUHRE_ERR_DEVICE, "guided iteration failed");
    }
    IterState s = *c->h_state;
    out->eval_launches++;
    out->iterations = it + 1;
    out->nregions += n;
    out->neval += n * p;
    out->peak_regions = std::max(out->peak_regions, n);
    out->integral = s.rG_s + s.rG_c;
    out->error = s.eG_s + s.eG_c;
    bool rerun = sampling;
    if (s.stop == 5) {
      // children would not fit: grow the lists (keeping this iteration) and
      // the guide's per-region buffers, then run the guide step again
      const int64_t ncap = std::min<int64_t>(capacity, std::max<int64_t>(2 * s.tot.nactive, 2 * c->cap));
      const int64_t nb = blocks_for(n, S);
      if ((rc = ctx_reserve(c, ncap, dim, n, cur, nb))) return rc;
      if ((rc = guide_reserve_regions(g, c->cap))) return rc;
      if (need_aux) {
        // i5 / dvec were computed into the old aux block: re-evaluate this list
        if ((rc = ctx_reserve_aux(c, c->cap, dim))) return rc;
        fill_eval_args(c, A, cur, n, dim, epsrel, epsabs);
        A.loc_epsrel = loc_rel;
        A.loc_epsabs = loc_abs;
        A.finalize = 0;
        A.i5 = c->aux;
        A.dvec = c->aux + c->cap;
        find_eval_launcher(f->family, dim)(A, P, R, S, st);
      }
      s.stop = 0;
      rerun = true;
    }
    const double abs_contrib = fin_abs + s.tot.abs_all;
    const double aest_contrib = fin_aest + s.tot.aest_all;
    out->fp64_err_bound = fp64_rounding_bound(p, dim, nfin + n, abs_contrib, aest_contrib, out->integral);
    if (s.stop == 4) {
      result = set_error(CUHRE_ERR_NONFINITE, "integrand returned a non-finite value");
      break;
    }
    if (s.stop == 1) {
      result = CUHRE_OK;
      break;
    }
    if (s.stop == 2 || s.stop == 3) {
      result = CUHRE_NOT_CONVERGED;
      break;
    }
    if (rerun) {
      guide_step(false);
      if (hipMemcpyAsync(g->h_sums, g->d_sums, sizeof(GuideSums), hipMemcpyDeviceToHost, st) != hipSuccess ||
          hipStreamSynchronize(st) != hipSuccess)
        return set_error(CUHRE_ERR_DEVICE, "guide step failed");
    }
    const GuideSums gsum = rerun ? *g->h_sums : gsum_pub;
    if (G.log_stride > 0 && gsum.n_logged > 0) {
      const int64_t L = std::min<int64_t>(gsum.n_logged, run.collect_max - logged);
      if ((rc = guide_reserve_dataset(g, g->ds_n + gsum.n_logged))) return rc;
      if ((rc = guide_reserve_lookahead(g, gsum.n_logged * 2 * dim, dim))) return rc;
      CollectArgs CA;
      CA.G = G;
      CA.logoff = g->tile_logoff;
      CA.base = g->ds_n;
      CA.feat = g->ds_feat;
      CA.s_err = g->ds_err;
      CA.s_vrel = g->ds_vrel;
      CA.la_lo = g->la_lo;
      CA.la_wd = g->la_wd;
      CA.la_ld = g->la_cap;
      hipLaunchKernelGGL(k_collect_write, dim3((unsigned)ntiles), dim3(64), 0, st, CA);
      EvalArgs LA;
      std::memset(&LA, 0, sizeof LA);
      LA.lower = g->la_lo;
      LA.width = g->la_wd;
      LA.ld = g->la_cap;
      LA.n = gsum.n_logged * 2 * dim;
      LA.epsrel = LA.loc_epsrel = epsrel;
      LA.epsabs = LA.loc_epsabs = epsabs;
      LA.est = g->la_est;
      LA.err = g->la_err;
      LA.axis = g->la_axis;
      LA.abssum = g->la_ab;
      LA.finalize = 0;
      find_eval_launcher(f->family, dim)(LA, P, R, pick_segments(c, LA.n, 7, f->family, dim), st);
      hipLaunchKernelGGL(k_label_axis, dim3((unsigned)((gsum.n_logged + 255) / 256)), dim3(256), 0, st, g->la_err,
                         gsum.n_logged, dim, g->ds_n, g->ds_axis);
      if (hipGetLastError() != hipSuccess) return set_error(CUHRE_ERR_DEVICE, "sample collection failed");
      g->ds_n += L;  // keep at most collect_max (the extra rows are overwritten later)
      logged += L;
    }
    if (gs) {
      gs->retired += gsum.n_retired;
      gs->guide_iterations += 1;
      gs->overridden += gsum.n_overridden;
    }
    out->split_launches++;
    dd_add_host(fin_s, fin_c, s.tot.fest_s, s.tot.fest_c);
    dd_add_host(fer_s, fer_c, s.tot.ferr_s, s.tot.ferr_c);
    dd_add_host(fin_s, fin_c, gsum.est_s, gsum.est_c);
    dd_add_host(fer_s, fer_c, gsum.err_s, gsum.err_c);
    ret_err += gsum.err_s + gsum.err_c;
    fin_abs += s.tot.abs_fin + gsum.abs_r;
    fin_aest += s.tot.aest_fin + gsum.aest_r;
    nfin += n - gsum.n_split;
    out->nfinished = nfin;
    if (gsum.n_split == 0) {
      // every region is finished or retired: the estimate is final
      out->integral = fin_s + fin_c;
      out->error = fer_s + fer_c;
      result = out->error <= std::max(epsrel * std::fabs(out->integral), epsabs) ? CUHRE_OK : CUHRE_NOT_CONVERGED;
      break;
    }
    bound = 2 * n;
    n = 2 * gsum.n_split;
    cur = 1 - cur;
  }
  if (result >= 0 && g->ds_n > ds_base0) {
    // finish labels of this run's samples against its final estimate
    const double tol = epsrel * std::fabs(out->integral);
    hipLaunchKernelGGL(k_label_finish, dim3((unsigned)((g->ds_n - ds_base0 + 255) / 256)), dim3(256), 0, st,
                       g->ds_err, g->ds_vrel, ds_base0, g->ds_n, tol, g->ds_fin);
    if (hipGetLastError() != hipSuccess || hipStreamSynchronize(st) != hipSuccess)
      return set_error(CUHRE_ERR_DEVICE, "finish labelling failed");
  }
  if (gs) gs->samples = g->ds_n - ds_base0;
  // the last iteration's tile scan and split (they skip themselves once the
  // run stopped) may still be queued
  if (hipStreamSynchronize(st) != hipSuccess) return set_error(CUHRE_ERR_DEVICE, "guided run failed");
  out->fail = result > 0 ? result : 0;
  out->ms_total = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  return result;
}

}  // namespace

namespace cuhre {
void guide_detach(cuhre_guide* g) { g->ctx = nullptr; }
}  // namespace cuhre

namespace {

// A guide is usable while its context lives.
int guide_check(const cuhre_guide* g)
{
  if (!g) return set_error(CUHRE_ERR_ARG, "null guide");
  if (!g->ctx) return set_error(CUHRE_ERR_ARG, "the guide's context was destroyed");
  return 0;
}

}  // namespace

extern "C" {

int cuhre_guide_create(cuhre_context* ctx, uint64_t seed, cuhre_guide** out)
{
  if (!ctx || !out) return set_error(CUHRE_ERR_ARG, "null argument");
  *out = nullptr;
  if (hipSetDevice(ctx->device) != hipSuccess) return set_error(CUHRE_ERR_DEVICE, "hipSetDevice failed");
  cuhre_guide* g = new cuhre_guide();
  g->ctx = ctx;
  if (hipMalloc(&g->params, sizeof(float) * kMlpParams) != hipSuccess ||
      hipMalloc(&g->m, sizeof(float) * kMlpParams) != hipSuccess ||
      hipMalloc(&g->v, sizeof(float) * kMlpParams) != hipSuccess ||
      hipMalloc(&g->grad, sizeof(float) * kMlpParams) != hipSuccess ||
      hipMalloc(&g->W, sizeof(uint16_t) * kMlpBOff[0]) != hipSuccess ||
      hipMalloc(&g->WT, sizeof(uint16_t) * kMlpBOff[0]) != hipSuccess ||
      hipMalloc(&g->img, kImgBytes) != hipSuccess || hipMemset(g->img, 0, kImgBytes) != hipSuccess ||
      hipMalloc(&g->d_sums, sizeof(GuideSums)) != hipSuccess ||
      hipMalloc(&g->d_retire_on, sizeof(int)) != hipSuccess ||
      hipHostMalloc(&g->h_sums, sizeof(GuideSums), hipHostMallocDefault) != hipSuccess ||
      hipHostMalloc(&g->pub, sizeof(GuidePub), hipHostMallocCoherent | hipHostMallocMapped) != hipSuccess ||
      hipHostMalloc(&g->pub_seq, sizeof(uint32_t), hipHostMallocCoherent | hipHostMallocMapped) != hipSuccess ||
      hipHostGetDevicePointer((void**)&g->pub_dev, g->pub, 0) != hipSuccess ||
      hipHostGetDevicePointer((void**)&g->pub_seq_dev, g->pub_seq, 0) != hipSuccess) {
    (void)hipGetLastError();
    cuhre_guide_destroy(g);
    return set_error(CUHRE_ERR_ALLOC, "cannot allocate the guide");
  }
  *g->pub_seq = 0;
  std::vector<float> p;
  init_params(p, seed);
  const int rc = upload_params(g, p.data());
  if (rc) {
    cuhre_guide_destroy(g);
    return rc;
  }
  ctx->guides.push_back(g);
  *out = g;
  return 0;
}

int cuhre_guide_destroy(cuhre_guide* g)
{
  if (!g) return 0;
  if (g->ctx) {
    (void)hipStreamSynchronize(g->ctx->stream);
    auto& v = g->ctx->guides;
    v.erase(std::remove(v.begin(), v.end(), g), v.end());
  }
  for (void* q : {(void*)g->params, (void*)g->m, (void*)g->v, (void*)g->W, (void*)g->WT, (void*)g->img, (void*)g->grad,
                  (void*)g->part, (void*)g->ds_feat, (void*)g->ds_axis, (void*)g->ds_fin, (void*)g->ds_err,
                  (void*)g->ds_vrel, (void*)g->la_lo, (void*)g->la_wd, (void*)g->la_est, (void*)g->la_err,
                  (void*)g->la_ab, (void*)g->la_axis, (void*)g->retire, (void*)g->tile_cnt, (void*)g->tile_off,
                  (void*)g->tile_logcnt, (void*)g->tile_logoff, (void*)g->tile_ret, (void*)g->d_sums, (void*)g->d_retire_on,
                  (void*)g->chunk_part, (void*)g->chunk_off, (void*)g->chunk_log,
                  (void*)g->order, (void*)g->d_stats})
    gfree(q);
  if (g->h_sums) (void)hipHostFree(g->h_sums);
  if (g->pub) (void)hipHostFree(g->pub);
  if (g->pub_seq) (void)hipHostFree(g->pub_seq);
  delete g;
  return 0;
}

int cuhre_guide_set_params(cuhre_guide* g, const float* params)
{
  if (guide_check(g)) return CUHRE_ERR_ARG;
  if (!params) return set_error(CUHRE_ERR_ARG, "null argument");
  return upload_params(g, params);
}

int cuhre_guide_get_params(cuhre_guide* g, float* params)
{
  if (guide_check(g)) return CUHRE_ERR_ARG;
  if (!params) return set_error(CUHRE_ERR_ARG, "null argument");
  if (hipMemcpy(params, g->params, sizeof(float) * kMlpParams, hipMemcpyDeviceToHost) != hipSuccess)
    return set_error(CUHRE_ERR_DEVICE, "parameter download failed");
  return 0;
}

int cuhre_guide_forward(cuhre_guide* g, int64_t n, const float* features, float* logits)
{
  if (guide_check(g)) return CUHRE_ERR_ARG;
  if (!features || !logits) return set_error(CUHRE_ERR_ARG, "null argument");
  if (n <= 0) return 0;
  hipStream_t st = g->ctx->stream;
  float *dx, *dy;
  if (hipMalloc(&dx, sizeof(float) * 32 * n) != hipSuccess || hipMalloc(&dy, sizeof(float) * 16 * n) != hipSuccess)
    return set_error(CUHRE_ERR_ALLOC, "cannot allocate forward buffers");
  int rc = 0;
  if (hipMemcpyAsync(dx, features, sizeof(float) * 32 * n, hipMemcpyHostToDevice, st) != hipSuccess) rc = 1;
  hipLaunchKernelGGL(k_mlp_forward, dim3((unsigned)((n + kMlpRows - 1) / kMlpRows)), dim3(kWave), 0, st, mlp_dev(g),
                     dx, n, dy);
  if (hipGetLastError() != hipSuccess ||
      hipMemcpyAsync(logits, dy, sizeof(float) * 16 * n, hipMemcpyDeviceToHost, st) != hipSuccess ||
      hipStreamSynchronize(st) != hipSuccess)
    rc = 1;
  (void)hipFree(dx);
  (void)hipFree(dy);
  return rc ? set_error(CUHRE_ERR_DEVICE, "guide forward failed") : 0;
}

int cuhre_guide_gradient(cuhre_guide* g, int64_t n, const float* features, const int32_t* axis,
                         const float* finish, int dim, float* grad, double* loss)
{
  if (guide_check(g)) return CUHRE_ERR_ARG;
  if (!features || !axis || !finish || !grad) return set_error(CUHRE_ERR_ARG, "null argument");
  if (n <= 0 || n > INT32_MAX || dim < 2 || dim > 8) return set_error(CUHRE_ERR_ARG, "bad size or dimension");
  hipStream_t st = g->ctx->stream;
  int rc = reserve_train(g, n, n, 1);
  if (rc) return rc;
  float *dx, *df;
  int8_t* da;
  std::vector<int8_t> ax8(n);
  std::vector<int32_t> ord(n);
  for (int64_t i = 0; i < n; ++i) {
    if (axis[i] < 0 || axis[i] >= dim) return set_error(CUHRE_ERR_ARG, "axis label out of range");
    ax8[i] = (int8_t)axis[i];
    ord[i] = (int32_t)i;
  }
  if (hipMalloc(&dx, sizeof(float) * 32 * n) != hipSuccess || hipMalloc(&df, sizeof(float) * n) != hipSuccess ||
      hipMalloc(&da, n) != hipSuccess)
    return set_error(CUHRE_ERR_ALLOC, "cannot allocate gradient inputs");
  double acc[4] = {0, 0, 0, 0};
  bool ok = hipMemcpyAsync(dx, features, sizeof(float) * 32 * n, hipMemcpyHostToDevice, st) == hipSuccess &&
            hipMemcpyAsync(df, finish, sizeof(float) * n, hipMemcpyHostToDevice, st) == hipSuccess &&
            hipMemcpyAsync(da, ax8.data(), n, hipMemcpyHostToDevice, st) == hipSuccess &&
            hipMemcpyAsync(g->order, ord.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice, st) == hipSuccess &&
            hipMemsetAsync(g->d_stats, 0, sizeof(double) * 4, st) == hipSuccess;
  if (ok) {
    train_step(g, dx, da, df, g->order, Shuffle{}, 0, n, dim, 0.f, false, st, 0, g->d_stats);
    ok = hipGetLastError() == hipSuccess &&
         hipMemcpyAsync(grad, g->grad, sizeof(float) * kMlpParams, hipMemcpyDeviceToHost, st) == hipSuccess &&
         hipMemcpyAsync(acc, g->d_stats, sizeof acc, hipMemcpyDeviceToHost, st) == hipSuccess &&
         hipStreamSynchronize(st) == hipSuccess;
    if (ok && loss) *loss = acc[0] / (double)n;
  }
  (void)hipFree(dx);
  (void)hipFree(df);
  (void)hipFree(da);
  return ok ? 0 : set_error(CUHRE_ERR_DEVICE, "guide gradient failed");
}

int64_t cuhre_guide_samples(const cuhre_guide* g) { return g ? g->ds_n : -1; }

int cuhre_guide_get_samples(cuhre_guide* g, int64_t first, int64_t count, float* features, int32_t* axis,
                            float* finish)
{
  if (guide_check(g)) return CUHRE_ERR_ARG;
  if (first < 0 || count < 0 || first + count > g->ds_n) return set_error(CUHRE_ERR_ARG, "bad sample range");
  if (count == 0) return 0;
  std::vector<int8_t> a8(count);
  if ((features && hipMemcpy(features, g->ds_feat + first * 32, sizeof(float) * 32 * count,
                             hipMemcpyDeviceToHost) != hipSuccess) ||
      hipMemcpy(a8.data(), g->ds_axis + first, count, hipMemcpyDeviceToHost) != hipSuccess ||
      (finish && hipMemcpy(finish, g->ds_fin + first, sizeof(float) * count, hipMemcpyDeviceToHost) != hipSuccess))
    return set_error(CUHRE_ERR_DEVICE, "sample download failed");
  if (axis)
    for (int64_t i = 0; i < count; ++i) axis[i] = a8[i];
  return 0;
}

int cuhre_guide_collect(cuhre_guide* g, const cuhre_integrand* f, int dim, double epsrel, double epsabs,
                        int64_t maxeval, const double* lows, const double* highs, int64_t max_samples,
                        const cuhre_guide_policy* pol, cuhre_result* out)
{
  if (guide_check(g)) return CUHRE_ERR_ARG;
  if (!out) return set_error(CUHRE_ERR_ARG, "null argument");
  // pol NULL: the plain method decides every axis; with axis_mode 1 the
  // network decides where confident, so the log holds the regions its own
  // choices produce (labelled by the look-ahead as always)
  if (pol && (pol->axis_mode < 0 || pol->axis_mode > 1)) return set_error(CUHRE_ERR_ARG, "bad guide policy");
  const int am = pol ? pol->axis_mode : 0;
  GuidedRun run{am, 0, am, pol ? pol->axis_confidence : 1.f, 1.f, 0.0, max_samples};
  cuhre_guide_stats gs;
  return guided_run(g->ctx, g, f, dim, epsrel, epsabs, maxeval, lows, highs, run, out, &gs);
}

int cuhre_guide_train(cuhre_guide* g, int dim, int epochs, int batch, float lr, uint64_t seed,
                      cuhre_train_stats* stats)
{
  if (guide_check(g)) return CUHRE_ERR_ARG;
  if (!stats) return set_error(CUHRE_ERR_ARG, "null argument");
  std::memset(stats, 0, sizeof *stats);
  if (dim < 2 || dim > 8 || epochs <= 0 || batch < kMlpRows) return set_error(CUHRE_ERR_ARG, "bad training settings");
  const int64_t n = g->ds_n;
  if (n < batch) return set_error(CUHRE_ERR_ARG, "fewer samples than one batch");
  hipStream_t st = g->ctx->stream;
  int rc = reserve_train(g, batch, 0, epochs);
  if (rc) return rc;
  // Every epoch visits the samples in the order of its own permutation,
  // computed per row on the device; the host queues the whole run's steps
  // and reads the per-epoch statistics once at the end.
  const int64_t steps = n / batch;
  const auto t0 = std::chrono::steady_clock::now();
  if (hipMemsetAsync(g->d_stats, 0, sizeof(double) * 4 * epochs, st) != hipSuccess)
    return set_error(CUHRE_ERR_DEVICE, "statistics reset failed");
  for (int ep = 0; ep < epochs; ++ep) {
    const Shuffle P = make_shuffle((uint64_t)n, mix64(seed ^ (0x2545F4914F6CDD1Dull * (uint64_t)(ep + 1))));
    for (int64_t k = 0; k < steps; ++k)
      train_step(g, g->ds_feat, g->ds_axis, g->ds_fin, nullptr, P, k * batch, batch, dim, lr, true, st,
                 mix64(seed ^ (0xA0761D6478BD642Full * (uint64_t)(ep * steps + k + 1))), g->d_stats + 4 * ep);
  }
  std::vector<double> acc((size_t)4 * epochs);
  if (hipGetLastError() != hipSuccess ||
      hipMemcpyAsync(acc.data(), g->d_stats, sizeof(double) * 4 * epochs, hipMemcpyDeviceToHost, st) != hipSuccess ||
      hipStreamSynchronize(st) != hipSuccess)
    return set_error(CUHRE_ERR_DEVICE, "training step failed");
  const double rows = (double)(steps * batch);
  stats->loss_first = acc[0] / rows;
  stats->loss_last = acc[4 * (epochs - 1)] / rows;
  stats->axis_accuracy = acc[4 * (epochs - 1) + 1] / rows;
  stats->finish_accuracy = acc[4 * (epochs - 1) + 2] / rows;
  stats->epochs = epochs;
  stats->steps = (int32_t)(steps * epochs);
  stats->samples = n;
  stats->ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  return 0;
}

int cuhre_integrate_guided(cuhre_context* ctx, cuhre_guide* g, const cuhre_integrand* f, int dim, double epsrel,
                           double epsabs, int64_t maxeval, const double* lows, const double* highs,
                           const cuhre_guide_policy* pol, cuhre_result* out, cuhre_guide_stats* stats)
{
  if (!ctx || !g || !out) return set_error(CUHRE_ERR_ARG, "null argument");
  if (g->ctx != ctx) return set_error(CUHRE_ERR_ARG, "the guide belongs to another context");
  cuhre_guide_policy d = {1, 1, 0.6f, 0.5f, 0.25};
  const cuhre_guide_policy& P = pol ? *pol : d;
  if (P.axis_mode < 0 || P.axis_mode > 1 || P.finish_mode < 0 || P.finish_mode > 4 || !(P.budget >= 0.0) ||
      P.budget > 1.0)
    return set_error(CUHRE_ERR_ARG, "bad guide policy");
  GuidedRun run{P.axis_mode, P.finish_mode, (P.axis_mode == 1 || P.finish_mode == 1) ? 1 : 0, P.axis_confidence,
                P.finish_threshold, P.budget, 0};
  return guided_run(ctx, g, f, dim, epsrel, epsabs, maxeval, lows, highs, run, out, stats);
}

}  // extern "C"
