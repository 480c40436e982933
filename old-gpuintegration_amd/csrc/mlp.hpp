// mlp.hpp -- the learned split guide's network on the MFMA units.
//
// 32 -> 64 -> 64 -> 64 -> 16, ReLU.  BF16 weights and BF16 activations enter
// v_mfma_f32_16x16x16_bf16 (FP32 accumulation).  Two forms of the forward
// pass: inference (mlp_forward64_wave) runs 64 rows on one wavefront, layer
// by layer, in LDS of its own; training (mlp_forward16_wave) gives each
// wavefront of a 256-thread block 16 of the block's 64 rows, with the
// weights staged in LDS.  Output columns 0..d-1
// are split-axis logits, column 15 the finish logit.
//
// Operand layout of v_mfma_f32_16x16x16_bf16 (lane l, item v = 0..3):
//   A[i][k]: i = l % 16, k = 4*(l/16) + v
//   B[k][j]: j = l % 16, k = 4*(l/16) + v
//   D[i][j]: j = l % 16, i = 4*(l/16) + v
#pragma once

#include "common.hpp"

namespace cuhre {

constexpr int kMlpIn = 32, kMlpHid = 64, kMlpOut = 16, kMlpFinish = 15;
constexpr int kMlpRows = 64;  // rows per tile (a training block, an inference wavefront)
// parameter vector: W0 [32x64], W1 [64x64], W2 [64x64], W3 [64x16], b0, b1, b2, b3
constexpr int kMlpK[4] = {32, 64, 64, 64};
constexpr int kMlpN[4] = {64, 64, 64, 16};
constexpr int kMlpWOff[4] = {0, 2048, 6144, 10240};
constexpr int kMlpBOff[4] = {11264, 11328, 11392, 11456};
constexpr int kMlpParams = 11472;

// LDS rows (activations and gradients) carry kPad extra BF16 (8 bytes):
// the 16 rows one MFMA operand read touches then start in different banks
// instead of all in the same one (a row of 64 BF16 is exactly one pass over
// the 32 banks).
constexpr int kPad = 4;
constexpr int kLdIn = kMlpIn + kPad, kLdHid = kMlpHid + kPad, kLdOut = kMlpOut + kPad;

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ uint16_t f2bf(float f)
{
  const uint32_t u = __float_as_uint(f);
  return (uint16_t)((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

// Device view of the weights: W[l] row-major [K][N] and WT[l] = W^T [N][K], both BF16.
struct MlpDev {
  const uint16_t* W;   // concatenated at kMlpWOff
  const uint16_t* WT;  // concatenated at kMlpWOff
  const float* b;      // biases at kMlpBOff - kMlpBOff[0]
};

__device__ __forceinline__ f32x4 mfma16(bf16x4 a, bf16x4 b, f32x4 c)
{
  return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

// Y[16 x N] = X[16 x K] . W for this wave's 16 rows; X in LDS (row stride
// K + kPad, BF16), W read from W^T ([N][K], row stride LDW) so each B
// operand is 8 contiguous bytes.
template <int K, int N, int LDW>
__device__ __forceinline__ void mlp_layer_fwd(const uint16_t* Xs, const uint16_t* WT, f32x4 (&acc)[N / 16])
{
  const int l = __lane_id(), row = l & 15, kq = (l >> 4) * 4;
#pragma unroll
  for (int nt = 0; nt < N / 16; ++nt) acc[nt] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int kt = 0; kt < K / 16; ++kt) {
    const bf16x4 a = *reinterpret_cast<const bf16x4*>(Xs + row * (K + kPad) + kt * 16 + kq);
#pragma unroll
    for (int nt = 0; nt < N / 16; ++nt) {
      const bf16x4 b = *reinterpret_cast<const bf16x4*>(WT + (nt * 16 + row) * LDW + kt * 16 + kq);
      acc[nt] = mfma16(a, b, acc[nt]);
    }
  }
}

// Bias + ReLU + BF16 store of a layer's output tile into LDS [16][N + kPad].
template <int N>
__device__ __forceinline__ void mlp_store_relu(const f32x4 (&acc)[N / 16], const float* bias, uint16_t* Hs)
{
  const int l = __lane_id(), col = l & 15, r0 = (l >> 4) * 4;
#pragma unroll
  for (int nt = 0; nt < N / 16; ++nt)
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const float z = acc[nt][v] + bias[nt * 16 + col];
      Hs[(r0 + v) * (N + kPad) + nt * 16 + col] = f2bf(z > 0.f ? z : 0.f);
    }
}

// Training image: the weights as the training kernel holds them in LDS, so
// one linear copy stages them.  W^T of layer l as [N][K + kPad] BF16 at
// element kImgWT[l], W of layer l as [K][N + kPad] BF16 at kImgW[l] (the
// padding spreads the 16 rows of an operand read over the banks), then the
// 208 FP32 biases at byte kImgBiasByte.  Padding elements stay zero.
constexpr int kImgWT[4] = {0, 2304, 6656, 11008};
constexpr int kImgW[4] = {12096, 14272, 18624, 22976};
constexpr int kImgBf16 = 24256;
constexpr int kImgBiasByte = kImgBf16 * 2;
constexpr int kImgBytes = kImgBiasByte + (kMlpParams - kMlpBOff[0]) * 4;
static_assert(kImgWT[1] == kImgWT[0] + kMlpN[0] * (kMlpK[0] + kPad) &&
                  kImgWT[2] == kImgWT[1] + kMlpN[1] * (kMlpK[1] + kPad) &&
                  kImgWT[3] == kImgWT[2] + kMlpN[2] * (kMlpK[2] + kPad) &&
                  kImgW[0] == kImgWT[3] + kMlpN[3] * (kMlpK[3] + kPad) &&
                  kImgW[1] == kImgW[0] + kMlpK[0] * (kMlpN[0] + kPad) &&
                  kImgW[2] == kImgW[1] + kMlpK[1] * (kMlpN[1] + kPad) &&
                  kImgW[3] == kImgW[2] + kMlpK[2] * (kMlpN[2] + kPad) &&
                  kImgBf16 == kImgW[3] + kMlpK[3] * (kMlpN[3] + kPad),
              "training image layout");
static_assert(kImgBiasByte % 16 == 0 && kImgBytes % 16 == 0, "training image alignment");

__host__ __device__ constexpr int img_wt(int l, int k, int n) { return kImgWT[l] + n * (kMlpK[l] + kPad) + k; }
__host__ __device__ constexpr int img_w(int l, int k, int n) { return kImgW[l] + k * (kMlpN[l] + kPad) + n; }

// Forward pass of this wave's 16 rows (the training kernel's form, whose
// backward pass reuses the activations).  Xs..H3s are this wave's LDS rows
// (padded strides); Ys receives the 16 x 16 logits in FP32.  The weights
// and biases come from the training image in LDS.  Wave-local: the layers
// are ordered by wavefront fences.
__device__ __forceinline__ void mlp_forward16_wave(const uint16_t* img, const float* bias, const uint16_t* Xs,
                                                   uint16_t* H1s, uint16_t* H2s, uint16_t* H3s, float* Ys)
{
  {
    f32x4 acc[4];
    mlp_layer_fwd<32, 64, 32 + kPad>(Xs, img + kImgWT[0], acc);
    mlp_store_relu<64>(acc, bias + 0, H1s);
  }
  __syncwarp();
  {
    f32x4 acc[4];
    mlp_layer_fwd<64, 64, 64 + kPad>(H1s, img + kImgWT[1], acc);
    mlp_store_relu<64>(acc, bias + 64, H2s);
  }
  __syncwarp();
  {
    f32x4 acc[4];
    mlp_layer_fwd<64, 64, 64 + kPad>(H2s, img + kImgWT[2], acc);
    mlp_store_relu<64>(acc, bias + 128, H3s);
  }
  __syncwarp();
  {
    f32x4 acc[1];
    mlp_layer_fwd<64, 16, 64 + kPad>(H3s, img + kImgWT[3], acc);
    const int l = __lane_id(), col = l & 15, r0 = (l >> 4) * 4;
#pragma unroll
    for (int v = 0; v < 4; ++v) Ys[(r0 + v) * 16 + col] = acc[0][v] + bias[192 + col];
  }
  __syncwarp();
}

// One layer over a wave's 64 rows: the B operands (W^T rows) are loaded once
// and used for the four 16-row groups, each followed by bias + ReLU into Out
// [64][N + kPad] BF16 (LAST: the plain logits into Yout [64][16] FP32).  A
// group whose bit in `work` is clear (wave-uniform) is skipped.
template <int K, int N, int LDW, bool LAST>
__device__ __forceinline__ void mlp_layer64(const uint16_t* In, int ldin, const uint16_t* WT, const float* bias,
                                            uint16_t* Out, float* Yout, unsigned work)
{
  const int l = __lane_id(), row = l & 15, kq = (l >> 4) * 4, col = l & 15, r0 = (l >> 4) * 4;
  bf16x4 b[K / 16][N / 16];
#pragma unroll
  for (int kt = 0; kt < K / 16; ++kt)
#pragma unroll
    for (int nt = 0; nt < N / 16; ++nt)
      b[kt][nt] = *reinterpret_cast<const bf16x4*>(WT + (nt * 16 + row) * LDW + kt * 16 + kq);
  float bs[N / 16];
#pragma unroll
  for (int nt = 0; nt < N / 16; ++nt) bs[nt] = bias[nt * 16 + col];
  for (int s = 0; s < 4; ++s) {
    if (!((work >> s) & 1u)) continue;
    f32x4 acc[N / 16];
#pragma unroll
    for (int nt = 0; nt < N / 16; ++nt) acc[nt] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int kt = 0; kt < K / 16; ++kt) {
      const bf16x4 a = *reinterpret_cast<const bf16x4*>(In + (s * 16 + row) * ldin + kt * 16 + kq);
#pragma unroll
      for (int nt = 0; nt < N / 16; ++nt) acc[nt] = mfma16(a, b[kt][nt], acc[nt]);
    }
#pragma unroll
    for (int nt = 0; nt < N / 16; ++nt)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const float z = acc[nt][v] + bs[nt];
        if (LAST)
          Yout[(s * 16 + r0 + v) * kMlpOut + nt * 16 + col] = z;
        else
          Out[(s * 16 + r0 + v) * (N + kPad) + nt * 16 + col] = f2bf(z > 0.f ? z : 0.f);
      }
  }
}

// LDS of one wave's 64-row forward pass, in two buffers used in turn:
// A holds the inputs [64][kLdIn], then the second hidden layer, then the
// logits [64][16] FP32; B the first hidden layer, then the third.
constexpr int kMlpWaveBuf = kMlpRows * kLdHid;  // uint16 elements per buffer
static_assert(kMlpRows * kLdIn <= kMlpWaveBuf && kMlpRows * kMlpOut * 2 <= kMlpWaveBuf, "wave buffers");

// Forward pass of a wave's 64 rows with no block barrier: the wave owns A
// and B, and the layers are ordered by wavefront-scope fences.  Bit s of
// `work` asks for rows 16s .. 16s + 15.  On return the logits of row i are
// reinterpret_cast<float*>(A)[i * 16 .. i * 16 + 15].
__device__ __forceinline__ void mlp_forward64_wave(const MlpDev& M, uint16_t* A, uint16_t* B, unsigned work)
{
  mlp_layer64<32, 64, 32, false>(A, kLdIn, M.WT + kMlpWOff[0], M.b + 0, B, nullptr, work);
  __syncwarp();
  mlp_layer64<64, 64, 64, false>(B, kLdHid, M.WT + kMlpWOff[1], M.b + 64, A, nullptr, work);
  __syncwarp();
  mlp_layer64<64, 64, 64, false>(A, kLdHid, M.WT + kMlpWOff[2], M.b + 128, B, nullptr, work);
  __syncwarp();
  mlp_layer64<64, 16, 64, true>(B, kLdHid, M.WT + kMlpWOff[3], M.b + 192, nullptr, reinterpret_cast<float*>(A),
                                work);
  __syncwarp();
}

}  // namespace cuhre
