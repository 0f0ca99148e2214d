// Shared layout of the learned deferred shader (MLP 12 -> 64 -> 64 -> 64 ->
// 64 -> 3, ReLU hidden layers, bfloat16 matrix-core operands, float32
// accumulation) between the inference/packing kernels (mlp.hip) and the
// training kernel (mlp_train.hip).
//
// Matrix-core mapping (v_mfma_f32_32x32x16_bf16, one wave):
//   A[32 x 16]: lane l holds row l%32, k = 8(l/32) + j, j < 8
//   B[16 x 32]: lane l holds col l%32, k = 8(l/32) + j
//   C[32 x 32]: lane l holds col l%32, reg r -> row (r/4)8 + 4(l/32) + r%4
// Each layer is computed transposed, out^T[unit][pixel] = W^T in^T, on
// chunks of 32 pixels (the N dimension).  Accumulator register r of unit
// tile m is fed straight into the next layer's B operand, k-step 2m + r/8,
// slot r%8.  That puts previous-layer unit u at physical k index sigma(u)
// with sigma swapping bits 2 and 3 of the index inside each 16-block, so the
// packed weights of every layer fed by an accumulator are stored with that K
// permutation and no lane shuffles or LDS round trips are needed between
// layers.
//
// Layer 1 carries its bias inside the matrix product: input k-slots 12, 13
// and 14 hold 1.0 and the packed W1 block holds b1 split into three
// bfloat16 parts (hi + mid + lo == b1 exactly), so layer 1 starts from a
// zero accumulator.  Layers 2..5 start from their float32 bias.
//
// Flat float32 parameters: W1[12,64] b1[64] W2[64,64] b2 W3 b3 W4 b4
// W5[64,3] b5[3], W[k_in][unit_out].
#pragma once

#include <cstdint>

namespace dr {
namespace mlp {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

constexpr int kIn = 12, kHid = 64, kOut = 3;
constexpr int kW[5] = {0, 832, 4992, 9152, 13312};
constexpr int kB[5] = {768, 4928, 9088, 13248, 13504};
constexpr int kParams = 13507;
constexpr int kInSize[5] = {kIn, kHid, kHid, kHid, kHid};
constexpr int kOutSize[5] = {kHid, kHid, kHid, kHid, kOut};
// k-slot of layer 1 holding the first bias part (then mid, lo).
constexpr int kBiasSlot = 12;
// Forward A blocks: L1 m<2 (1 k-step), L2..L4 m<2 s<4, L5 s<4.
constexpr int kFwdBase[5] = {0, 2, 10, 18, 26};
constexpr int kFwdBlocks = 30;
// Backward (d in = W d out) A blocks: L2..L4 m<2 s<4, L5 m<2 (1 k-step).
constexpr int kBwdBase[5] = {-1, 0, 8, 16, 24};
constexpr int kBwdBlocks = 26;
// Bias A blocks (training): layers 2..4 m<2, layer 5; row u, k-slots 0..2
// hold the bias of unit 32m + u as hi + mid + lo (see above), so that a
// product with a B operand of ones in k-slots 0..2 starts the accumulator
// at the float32 bias.
constexpr int kBiasBase = kFwdBlocks + kBwdBlocks;
constexpr int kBiasBlocks = 7;
constexpr int kPackedBlocks = kBiasBase + kBiasBlocks;

__device__ __forceinline__ int sigma(int k) {
  const int lo = k & 15;
  return (k & ~15) | (((lo >> 2) & 1) << 3) | (((lo >> 3) & 1) << 2) |
         (lo & 3);
}

// Unit held by accumulator register r of unit tile m in lane half h.
__host__ __device__ __forceinline__ int c_unit(int m, int r, int h) {
  return 32 * m + (r >> 2) * 8 + 4 * h + (r & 3);
}

__device__ __forceinline__ f32x16 mfma(bf16x8 a, bf16x8 b, f32x16 c) {
  return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
}

// b = hi + mid + lo, each exactly a bfloat16 (b's 24 significant bits in
// three 8-bit pieces; exact for normal b).
__device__ __forceinline__ float bias_part(float b, int part) {
  const float hi = static_cast<float>(static_cast<__bf16>(b));
  const float r1 = b - hi;
  const float mid = static_cast<float>(static_cast<__bf16>(r1));
  const float lo = r1 - mid;
  return part == 0 ? hi : (part == 1 ? mid : lo);
}

__device__ __forceinline__ f32x16 zero16() {
  f32x16 z;
#pragma unroll
  for (int r = 0; r < 16; ++r) z[r] = 0.0f;
  return z;
}

}  // namespace mlp
}  // namespace dr
