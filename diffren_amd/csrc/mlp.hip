// Learned deferred shader for gfx950: an MLP 12 -> 64 -> 64 -> 64 -> 64 -> 3
// (ReLU hidden layers) on bfloat16 MFMA with float32 accumulation: weight
// packing, inference and Adam kernels (the training step is in
// mlp_train.hip; the shared operand layout in mlp_common.h).  No reference
// counterpart: the reference shades with closed-form shaders only
// (diffren/jax/utils/shaders.py).
#pragma clang fp contract(off)

#include <algorithm>
#include <cmath>

#include "dr_common.h"
#include "mlp_common.h"
#include "../../include/diffren_amd.h"

namespace dr {
namespace mlp {
namespace {

constexpr int kWaves = 4, kBlock = 64 * kWaves;

// ---------------------------------------------------------------------------
// Packing float32 master weights into MFMA operand order.

__global__ void pack_kernel(const float* __restrict__ p,
                            bf16x8* __restrict__ fwd,
                            bf16x8* __restrict__ bwd) {
  const int blk = blockIdx.x, lane = threadIdx.x;
  const int row = lane & 31, h = lane >> 5;
  bf16x8 v;
  if (blk < kFwdBlocks) {
    int layer, m, s;
    if (blk < 2) { layer = 0; m = blk; s = 0; }
    else if (blk < 26) { layer = 1 + (blk - 2) / 8; m = ((blk - 2) % 8) / 4; s = (blk - 2) % 4; }
    else { layer = 4; m = 0; s = blk - 26; }
    const int unit = 32 * m + row;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const int k = 16 * s + 8 * h + j;
      float w = 0.0f;
      if (layer == 0) {
        if (k < kIn) {
          w = p[kW[0] + k * kHid + unit];
        } else if (k < kBiasSlot + 3) {
          w = bias_part(p[kB[0] + unit], k - kBiasSlot);
        }
      } else {
        const int kk = sigma(k);
        if (unit < kOutSize[layer])
          w = p[kW[layer] + kk * kOutSize[layer] + unit];
      }
      v[j] = static_cast<__bf16>(w);
    }
    fwd[blk * 64 + lane] = v;
  } else if (blk >= kBiasBase) {
    // Layer 2..5 bias parts in k-slots 0..2 (lane half 0).
    const int b = blk - kBiasBase;
    const int layer = 1 + (b < 6 ? b / 2 : 3), m = b < 6 ? b % 2 : 0;
    const int unit = 32 * m + row;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      float w = 0.0f;
      if (h == 0 && j < 3 && unit < kOutSize[layer])
        w = bias_part(p[kB[layer] + unit], j);
      v[j] = static_cast<__bf16>(w);
    }
    fwd[blk * 64 + lane] = v;
  } else {
    const int b = blk - kFwdBlocks;
    int layer, m, s;
    if (b < 24) { layer = 1 + b / 8; m = (b % 8) / 4; s = b % 4; }
    else { layer = 4; m = b - 24; s = 0; }
    const int in = 32 * m + row;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const int kk = sigma(16 * s + 8 * h + j);
      float w = 0.0f;
      if (kk < kOutSize[layer])
        w = p[kW[layer] + in * kOutSize[layer] + kk];
      v[j] = static_cast<__bf16>(w);
    }
    bwd[b * 64 + lane] = v;
  }
}

// ---------------------------------------------------------------------------
// Forward building blocks.

// (raw - shift) * scale as bfloat16, then 1 in the bias slots 12..14; nrm =
// [shift 12 | scale 12] in LDS.
__device__ __forceinline__ bf16x8 input_operand(const float raw[8],
                                                const float* nrm, int h) {
  bf16x8 b;
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    const int f = 8 * h + j;
    const float v = f < kIn ? (raw[j] - nrm[f]) * nrm[kIn + f]
                            : (f < kBiasSlot + 3 ? 1.0f : 0.0f);
    b[j] = static_cast<__bf16>(v);
  }
  return b;
}

__device__ __forceinline__ void stage_norm(const float* __restrict__ norm,
                                           float* nrm) {
  for (int i = threadIdx.x; i < 2 * kIn; i += blockDim.x)
    nrm[i] = norm ? norm[i] : (i < kIn ? 0.0f : 1.0f);
}

// ReLU of two accumulator tiles as the next layer's four B operands.
__device__ __forceinline__ void relu_operands(const f32x16 acc[2],
                                              bf16x8 nb[4]) {
#pragma unroll
  for (int m = 0; m < 2; ++m)
#pragma unroll
    for (int s = 0; s < 2; ++s)
#pragma unroll
      for (int j = 0; j < 8; ++j)
        nb[2 * m + s][j] = static_cast<__bf16>(fmaxf(acc[m][8 * s + j], 0.0f));
}

__device__ __forceinline__ void stage_bias(const float* __restrict__ params,
                                           float* sb) {
  // sb[layer * 64 + u] = b_layer[u]
  for (int i = threadIdx.x; i < 5 * 64; i += blockDim.x) {
    const int layer = i / 64, u = i % 64;
    sb[i] = u < kOutSize[layer] ? params[kB[layer] + u] : 0.0f;
  }
}

__device__ __forceinline__ f32x16 bias_init(const float* sb, int layer, int m,
                                            int h) {
  f32x16 c;
#pragma unroll
  for (int r = 0; r < 16; ++r) c[r] = sb[layer * 64 + c_unit(m, r, h)];
  return c;
}

// ---------------------------------------------------------------------------
// Inference: persistent waves, each on two 32-pixel chunks at a time (every
// weight and bias read from LDS feeds two independent accumulator chains);
// weights and biases in LDS (keeping the 30 operand blocks in registers would
// leave one wave per SIMD).  A chunk pair's 64 x 12 input floats are one
// contiguous 3 KB span: they are loaded as three coalesced 16-byte loads per
// lane (issued one pair ahead, while the current pair runs on the matrix
// cores) and regrouped through a per-wave LDS stage into operand order.

constexpr int kStage = 64 * kIn / 4;  // float4 per wave stage

struct Fetch {
  float4 v0, v1, v2;
};

__device__ __forceinline__ float4 fetch4(const float4* __restrict__ x4,
                                         int64_t k, int64_t N) {
  // float4 k lies inside pixel k / 3.
  float4 t = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  if (k < 3 * N) t = x4[k];
  return t;
}

__device__ __forceinline__ Fetch fetch_pair(const float4* __restrict__ x4,
                                            int64_t pair, int64_t N,
                                            int lane) {
  const int64_t k = pair * kStage + lane;
  return Fetch{fetch4(x4, k, N), fetch4(x4, k + 64, N),
               fetch4(x4, k + 128, N)};
}

__global__ __launch_bounds__(kBlock) void forward_kernel(
    const bf16x8* __restrict__ wf, const float* __restrict__ params,
    const float* __restrict__ x, int64_t N, float* __restrict__ y,
    const float* __restrict__ norm, int flags) {
  __shared__ float sb[5 * 64];
  __shared__ float nrm[2 * kIn];
  __shared__ bf16x8 sw[kFwdBlocks * 64];
  __shared__ float4 stage_all[kWaves * kStage];
  stage_bias(params, sb);
  stage_norm(norm, nrm);
  for (int i = threadIdx.x; i < kFwdBlocks * 64; i += kBlock) sw[i] = wf[i];
  __syncthreads();
  const bool modulate = (flags & DR_MLP_ALBEDO_MODULATE) != 0;
  const int lane = threadIdx.x & 63, h = lane >> 5, col = lane & 31;
  float4* stage = stage_all + (threadIdx.x >> 6) * kStage;
  const float* stage_f = reinterpret_cast<const float*>(stage);
  const float4* x4 = reinterpret_cast<const float4*>(x);
  const int64_t waves = int64_t(gridDim.x) * kWaves;
  int64_t pair = int64_t(blockIdx.x) * kWaves + (threadIdx.x >> 6);
  Fetch v = fetch_pair(x4, pair, N, lane);
  for (; pair * 64 < N; pair += waves) {
    // Keeps the weight reads from LDS inside the loop (hoisting them would
    // pin all 30 operand blocks in registers again).
    __asm__ volatile("" ::: "memory");
    stage[lane] = v.v0;
    stage[64 + lane] = v.v1;
    stage[128 + lane] = v.v2;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    bf16x8 b0[2];
    float a[2][3];
#pragma unroll
    for (int i = 0; i < 2; ++i) {
      const float* px = stage_f + (32 * i + col) * kIn;
      float raw[8];
#pragma unroll
      for (int j = 0; j < 8; ++j) raw[j] = 8 * h + j < kIn ? px[8 * h + j] : 0.0f;
      b0[i] = input_operand(raw, nrm, h);
#pragma unroll
      for (int c = 0; c < 3; ++c) a[i][c] = px[9 + c];
    }
    // The next pair's inputs (in bounds or zero) while this pair computes.
    v = fetch_pair(x4, pair + waves, N, lane);
    f32x16 acc[2][2];
#pragma unroll
    for (int m = 0; m < 2; ++m) {
      const bf16x8 w = sw[(kFwdBase[0] + m) * 64 + lane];
#pragma unroll
      for (int i = 0; i < 2; ++i) acc[i][m] = mfma(w, b0[i], zero16());
    }
    bf16x8 nb[2][4];
#pragma unroll
    for (int layer = 1; layer < 4; ++layer) {
#pragma unroll
      for (int i = 0; i < 2; ++i) relu_operands(acc[i], nb[i]);
#pragma unroll
      for (int m = 0; m < 2; ++m) {
        f32x16 c[2];
        c[0] = bias_init(sb, layer, m, h);
        c[1] = c[0];
#pragma unroll
        for (int s = 0; s < 4; ++s) {
          const bf16x8 w = sw[(kFwdBase[layer] + 4 * m + s) * 64 + lane];
#pragma unroll
          for (int i = 0; i < 2; ++i) c[i] = mfma(w, nb[i][s], c[i]);
        }
#pragma unroll
        for (int i = 0; i < 2; ++i) acc[i][m] = c[i];
      }
    }
#pragma unroll
    for (int i = 0; i < 2; ++i) relu_operands(acc[i], nb[i]);
    f32x16 c[2];
    c[0] = bias_init(sb, 4, 0, h);
    c[1] = c[0];
#pragma unroll
    for (int s = 0; s < 4; ++s) {
      const bf16x8 w = sw[(kFwdBase[4] + s) * 64 + lane];
#pragma unroll
      for (int i = 0; i < 2; ++i) c[i] = mfma(w, nb[i][s], c[i]);
    }
#pragma unroll
    for (int i = 0; i < 2; ++i) {
      const int64_t p = (2 * pair + i) * 32 + col;
      if (h == 0 && p < N) {
#pragma unroll
        for (int r = 0; r < 3; ++r)
          y[p * 3 + r] = modulate ? a[i][r] * c[i][r] : c[i][r];
      }
    }
  }
}

// Adam on float32 master weights (Kingma & Ba, 2015).
__global__ void adam_kernel(float* __restrict__ p, const float* __restrict__ g,
                            float* __restrict__ m, float* __restrict__ v,
                            int n, float lr, float b1, float b2, float eps,
                            float c1, float c2) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.0f - b1) * gi;
  const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  p[i] = p[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

}  // namespace
}  // namespace mlp
}  // namespace dr

extern "C" {

int32_t dr_mlp_num_params(void) { return dr::mlp::kParams; }

size_t dr_mlp_packed_bytes(void) {
  return size_t(dr::mlp::kPackedBlocks) * 64 * 16;
}

int dr_mlp_pack(const float* params, void* packed, dr_stream_t stream) {
  DR_CHECK_ARG(params && packed, "null pointer");
  DR_CHECK_ARG((reinterpret_cast<uintptr_t>(packed) & 15) == 0,
               "packed weights must be 16-byte aligned");
  auto* f = static_cast<dr::mlp::bf16x8*>(packed);
  dr::mlp::pack_kernel<<<dr::mlp::kPackedBlocks, 64, 0,
                    static_cast<hipStream_t>(stream)>>>(
      params, f, f + dr::mlp::kFwdBlocks * 64);
  DR_CHECK_LAUNCH("mlp pack");
  return DR_OK;
}

int dr_mlp_forward(const void* packed, const float* params, const float* x,
                   int64_t num_pixels, float* y, const float* input_norm,
                   int32_t flags, int32_t num_blocks, dr_stream_t stream) {
  DR_CHECK_ARG(num_pixels >= 0, "negative size");
  if (num_pixels == 0) return DR_OK;
  DR_CHECK_ARG(packed && params && x && y, "null pointer");
  DR_CHECK_ARG((reinterpret_cast<uintptr_t>(x) & 15) == 0,
               "features must be 16-byte aligned");
  const int64_t pairs = (num_pixels + 63) / 64;
  int64_t nb = num_blocks > 0 ? num_blocks : 2048;
  nb = std::min<int64_t>(nb, (pairs + dr::mlp::kWaves - 1) / dr::mlp::kWaves);
  dr::mlp::forward_kernel<<<unsigned(nb), dr::mlp::kBlock, 0,
                       static_cast<hipStream_t>(stream)>>>(
      static_cast<const dr::mlp::bf16x8*>(packed), params, x, num_pixels, y,
      input_norm, flags);
  DR_CHECK_LAUNCH("mlp forward");
  return DR_OK;
}

int dr_mlp_adam(float* params, const float* grad, float* m, float* v,
                int32_t step, float lr, float beta1, float beta2, float eps,
                dr_stream_t stream) {
  DR_CHECK_ARG(params && grad && m && v, "null pointer");
  DR_CHECK_ARG(step >= 1, "step counts from 1");
  const float c1 = 1.0f - std::pow(beta1, float(step));
  const float c2 = 1.0f - std::pow(beta2, float(step));
  dr::mlp::adam_kernel<<<(dr::mlp::kParams + 255) / 256, 256, 0,
                    static_cast<hipStream_t>(stream)>>>(
      params, grad, m, v, dr::mlp::kParams, lr, beta1, beta2, eps, c1, c2);
  DR_CHECK_LAUNCH("mlp adam");
  return DR_OK;
}

}  // extern "C"
