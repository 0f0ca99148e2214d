// Learned deferred shader for gfx950: an MLP 12 -> 64 -> 64 -> 64 -> 64 -> 3
// (ReLU hidden layers) on bfloat16 MFMA with float32 accumulation, with
// hand-written inference, training (forward + backward + weight gradients)
// and Adam kernels.  No reference counterpart: the reference shades with
// closed-form shaders only (diffren/jax/utils/shaders.py).
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
// Flat float32 parameters: W1[12,64] b1[64] W2[64,64] b2 W3 b3 W4 b4
// W5[64,3] b5[3], W[k_in][unit_out].
#pragma clang fp contract(off)

#include <algorithm>
#include <cmath>

#include "dr_common.h"
#include "../../include/diffren_amd.h"

namespace dr {
namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

constexpr int kIn = 12, kHid = 64, kOut = 3;
constexpr int kW[5] = {0, 832, 4992, 9152, 13312};
constexpr int kB[5] = {768, 4928, 9088, 13248, 13504};
constexpr int kParams = 13507;
constexpr int kInSize[5] = {kIn, kHid, kHid, kHid, kHid};
constexpr int kOutSize[5] = {kHid, kHid, kHid, kHid, kOut};
// Forward A blocks: L1 m<2 (1 k-step), L2..L4 m<2 s<4, L5 s<4.
constexpr int kFwdBase[5] = {0, 2, 10, 18, 26};
constexpr int kFwdBlocks = 30;
// Backward (d in = W d out) A blocks: L2..L4 m<2 s<4, L5 m<2 (1 k-step).
constexpr int kBwdBase[5] = {-1, 0, 8, 16, 24};
constexpr int kBwdBlocks = 26;
constexpr int kWaves = 4, kBlock = 64 * kWaves;

__device__ __forceinline__ int sigma(int k) {
  const int lo = k & 15;
  return (k & ~15) | (((lo >> 2) & 1) << 3) | (((lo >> 3) & 1) << 2) |
         (lo & 3);
}

__device__ __forceinline__ int c_unit(int m, int r, int h) {
  return 32 * m + (r >> 2) * 8 + 4 * h + (r & 3);
}

__device__ __forceinline__ f32x16 mfma(bf16x8 a, bf16x8 b, f32x16 c) {
  return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
}

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
        if (k < kIn) w = p[kW[0] + k * kHid + unit];
      } else {
        const int kk = sigma(k);
        if (unit < kOutSize[layer])
          w = p[kW[layer] + kk * kOutSize[layer] + unit];
      }
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

// Raw inputs of the layer-1 B operand: pixel p, features 8h..8h+7 (zero
// beyond the 12 features or the last pixel).
__device__ __forceinline__ void load_inputs(const float* __restrict__ x,
                                            int64_t p, int64_t N, int h,
                                            float raw[8]) {
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    const int f = 8 * h + j;
    raw[j] = (p < N && f < kIn) ? x[p * kIn + f] : 0.0f;
  }
}

// (raw - shift) * scale as bfloat16; nrm = [shift 12 | scale 12] in LDS.
__device__ __forceinline__ bf16x8 input_operand(const float raw[8],
                                                const float* nrm, int h) {
  bf16x8 b;
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    const int f = 8 * h + j;
    const float v = f < kIn ? (raw[j] - nrm[f]) * nrm[kIn + f] : 0.0f;
    b[j] = static_cast<__bf16>(v);
  }
  return b;
}

__device__ __forceinline__ void stage_norm(const float* __restrict__ norm,
                                           float* nrm) {
  for (int i = threadIdx.x; i < 2 * kIn; i += blockDim.x)
    nrm[i] = norm ? norm[i] : (i < kIn ? 0.0f : 1.0f);
}

// Albedo (features 9..11) of pixel p for the output lanes.
__device__ __forceinline__ void load_albedo(const float* __restrict__ x,
                                            int64_t p, int64_t N,
                                            float alb[3]) {
#pragma unroll
  for (int c = 0; c < 3; ++c) alb[c] = p < N ? x[p * kIn + 9 + c] : 0.0f;
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
// Inference: persistent waves; weights and biases in LDS (keeping the 30
// operand blocks in registers would leave one wave per SIMD); the next
// chunk's inputs are loaded while the current chunk runs on the matrix cores.

__global__ __launch_bounds__(kBlock) void forward_kernel(
    const bf16x8* __restrict__ wf, const float* __restrict__ params,
    const float* __restrict__ x, int64_t N, float* __restrict__ y,
    const float* __restrict__ norm, int flags) {
  __shared__ float sb[5 * 64];
  __shared__ float nrm[2 * kIn];
  __shared__ bf16x8 sw[kFwdBlocks * 64];
  stage_bias(params, sb);
  stage_norm(norm, nrm);
  for (int i = threadIdx.x; i < kFwdBlocks * 64; i += kBlock) sw[i] = wf[i];
  __syncthreads();
  const bool modulate = (flags & DR_MLP_ALBEDO_MODULATE) != 0;
  const int lane = threadIdx.x & 63, h = lane >> 5, col = lane & 31;
  const int64_t waves = int64_t(gridDim.x) * kWaves;
  int64_t chunk = int64_t(blockIdx.x) * kWaves + (threadIdx.x >> 6);
  float raw[8], alb[3];
  load_inputs(x, chunk * 32 + col, N, h, raw);
  if (modulate) load_albedo(x, chunk * 32 + col, N, alb);
  for (; chunk * 32 < N; chunk += waves) {
    // Keeps the weight reads from LDS inside the loop (hoisting them would
    // pin all 30 operand blocks in registers again).
    __asm__ volatile("" ::: "memory");
    const int64_t p = chunk * 32 + col;
    const bf16x8 b0 = input_operand(raw, nrm, h);
    float a[3] = {alb[0], alb[1], alb[2]};
    // Prefetch the next chunk (its loads are in bounds or masked).
    load_inputs(x, (chunk + waves) * 32 + col, N, h, raw);
    if (modulate) load_albedo(x, (chunk + waves) * 32 + col, N, alb);
    f32x16 acc[2];
#pragma unroll
    for (int m = 0; m < 2; ++m)
      acc[m] = mfma(sw[(kFwdBase[0] + m) * 64 + lane], b0,
                    bias_init(sb, 0, m, h));
    bf16x8 nb[4];
#pragma unroll
    for (int layer = 1; layer < 4; ++layer) {
      relu_operands(acc, nb);
#pragma unroll
      for (int m = 0; m < 2; ++m) {
        f32x16 c = bias_init(sb, layer, m, h);
#pragma unroll
        for (int s = 0; s < 4; ++s)
          c = mfma(sw[(kFwdBase[layer] + 4 * m + s) * 64 + lane], nb[s], c);
        acc[m] = c;
      }
    }
    relu_operands(acc, nb);
    f32x16 c = bias_init(sb, 4, 0, h);
#pragma unroll
    for (int s = 0; s < 4; ++s)
      c = mfma(sw[(kFwdBase[4] + s) * 64 + lane], nb[s], c);
    if (h == 0 && p < N) {
#pragma unroll
      for (int r = 0; r < 3; ++r)
        y[p * 3 + r] = modulate ? a[r] * c[r] : c[r];
    }
  }
}

// ---------------------------------------------------------------------------
// Training step: forward, masked L2 loss, backward, dW/db.
//
// LDS per wave: activations a0 [16][32], a1..a4 [64][32] and the current
// delta [64][32], all bfloat16 as [unit][pixel] so that 8 consecutive pixels
// of one unit -- an A or B operand with K = pixels -- are one 16-byte read.
// Per block: float32 gradient accumulators for all parameters.

constexpr int kActWave = (16 + 4 * 64) * 32;  // bf16 elements
constexpr int kDltWave = 64 * 32;

__device__ __forceinline__ void lds_fence() {
  // Every lane's LDS writes complete before any lane reads them back.
  __asm__ volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
  __builtin_amdgcn_wave_barrier();
}

__device__ __forceinline__ __bf16* act_ptr(__bf16* base, int layer) {
  // layer 0: a0 [16][32]; layer i>0: a_i [64][32]
  return layer == 0 ? base : base + 16 * 32 + (layer - 1) * 64 * 32;
}

__device__ __forceinline__ void store_tile(__bf16* arr, const f32x16& c, int m,
                                           int h, int col, bool relu) {
#pragma unroll
  for (int r = 0; r < 16; ++r) {
    const float v = relu ? fmaxf(c[r], 0.0f) : c[r];
    arr[c_unit(m, r, h) * 32 + col] = static_cast<__bf16>(v);
  }
}

__device__ __forceinline__ bf16x8 lds_operand(const __bf16* arr, int row,
                                              int nrows, int s, int h) {
  if (row >= nrows) {
    bf16x8 z;
#pragma unroll
    for (int j = 0; j < 8; ++j) z[j] = static_cast<__bf16>(0.0f);
    return z;
  }
  return *reinterpret_cast<const bf16x8*>(arr + row * 32 + 16 * s + 8 * h);
}

__global__ __launch_bounds__(kBlock) void train_kernel(
    const bf16x8* __restrict__ wf, const bf16x8* __restrict__ wb,
    const float* __restrict__ params, const float* __restrict__ x,
    const float* __restrict__ target, const float* __restrict__ mask,
    int64_t N, float scale, float* __restrict__ partials,
    const float* __restrict__ norm, int flags) {
  extern __shared__ __attribute__((aligned(16))) float lds[];
  float* g = lds;                       // [kParams] gradient accumulators
  float* sb = g + kParams + 1;          // [5 * 64] biases
  // kParams + 1 + 5 * 64 = 13828 floats: the bf16 arrays start 16-byte
  // aligned, as their 16-byte operand reads need.
  __bf16* act_all = reinterpret_cast<__bf16*>(sb + 5 * 64);
  const int wave = threadIdx.x >> 6;
  __bf16* act = act_all + wave * (kActWave + kDltWave);
  __bf16* dlt = act + kActWave;
  __shared__ float nrm[2 * kIn];
  __shared__ float wave_loss[kWaves];
  for (int i = threadIdx.x; i < kParams; i += kBlock) g[i] = 0.0f;
  stage_bias(params, sb);
  stage_norm(norm, nrm);
  __syncthreads();
  const bool modulate = (flags & DR_MLP_ALBEDO_MODULATE) != 0;

  const int lane = threadIdx.x & 63, h = lane >> 5, col = lane & 31;
  const int64_t waves = int64_t(gridDim.x) * kWaves;
  float partial = 0.0f;
  for (int64_t chunk = int64_t(blockIdx.x) * kWaves + wave; chunk * 32 < N;
       chunk += waves) {
    const int64_t p = chunk * 32 + col;
    // ---- forward, keeping activations
    float raw[8];
    load_inputs(x, p, N, h, raw);
    const bf16x8 b0 = input_operand(raw, nrm, h);
#pragma unroll
    for (int j = 0; j < 8; ++j) act[(8 * h + j) * 32 + col] = b0[j];
    f32x16 acc[2];
#pragma unroll
    for (int m = 0; m < 2; ++m)
      acc[m] = mfma(wf[(kFwdBase[0] + m) * 64 + lane], b0,
                    bias_init(sb, 0, m, h));
    bf16x8 nb[4];
    for (int layer = 1; layer < 4; ++layer) {
      relu_operands(acc, nb);
      for (int m = 0; m < 2; ++m) store_tile(act_ptr(act, layer), acc[m], m,
                                             h, col, true);
#pragma unroll
      for (int m = 0; m < 2; ++m) {
        f32x16 c = bias_init(sb, layer, m, h);
#pragma unroll
        for (int s = 0; s < 4; ++s)
          c = mfma(wf[(kFwdBase[layer] + 4 * m + s) * 64 + lane], nb[s], c);
        acc[m] = c;
      }
    }
    relu_operands(acc, nb);
    for (int m = 0; m < 2; ++m) store_tile(act_ptr(act, 4), acc[m], m, h, col,
                                           true);
    f32x16 y = bias_init(sb, 4, 0, h);
#pragma unroll
    for (int s = 0; s < 4; ++s)
      y = mfma(wf[(kFwdBase[4] + s) * 64 + lane], nb[s], y);
    // ---- loss gradient (output units 0..2 live in lanes h == 0, regs 0..2)
    f32x16 d[2];
#pragma unroll
    for (int r = 0; r < 16; ++r) { d[0][r] = 0.0f; d[1][r] = 0.0f; }
    if (h == 0 && p < N) {
      const float mk = mask ? mask[p] : 1.0f;
      float alb[3] = {1.0f, 1.0f, 1.0f};
      if (modulate) load_albedo(x, p, N, alb);
#pragma unroll
      for (int r = 0; r < 3; ++r) {
        // image = albedo * network output when modulating.
        const float e = alb[r] * y[r] - target[p * 3 + r];
        partial += mk * (e * e) * scale;
        d[0][r] = (((2.0f * e) * scale) * mk) * alb[r];
      }
    }
    // ---- backward, layer 5 .. 1
    for (int layer = 4; layer >= 0; --layer) {
      const int mt = layer == 4 ? 1 : 2;
      // bias gradient and delta to LDS
      for (int m = 0; m < mt; ++m) {
#pragma unroll
        for (int r = 0; r < 16; ++r) {
          const int u = c_unit(m, r, h);
          if (u < kOutSize[layer] && d[m][r] != 0.0f)
            atomicAdd(&g[kB[layer] + u], d[m][r]);
        }
        store_tile(dlt, d[m], m, h, col, false);
      }
      lds_fence();
      // dW[k_in][u_out] += sum over pixels a_{layer}[k_in] * delta[u_out]
      const __bf16* ain = act_ptr(act, layer);
      const int in_rows = layer == 0 ? 16 : 64;
      const int nt = layer == 0 ? 1 : 2;
      for (int m = 0; m < mt; ++m) {
        const bf16x8 a0 = lds_operand(dlt, 32 * m + col, 32 * mt, 0, h);
        const bf16x8 a1 = lds_operand(dlt, 32 * m + col, 32 * mt, 1, h);
        for (int n = 0; n < nt; ++n) {
          const int kin = 32 * n + col;
          f32x16 c;
#pragma unroll
          for (int r = 0; r < 16; ++r) c[r] = 0.0f;
          c = mfma(a0, lds_operand(ain, kin, in_rows, 0, h), c);
          c = mfma(a1, lds_operand(ain, kin, in_rows, 1, h), c);
          if (kin < kInSize[layer]) {
#pragma unroll
            for (int r = 0; r < 16; ++r) {
              const int u = c_unit(m, r, h);
              if (u < kOutSize[layer] && c[r] != 0.0f)
                atomicAdd(&g[kW[layer] + kin * kOutSize[layer] + u], c[r]);
            }
          }
        }
      }
      if (layer == 0) break;
      // d a_{layer} = W_layer d_out, then ReLU' from the stored activation.
      bf16x8 db[4];
      const int ks = layer == 4 ? 1 : 4;
#pragma unroll
      for (int m = 0; m < 2; ++m)
#pragma unroll
        for (int s = 0; s < 2; ++s)
#pragma unroll
          for (int j = 0; j < 8; ++j)
            db[2 * m + s][j] = static_cast<__bf16>(d[m][8 * s + j]);
      const __bf16* aprev = act_ptr(act, layer);
      f32x16 nd[2];
      for (int m = 0; m < 2; ++m) {
        f32x16 c;
#pragma unroll
        for (int r = 0; r < 16; ++r) c[r] = 0.0f;
        for (int s = 0; s < ks; ++s)
          c = mfma(wb[(kBwdBase[layer] + (layer == 4 ? m : 4 * m + s)) * 64 +
                      lane],
                   db[s], c);
#pragma unroll
        for (int r = 0; r < 16; ++r) {
          const float a = static_cast<float>(aprev[c_unit(m, r, h) * 32 + col]);
          c[r] = a > 0.0f ? c[r] : 0.0f;
        }
        nd[m] = c;
      }
      d[0] = nd[0];
      d[1] = nd[1];
      lds_fence();  // dlt is rewritten by the next layer
    }
  }
  for (int off = 32; off > 0; off >>= 1) partial += __shfl_xor(partial, off);
  if (lane == 0) wave_loss[wave] = partial;
  __syncthreads();
  // This block's gradient and loss, stored (reduced over blocks in a fixed
  // order by reduce_kernel: no same-address atomics across blocks).
  float* out = partials + int64_t(blockIdx.x) * (kParams + 1);
  for (int i = threadIdx.x; i < kParams; i += kBlock) out[i] = g[i];
  if (threadIdx.x == 0) {
    float t = 0.0f;
    for (int w = 0; w < kWaves; ++w) t += wave_loss[w];
    out[kParams] = t;
  }
}

// grad[i] += sum over blocks of partials[b][i]; *loss += sum of the losses.
__global__ __launch_bounds__(256) void reduce_kernel(
    const float* __restrict__ partials, int nblocks, float* __restrict__ grad,
    float* __restrict__ loss) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i > kParams) return;
  float t = 0.0f;
  for (int b = 0; b < nblocks; ++b) t += partials[int64_t(b) * (kParams + 1) + i];
  if (i < kParams) grad[i] += t;
  else if (loss) *loss += t;
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

static_assert((kParams + 1 + 5 * 64) % 4 == 0, "LDS bf16 stage alignment");
static_assert(((kActWave + kDltWave) * 2) % 16 == 0, "per-wave alignment");

size_t train_lds_bytes() {
  return sizeof(float) * (kParams + 1 + 5 * 64) +
         sizeof(__bf16) * kWaves * (kActWave + kDltWave);
}

}  // namespace
}  // namespace dr

extern "C" {

int32_t dr_mlp_num_params(void) { return dr::kParams; }

size_t dr_mlp_packed_bytes(void) {
  return size_t(dr::kFwdBlocks + dr::kBwdBlocks) * 64 * 16;
}

int dr_mlp_pack(const float* params, void* packed, dr_stream_t stream) {
  DR_CHECK_ARG(params && packed, "null pointer");
  DR_CHECK_ARG((reinterpret_cast<uintptr_t>(packed) & 15) == 0,
               "packed weights must be 16-byte aligned");
  auto* f = static_cast<dr::bf16x8*>(packed);
  dr::pack_kernel<<<dr::kFwdBlocks + dr::kBwdBlocks, 64, 0,
                    static_cast<hipStream_t>(stream)>>>(
      params, f, f + dr::kFwdBlocks * 64);
  DR_CHECK_LAUNCH("mlp pack");
  return DR_OK;
}

int dr_mlp_forward(const void* packed, const float* params, const float* x,
                   int64_t num_pixels, float* y, const float* input_norm,
                   int32_t flags, int32_t num_blocks, dr_stream_t stream) {
  DR_CHECK_ARG(num_pixels >= 0, "negative size");
  if (num_pixels == 0) return DR_OK;
  DR_CHECK_ARG(packed && params && x && y, "null pointer");
  const int64_t chunks = (num_pixels + 31) / 32;
  int64_t nb = num_blocks > 0 ? num_blocks : 2048;
  nb = std::min<int64_t>(nb, (chunks + dr::kWaves - 1) / dr::kWaves);
  dr::forward_kernel<<<unsigned(nb), dr::kBlock, 0,
                       static_cast<hipStream_t>(stream)>>>(
      static_cast<const dr::bf16x8*>(packed), params, x, num_pixels, y,
      input_norm, flags);
  DR_CHECK_LAUNCH("mlp forward");
  return DR_OK;
}

size_t dr_mlp_train_workspace_size(int32_t num_blocks) {
  const int nb = num_blocks > 0 ? num_blocks : 256;
  return size_t(nb) * (dr::kParams + 1) * sizeof(float);
}

int dr_mlp_train_step(const void* packed, const float* params, const float* x,
                      const float* target, const float* mask,
                      int64_t num_pixels, float loss_scale,
                      const float* input_norm, int32_t flags, float* grad,
                      float* loss, int32_t num_blocks, void* workspace,
                      size_t workspace_bytes, dr_stream_t stream) {
  DR_CHECK_ARG(num_pixels >= 0, "negative size");
  if (num_pixels == 0) return DR_OK;
  DR_CHECK_ARG(packed && params && x && target && grad, "null pointer");
  DR_CHECK_ARG(workspace != nullptr, "workspace is null");
  const int nb0 = num_blocks > 0 ? num_blocks : 256;
  if (workspace_bytes < dr_mlp_train_workspace_size(nb0))
    return dr::fail(dr::kWorkspaceTooSmall, "mlp train workspace too small");
  static bool lds_set = false;  // one attribute call per process
  const size_t lds = dr::train_lds_bytes();
  if (!lds_set) {
    DR_CHECK_HIP(hipFuncSetAttribute(
        reinterpret_cast<const void*>(&dr::train_kernel),
        hipFuncAttributeMaxDynamicSharedMemorySize, int(lds)));
    lds_set = true;
  }
  const int64_t chunks = (num_pixels + 31) / 32;
  const int nb = int(std::min<int64_t>(nb0, (chunks + dr::kWaves - 1) /
                                                dr::kWaves));
  const auto* f = static_cast<const dr::bf16x8*>(packed);
  float* partials = static_cast<float*>(workspace);
  hipStream_t s = static_cast<hipStream_t>(stream);
  dr::train_kernel<<<unsigned(nb), dr::kBlock, lds, s>>>(
      f, f + dr::kFwdBlocks * 64, params, x, target, mask, num_pixels,
      loss_scale, partials, input_norm, flags);
  DR_CHECK_LAUNCH("mlp train step");
  dr::reduce_kernel<<<(dr::kParams + 1 + 255) / 256, 256, 0, s>>>(
      partials, nb, grad, loss);
  DR_CHECK_LAUNCH("mlp train reduce");
  return DR_OK;
}

int dr_mlp_adam(float* params, const float* grad, float* m, float* v,
                int32_t step, float lr, float beta1, float beta2, float eps,
                dr_stream_t stream) {
  DR_CHECK_ARG(params && grad && m && v, "null pointer");
  DR_CHECK_ARG(step >= 1, "step counts from 1");
  const float c1 = 1.0f - std::pow(beta1, float(step));
  const float c2 = 1.0f - std::pow(beta2, float(step));
  dr::adam_kernel<<<(dr::kParams + 255) / 256, 256, 0,
                    static_cast<hipStream_t>(stream)>>>(
      params, grad, m, v, dr::kParams, lr, beta1, beta2, eps, c1, c2);
  DR_CHECK_LAUNCH("mlp adam");
  return DR_OK;
}

}  // extern "C"
