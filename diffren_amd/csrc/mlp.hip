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
      const f32x16 bias = bias_init(sb, 0, m, h);
      const bf16x8 w = sw[(kFwdBase[0] + m) * 64 + lane];
#pragma unroll
      for (int i = 0; i < 2; ++i) acc[i][m] = mfma(w, b0[i], bias);
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

// ---------------------------------------------------------------------------
// Training step: forward, masked L2 loss, backward, dW/db.
//
// A block of four waves works on rounds of 128 pixels, 32 per wave.  Each
// wave runs the forward and the input-gradient (d in = W d out) chain of its
// own 32 pixels on the matrix cores, as in inference, and writes the layer
// inputs and output deltas of those pixels to LDS as bfloat16 [unit][pixel].
// The weight gradients dW_l[in][out] = sum_pixels a_l[in] delta_l[out] are
// 32x32 tiles with the 128 pixels of a round as K; every tile has one owner
// wave that keeps it in accumulator registers across all rounds of the block
// (layers 1..3: wave w owns out tile w&1, in tile w>>1; layer 0: waves 0, 1;
// layer 4: waves 2, 3), so no gradient is ever added through memory until
// the block writes its partial sums once at the end.  Bias gradients are the
// row sums of the delta operands the owners of in tile 0 read anyway.
//
// LDS: the 56 packed weight blocks, the five layer inputs a0 [16][128],
// a1..a4 [64][128] and two delta buffers [64][128] (layer l uses buffer
// l&1; one barrier per layer orders them).  Row r of an array keeps its
// 16-byte groups of 8 pixels XOR-swizzled by r&15: operand reads (32 rows,
// one group) and tile writes (rows r and r+4 from the two lane halves) are
// then spread over all banks.

constexpr int kRound = 32 * kWaves;  // pixels per block round
constexpr int kActRows = 16 + 4 * kHid;
constexpr size_t kTrainLdsBytes =
    size_t(kFwdBlocks + kBwdBlocks) * 64 * 16 +      // weights
    size_t(kActRows + 2 * kHid) * kRound * 2;        // a0..a4, 2 deltas

typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));

__device__ __forceinline__ int swz(int row, int px) {
  return row * kRound + ((((px >> 3) ^ (row & 15)) << 3) | (px & 7));
}

// The lane-dependent parts of every swz() address of one iteration, so that
// a tile write or operand read is a register plus a compile-time offset:
//  * C register r of tile m sits in row 32m + 8(r>>2) + 4h + (r&3), whose
//    low 4 bits are 4h | c with c = 8((r>>2)&1) + (r&3), disjoint from 4h:
//    w[j] holds 128 * 4h + the swizzled column for c = c(j);
//  * operand rows are 32 * tile + col, low 4 bits col & 15: r[s] holds
//    128 col + the swizzled 8-pixel group of k-step s.
struct LaneOfs {
  int w[8];
  int r[kRound / 16];
};

__device__ __forceinline__ LaneOfs lane_offsets(int h, int col, int px) {
  LaneOfs o;
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    const int c = 8 * (j >> 2) + (j & 3);
    o.w[j] = 512 * h + ((((px >> 3) ^ (4 * h) ^ c) << 3) | (px & 7));
  }
#pragma unroll
  for (int s = 0; s < kRound / 16; ++s)
    o.r[s] = col * kRound + (((2 * s + h) ^ (col & 15)) << 3);
  return o;
}

__device__ __forceinline__ void put_tile(__bf16* arr, const f32x16& c, int m,
                                         const LaneOfs& o, bool relu) {
#pragma unroll
  for (int r = 0; r < 16; ++r) {
    const float v = relu ? fmaxf(c[r], 0.0f) : c[r];
    arr[kRound * (32 * m + 8 * (r >> 2) + (r & 3)) +
        o.w[4 * ((r >> 2) & 1) + (r & 3)]] = static_cast<__bf16>(v);
  }
}

// Operand with K = pixels 16s..16s+15 of the round: lane row 32 tile + col,
// 8 pixels (zero beyond nrows).
__device__ __forceinline__ bf16x8 px_operand(const __bf16* arr, int tile,
                                             int nrows, int s,
                                             const LaneOfs& o, int col) {
  if (32 * tile + col >= nrows) {
    bf16x8 z;
#pragma unroll
    for (int j = 0; j < 8; ++j) z[j] = static_cast<__bf16>(0.0f);
    return z;
  }
  return *reinterpret_cast<const bf16x8*>(arr + 32 * tile * kRound + o.r[s]);
}

__device__ __forceinline__ float row_sum(const bf16x8& a, float acc) {
  const bf16x2 one = {static_cast<__bf16>(1.0f), static_cast<__bf16>(1.0f)};
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const bf16x2 v = {a[2 * j], a[2 * j + 1]};
    acc = __builtin_amdgcn_fdot2_f32_bf16(v, one, acc, false);
  }
  return acc;
}

__device__ __forceinline__ uint32_t relu_bits(const f32x16 acc[2]) {
  uint32_t b = 0;
#pragma unroll
  for (int m = 0; m < 2; ++m)
#pragma unroll
    for (int r = 0; r < 16; ++r)
      b |= (acc[m][r] > 0.0f ? 1u : 0u) << (16 * m + r);
  return b;
}

// dW_L tile (out tile m, in tile n) += delta_L a_L^T over the round's pixels;
// with `bias`, also the delta row sums.
template <int L>
__device__ __forceinline__ void weight_grad(f32x16& acc, float& bacc,
                                            const __bf16* dlt,
                                            const __bf16* ain, int m, int n,
                                            const LaneOfs& o, int col,
                                            bool bias) {
  constexpr int out_rows = kOutSize[L];
  constexpr int in_rows = L == 0 ? 16 : kHid;
#pragma unroll
  for (int s = 0; s < kRound / 16; ++s) {
    const bf16x8 a = px_operand(dlt, m, out_rows, s, o, col);
    const bf16x8 b = px_operand(ain, n, in_rows, s, o, col);
    acc = mfma(a, b, acc);
    if (bias) bacc = row_sum(a, bacc);
  }
}

// d a_L = W_L d out, as the delta of layer L-1 (ReLU' from its forward
// pre-activation signs, bit 16m + r of `bits`).
template <int L>
__device__ __forceinline__ void input_grad(const bf16x8* sw, const f32x16 d[2],
                                           uint32_t bits, int lane,
                                           f32x16 nd[2]) {
  bf16x8 db[4];
#pragma unroll
  for (int m = 0; m < 2; ++m)
#pragma unroll
    for (int s = 0; s < 2; ++s)
#pragma unroll
      for (int j = 0; j < 8; ++j)
        db[2 * m + s][j] = static_cast<__bf16>(d[m][8 * s + j]);
  const bf16x8* wb = sw + kFwdBlocks * 64;
#pragma unroll
  for (int m = 0; m < 2; ++m) {
    f32x16 c;
#pragma unroll
    for (int r = 0; r < 16; ++r) c[r] = 0.0f;
    if (L == 4) {
      c = mfma(wb[(kBwdBase[4] + m) * 64 + lane], db[0], c);
    } else {
#pragma unroll
      for (int s = 0; s < 4; ++s)
        c = mfma(wb[(kBwdBase[L] + 4 * m + s) * 64 + lane], db[s], c);
    }
#pragma unroll
    for (int r = 0; r < 16; ++r)
      c[r] = (bits >> (16 * m + r)) & 1u ? c[r] : 0.0f;
    nd[m] = c;
  }
}

// One backward layer L: barrier (delta_L complete in dlt[L&1]), the owned
// dW_L tile, then delta_{L-1} (computed before the barrier) to dlt[(L-1)&1].
template <int L>
__device__ __forceinline__ void backward_layer(
    f32x16 acc[4], float bacc[5], __bf16* act, __bf16* dlt, f32x16 d[2],
    const f32x16 nd[2], int wave, const LaneOfs& o, int col) {
  __syncthreads();
  const __bf16* ain = L == 0 ? act : act + (16 + (L - 1) * kHid) * kRound;
  const __bf16* dl = dlt + (L & 1) * kHid * kRound;
  if (L >= 1 && L <= 3) {
    weight_grad<L>(acc[L - 1], bacc[L], dl, ain, wave & 1, wave >> 1, o, col,
                   (wave >> 1) == 0);
  } else if (L == 0) {
    if (wave < 2) weight_grad<0>(acc[3], bacc[0], dl, ain, wave, 0, o, col,
                                 true);
  } else {
    if (wave >= 2) weight_grad<4>(acc[3], bacc[4], dl, ain, 0, wave - 2, o,
                                  col, wave == 2);
  }
  if (L > 0) {
    __bf16* next = dlt + ((L - 1) & 1) * kHid * kRound;
#pragma unroll
    for (int m = 0; m < 2; ++m) {
      put_tile(next, nd[m], m, o, false);
      d[m] = nd[m];
    }
  }
}

// Block partial sums of one dW tile and (if kept) of its bias rows.
template <int L>
__device__ __forceinline__ void write_tile(float* out, const f32x16& acc,
                                           int m, int n, int h, int col) {
  const int in = 32 * n + col;
  if (in >= kInSize[L]) return;
#pragma unroll
  for (int r = 0; r < 16; ++r) {
    const int u = c_unit(m, r, h);
    if (u < kOutSize[L]) out[kW[L] + in * kOutSize[L] + u] = acc[r];
  }
}

template <int L>
__device__ __forceinline__ void write_bias(float* out, float bacc, int m,
                                           int h, int col) {
  const float t = bacc + __shfl_xor(bacc, 32);
  const int u = 32 * m + col;
  if (h == 0 && u < kOutSize[L]) out[kB[L] + u] = t;
}

__global__ __launch_bounds__(kBlock) void train_kernel(
    const bf16x8* __restrict__ wf, const bf16x8* __restrict__ wb,
    const float* __restrict__ params, const float* __restrict__ x,
    const float* __restrict__ target, const float* __restrict__ mask,
    int64_t N, float scale, float* __restrict__ partials,
    const float* __restrict__ norm, int flags) {
  extern __shared__ __attribute__((aligned(16))) bf16x8 sw[];
  __bf16* act = reinterpret_cast<__bf16*>(sw + (kFwdBlocks + kBwdBlocks) * 64);
  __bf16* dlt = act + kActRows * kRound;
  __shared__ float sb[5 * 64];
  __shared__ float nrm[2 * kIn];
  __shared__ float wave_loss[kWaves];
  for (int i = threadIdx.x; i < kFwdBlocks * 64; i += kBlock) sw[i] = wf[i];
  for (int i = threadIdx.x; i < kBwdBlocks * 64; i += kBlock)
    sw[kFwdBlocks * 64 + i] = wb[i];
  stage_bias(params, sb);
  stage_norm(norm, nrm);
  const bool modulate = (flags & DR_MLP_ALBEDO_MODULATE) != 0;

  const int wave = threadIdx.x >> 6;
  const int lane = threadIdx.x & 63;
  f32x16 acc[4];
#pragma unroll
  for (int t = 0; t < 4; ++t)
#pragma unroll
    for (int r = 0; r < 16; ++r) acc[t][r] = 0.0f;
  float bacc[5] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
  float partial = 0.0f;

  for (int64_t q = blockIdx.x; q * kRound < N; q += gridDim.x) {
    // The previous round's owners are done with a*, delta buffers.
    __syncthreads();
    // Lane coordinates, opaque to the compiler so that it derives the LDS
    // addresses inside the loop instead of keeping hundreds of them live.
    int h = lane >> 5, col = lane & 31;
    __asm__ volatile("" : "+v"(h), "+v"(col));
    const int px = 32 * wave + col;  // pixel within the round
    const int64_t p = q * kRound + px;
    const LaneOfs o = lane_offsets(h, col, px);
    // ---- forward, keeping layer inputs in LDS and ReLU signs in bits
    float raw[8];
    load_inputs(x, p, N, h, raw);
    const bf16x8 b0 = input_operand(raw, nrm, h);
#pragma unroll
    for (int j = 0; j < 8; ++j) act[swz(8 * h + j, px)] = b0[j];
    f32x16 a[2];
#pragma unroll
    for (int m = 0; m < 2; ++m)
      a[m] = mfma(sw[(kFwdBase[0] + m) * 64 + lane], b0,
                  bias_init(sb, 0, m, h));
    uint32_t bits[4];
    bf16x8 nb[4];
#pragma unroll
    for (int layer = 1; layer < 4; ++layer) {
      bits[layer - 1] = relu_bits(a);
      relu_operands(a, nb);
      __bf16* al = act + (16 + (layer - 1) * kHid) * kRound;
#pragma unroll
      for (int m = 0; m < 2; ++m) put_tile(al, a[m], m, o, true);
#pragma unroll
      for (int m = 0; m < 2; ++m) {
        f32x16 c = bias_init(sb, layer, m, h);
#pragma unroll
        for (int s = 0; s < 4; ++s)
          c = mfma(sw[(kFwdBase[layer] + 4 * m + s) * 64 + lane], nb[s], c);
        a[m] = c;
      }
    }
    bits[3] = relu_bits(a);
    relu_operands(a, nb);
    {
      __bf16* a4 = act + (16 + 3 * kHid) * kRound;
#pragma unroll
      for (int m = 0; m < 2; ++m) put_tile(a4, a[m], m, o, true);
    }
    f32x16 y = bias_init(sb, 4, 0, h);
#pragma unroll
    for (int s = 0; s < 4; ++s)
      y = mfma(sw[(kFwdBase[4] + s) * 64 + lane], nb[s], y);
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
    if (h == 0) {
#pragma unroll
      for (int r = 0; r < 3; ++r)
        dlt[swz(r, px)] = static_cast<__bf16>(d[0][r]);
    }
    // ---- backward
    f32x16 nd[2];
    input_grad<4>(sw, d, bits[3], lane, nd);
    backward_layer<4>(acc, bacc, act, dlt, d, nd, wave, o, col);
    input_grad<3>(sw, d, bits[2], lane, nd);
    backward_layer<3>(acc, bacc, act, dlt, d, nd, wave, o, col);
    input_grad<2>(sw, d, bits[1], lane, nd);
    backward_layer<2>(acc, bacc, act, dlt, d, nd, wave, o, col);
    input_grad<1>(sw, d, bits[0], lane, nd);
    backward_layer<1>(acc, bacc, act, dlt, d, nd, wave, o, col);
    backward_layer<0>(acc, bacc, act, dlt, d, nd, wave, o, col);
  }
  for (int off = 32; off > 0; off >>= 1) partial += __shfl_xor(partial, off);
  if (lane == 0) wave_loss[wave] = partial;
  const int h = lane >> 5, col = lane & 31;
  // This block's gradient and loss (every parameter has exactly one writer;
  // reduce_kernel sums the blocks in a fixed order).
  float* out = partials + int64_t(blockIdx.x) * (kParams + 1);
  write_tile<1>(out, acc[0], wave & 1, wave >> 1, h, col);
  write_tile<2>(out, acc[1], wave & 1, wave >> 1, h, col);
  write_tile<3>(out, acc[2], wave & 1, wave >> 1, h, col);
  if (wave < 2) {
    write_tile<0>(out, acc[3], wave, 0, h, col);
    write_bias<0>(out, bacc[0], wave, h, col);
    write_bias<1>(out, bacc[1], wave, h, col);
    write_bias<2>(out, bacc[2], wave, h, col);
    write_bias<3>(out, bacc[3], wave, h, col);
  } else {
    write_tile<4>(out, acc[3], 0, wave - 2, h, col);
    if (wave == 2) write_bias<4>(out, bacc[4], 0, h, col);
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    float t = 0.0f;
    for (int w = 0; w < kWaves; ++w) t += wave_loss[w];
    out[kParams] = t;
  }
}

// grad[i] += sum over blocks of partials[b][i]; *loss += sum of the losses.
// A workgroup owns 64 consecutive entries; its four waves each sum every
// fourth block row (64 consecutive floats per wave read), then the four
// partial sums are added in a fixed order.
__global__ __launch_bounds__(256) void reduce_kernel(
    const float* __restrict__ partials, int nblocks, float* __restrict__ grad,
    float* __restrict__ loss) {
  __shared__ float part[4][64];
  const int c = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int i = blockIdx.x * 64 + c;
  float t = 0.0f;
  if (i <= kParams)
    for (int b = w; b < nblocks; b += 4)
      t += partials[int64_t(b) * (kParams + 1) + i];
  part[w][c] = t;
  __syncthreads();
  if (w == 0 && i <= kParams) {
    const float s = ((part[0][c] + part[1][c]) + part[2][c]) + part[3][c];
    if (i < kParams) grad[i] += s;
    else if (loss) *loss += s;
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

size_t train_lds_bytes() { return kTrainLdsBytes; }

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
  DR_CHECK_ARG((reinterpret_cast<uintptr_t>(x) & 15) == 0,
               "features must be 16-byte aligned");
  const int64_t pairs = (num_pixels + 63) / 64;
  int64_t nb = num_blocks > 0 ? num_blocks : 2048;
  nb = std::min<int64_t>(nb, (pairs + dr::kWaves - 1) / dr::kWaves);
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
  const int64_t rounds = (num_pixels + dr::kRound - 1) / dr::kRound;
  const int nb = int(std::min<int64_t>(nb0, rounds));
  const auto* f = static_cast<const dr::bf16x8*>(packed);
  float* partials = static_cast<float*>(workspace);
  hipStream_t s = static_cast<hipStream_t>(stream);
  dr::train_kernel<<<unsigned(nb), dr::kBlock, lds, s>>>(
      f, f + dr::kFwdBlocks * 64, params, x, target, mask, num_pixels,
      loss_scale, partials, input_norm, flags);
  DR_CHECK_LAUNCH("mlp train step");
  dr::reduce_kernel<<<(dr::kParams + 1 + 63) / 64, 256, 0, s>>>(
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
