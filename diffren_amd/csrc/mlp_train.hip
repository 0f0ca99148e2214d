// Training step of the learned deferred shader on gfx950: forward, masked L2
// loss, backward and the weight and bias gradients in one kernel, then a
// fixed-order two-pass sum of the block partials into the gradient.  No
// reference counterpart: the reference shades with closed-form shaders only
// (diffren/jax/utils/shaders.py).
//
// Every wave works alone, with no barrier inside its loop, on chunks of 32
// pixels taken round-robin over the grid.  It runs the forward and the
// input-gradient chain (d in = W d out) of a chunk on the matrix cores as
// inference does (pixels on the N dimension), and keeps all sixteen 32x32
// weight-gradient tiles of the network in accumulator registers across its
// chunks: dW_l[in][out] = sum_pixels a_l[in] delta_l[out] is a product with
// the chunk's pixels as K, so it needs a_l and delta_l with units on lanes
// and pixels in registers, the transpose of the layout the chain produces.
//  * a_l is produced in the forward and used in the backward, so it goes
//    through a wave-private LDS stash: bfloat16 rows, one per pixel (32
//    bytes per lane and tile), read back transposed with gfx950's
//    ds_read_b64_tr_b16 (two per operand), and as rows for the ReLU mask.
//  * delta_l and the network input are transposed on the matrix cores, by
//    products with one-hot B operands (identity slices): a product whose A
//    operand is a packed accumulator tile puts the tile's pixels in
//    registers and its units on lanes.  LDS bandwidth (writes especially)
//    bounds this kernel, matrix time does not, so two more products per
//    tile are cheaper than a stash round trip.
// Stash positions are the accumulator order (position 32m + 16h + r of a
// row holds the unit of register r of tile m in lane half h), so the
// gradient tiles come out with their rows and columns in that order and
// the final sum maps them back to parameter indices.
//
// ReLU works on packed bfloat16 (a signed 16-bit max with 0 after rounding,
// which equals rounding the float32 ReLU), and the backward mask is the
// nonzero test of the stored activation (min(a, 1) as unsigned 16-bit times
// the packed delta): a pre-activation is positive exactly when its rounded
// ReLU output is nonzero, except below bfloat16's smallest subnormal.
// Layer 1's bias rides in the matrix product (input slots 12..14 = 1), which
// also yields its gradient as column 12 of dW1; the other layers' bias
// gradients are the row sums of the delta operands (bfloat16 dot products
// with ones).
#pragma clang fp contract(off)

#include <algorithm>

#include "dr_common.h"
#include "mlp_common.h"
#include "../../include/diffren_amd.h"

namespace dr {
namespace mlp {
namespace {

#define DR_LDS __attribute__((address_space(3)))

typedef short s16x8 __attribute__((ext_vector_type(8)));
typedef unsigned short u16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));

constexpr int kWaves = 4, kBlock = 64 * kWaves;
// Stash rows (bytes): one per pixel of the chunk, 64 bfloat16 and 16 bytes
// of padding (row reads and transposed reads then spread over the banks).
constexpr int kRowH = 2 * kHid + 16;
constexpr int kSlotH = 32 * kRowH;
constexpr int kWaveStash = 4 * kSlotH;
constexpr int kWeightBytes = kPackedBlocks * 64 * 16;
constexpr int kNormOff = kWeightBytes;            // float [24]
constexpr int kStashOff = kNormOff + 2 * kIn * 4;
constexpr int kLoopLds = kStashOff + kWaves * kWaveStash;
static_assert(kStashOff % 16 == 0 && kWaveStash % 16 == 0, "alignment");

// Block partial sums, raw order: 16 weight-gradient tiles
// [tile][r/4][lane][r%4], per-lane bias sums of layers 2..5
// [layer][tile][lane] (the two lane halves hold different pixels), per-lane
// losses.
constexpr int kTiles = 16;
constexpr int kRawBias = kTiles * 1024;
constexpr int kRawLoss = kRawBias + 4 * 2 * 64;
constexpr int kRawUsed = kRawLoss + 64;
constexpr int kRawStride = kRawUsed;      // 16-byte rows
static_assert(kRawStride % 4 == 0 && kRawLoss % 256 == 0 &&
                  kRawUsed == kRawLoss + 64,
              "sizes");
constexpr int kHalfTiles = kTiles / 2;
constexpr int kRawTail = kRawUsed - kRawBias;  // bias sums and losses
constexpr int kSumLds = kWaves * (kHalfTiles * 1024 + kRawTail) * 4;
constexpr int kLdsBytes = kLoopLds > kSumLds ? kLoopLds : kSumLds;
static_assert(kLdsBytes <= 160 * 1024, "LDS per workgroup");

// Tile order: W1 (out tile), W2..W4 (out tile, in tile), W5 (in tile).
constexpr int tile_index(int layer, int mo, int ni) {
  return layer == 0 ? mo : (layer == 4 ? 14 + ni : 2 + 4 * (layer - 1) + 2 * mo + ni);
}

__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ bf16x8 zero8() {
  bf16x8 z;
#pragma unroll
  for (int j = 0; j < 8; ++j) z[j] = static_cast<__bf16>(0.0f);
  return z;
}

typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ uint32_t cvt2(float a, float b) {
  const f32x2 v = {a, b};
  return __builtin_bit_cast(uint32_t, __builtin_convertvector(v, bf16x2));
}

// Registers 8s..8s+7 of an accumulator tile as bfloat16, one packed
// conversion per pair (converting the 8-vector at once, the compiler
// converts each value alone and merges the halves).
__device__ __forceinline__ bf16x8 pack8(const f32x16& c, int s) {
  const int o = 8 * s;
  const u32x4 u = {cvt2(c[o], c[o + 1]), cvt2(c[o + 2], c[o + 3]),
                   cvt2(c[o + 4], c[o + 5]), cvt2(c[o + 6], c[o + 7])};
  return __builtin_bit_cast(bf16x8, u);
}

// 1 in both 16-bit halves, as a value the compiler cannot see: with a
// literal 1 it rewrites d * min(a, 1) as two compares and selects per pair.
__device__ __forceinline__ u16x8 opaque_ones() {
  uint32_t k;
  __asm__ volatile("v_mov_b32 %0, 0x10001" : "=v"(k));
  typedef unsigned short u16x2 __attribute__((ext_vector_type(2)));
  const u16x2 k2 = __builtin_bit_cast(u16x2, k);
  return __builtin_shufflevector(k2, k2, 0, 1, 0, 1, 0, 1, 0, 1);
}

__device__ __forceinline__ bf16x8 relu8(bf16x8 v) {
  const s16x8 z = {0, 0, 0, 0, 0, 0, 0, 0};
  return __builtin_bit_cast(
      bf16x8, __builtin_elementwise_max(__builtin_bit_cast(s16x8, v), z));
}

// d where act != 0, else 0 (act >= +0, as relu8 leaves it; one = 1s).
__device__ __forceinline__ bf16x8 mask8(bf16x8 d, bf16x8 act, u16x8 one) {
  const u16x8 m = __builtin_elementwise_min(__builtin_bit_cast(u16x8, act), one);
  return __builtin_bit_cast(bf16x8, __builtin_bit_cast(u16x8, d) * m);
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

// Tile m of the lane's pixel row (16 values, 32 bytes at position
// 32m + 16h).
__device__ __forceinline__ void put_row(char* row, int m, int h, bf16x8 lo,
                                        bf16x8 hi) {
  bf16x8* p = reinterpret_cast<bf16x8*>(row + 64 * m + 32 * h);
  p[0] = lo;
  p[1] = hi;
}

__device__ __forceinline__ void get_row(const char* row, int m, int h,
                                        bf16x8& lo, bf16x8& hi) {
  const bf16x8* p = reinterpret_cast<const bf16x8*>(row + 64 * m + 32 * h);
  lo = p[0];
  hi = p[1];
}

// Transposed operand: lane l gets position (l%32) of a 32-position tile for
// eight pixels of a 16-pixel k-step s: slot j holds pixel
// 16s + 8(j/4) + 4(l/32) + j%4, the order in which a matrix-core transpose
// leaves pixels in registers (see transpose_tile).  ds_read_b64_tr_b16 gives
// lane j of each 16-lane group the 16-bit element j%4 of the 8 bytes
// addressed by lane 4e + j/4, for e = 0..3; lane 4e + q of group g
// addresses pixel row 4(g/2) + e (+8 for the second load), positions
// 16(g%2) + 4q..+3.  `p` is that address for the first load.
__device__ __forceinline__ bf16x8 tr_operand(const char* p, int row_bytes) {
  const bf16x4 lo = __builtin_amdgcn_ds_read_tr16_b64_v4bf16(
      (DR_LDS bf16x4*)(p));
  const bf16x4 hi = __builtin_amdgcn_ds_read_tr16_b64_v4bf16(
      (DR_LDS bf16x4*)(p + 8 * row_bytes));
  return __builtin_shufflevector(lo, hi, 0, 1, 2, 3, 4, 5, 6, 7);
}

__device__ __forceinline__ int tr_offset(int lane, int row_bytes) {
  const int g = lane >> 4, e = (lane >> 2) & 3, q = lane & 3;
  return (4 * (g >> 1) + e) * row_bytes + 32 * (g & 1) + 8 * q;
}

// One-hot B operands.  A product of a packed accumulator tile (A operand:
// lane = pixel, k-slots = units) with pos[s] picks its k-step-s half into
// output columns 16(l/32) + 8s + j, i.e. column n = position n; the two
// products of a tile give its transpose: lane n holds position n, register
// r pixel (r/4)8 + 4(l/32) + r%4, and pack8(result, s) is the k-step-s
// operand of a weight-gradient product.  Products with one nonzero term are
// exact.  The input operand (k-slot 8h + j = feature 8h + j) times pos[0]
// puts feature f in column f for f < 8 and in column f + 8 for f = 8..15.
struct Ident {
  bf16x8 pos[2];
};

__device__ __forceinline__ Ident identities(int lane) {
  const int n = lane & 31, hb = lane >> 5;
  Ident I;
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    I.pos[0][j] = static_cast<__bf16>(n == 16 * hb + j ? 1.0f : 0.0f);
    I.pos[1][j] = static_cast<__bf16>(n == 16 * hb + 8 + j ? 1.0f : 0.0f);
  }
  return I;
}

// One chunk's global inputs: features 8h..8h+3 (xa), 4..7 (xb) and 8..11
// (xc, for the albedo); target and mask.  Every lane loads, from pixel
// min(p, N - 1), so that no load sits under a branch: the compiler cannot
// count loads issued under one, and would wait for the next chunk's loads
// as soon as this chunk's values are used.  Pixels past N get a zero delta,
// so their (finite) copied inputs do not reach any sum.
struct Fetch {
  float4 xa, xb, xc;
  float t0, t1, t2, mk;
};

__device__ __forceinline__ Fetch fetch(const float* __restrict__ x,
                                       const float* __restrict__ target,
                                       const float* __restrict__ mask,
                                       int64_t p, int64_t N, int h) {
  const int64_t q = p < N ? p : N - 1;
  Fetch f;
  f.xa = ldg4(x + q * kIn + 8 * h);
  f.xb = ldg4(x + q * kIn + 4);
  f.xc = ldg4(x + q * kIn + 8);
  f.t0 = ldg(target + q * 3);
  f.t1 = ldg(target + q * 3 + 1);
  f.t2 = ldg(target + q * 3 + 2);
  // The raw value (the target when there is no mask): selecting 1 here
  // would make the compiler wait for this load as soon as it is issued.
  f.mk = ldg((mask ? mask : target) + q);
  return f;
}

// gacc += a b on the matrix cores with the accumulator in AGPRs.  The
// builtin would keep matrix-core results in VGPRs (this file is built like
// the rest of the library, for the chain's sake), and 256 accumulator
// registers do not fit next to the chain there.  The compiler does not see
// inside the statement, so it leaves out the wait states it would insert
// for a matrix instruction: the operands come from LDS loads (ordered by
// the counter waits it does insert), the leading s_nop covers a VALU write
// of an operand just before, and the accumulators are read back only after
// the loop, behind an explicit drain at the end of each chunk.
__device__ __forceinline__ void dw_mfma(f32x16& acc, bf16x8 a, bf16x8 b) {
  __asm__ volatile("s_nop 1\n\tv_mfma_f32_32x32x16_bf16 %0, %1, %2, %0"
                   : "+a"(acc)
                   : "v"(a), "v"(b));
}

// Every stage of a chunk issues, before any of its arithmetic and behind a
// scheduling fence, the LDS loads of the next stage's weights and of its own
// transposed operands and stored activations: left alone, the compiler
// places each load next to its use and the wave, alone on its SIMD, waits
// out the LDS latency before every product.
__device__ __forceinline__ void sched_fence() {
  __builtin_amdgcn_sched_barrier(0);
}

// B operand of ones in k-slots 0..2: a product with a bias block starts an
// accumulator at the float32 bias.
__device__ __forceinline__ bf16x8 ones_operand(int h) {
  bf16x8 v = zero8();
  if (h == 0) {
    v[0] = static_cast<__bf16>(1.0f);
    v[1] = static_cast<__bf16>(1.0f);
    v[2] = static_cast<__bf16>(1.0f);
  }
  return v;
}

// Weights of forward layer L (1..4, W_{L+1}): bias block and four k-step
// blocks per out tile (two out tiles, one for L = 4).
struct FwdW {
  bf16x8 bias[2], w[8];
};

template <int L>
__device__ __forceinline__ FwdW load_fwd(const bf16x8* __restrict__ sw,
                                         int lane) {
  constexpr int M = L < 4 ? 2 : 1;
  FwdW f;
#pragma unroll
  for (int m = 0; m < M; ++m) {
    f.bias[m] = sw[(kBiasBase + 2 * (L - 1) + m) * 64 + lane];
#pragma unroll
    for (int s = 0; s < 4; ++s)
      f.w[4 * m + s] = sw[(kFwdBase[L] + 4 * m + s) * 64 + lane];
  }
  return f;
}

// Weights of d in = W d out through layer L (1..4): k-step blocks of two in
// tiles (one k-step for L = 4).
struct BwdW {
  bf16x8 w[8];
};

template <int L>
__device__ __forceinline__ BwdW load_bwd(const bf16x8* __restrict__ wb,
                                         int lane) {
  constexpr int KS = L < 4 ? 4 : 1;
  BwdW b;
#pragma unroll
  for (int m = 0; m < 2; ++m)
#pragma unroll
    for (int s = 0; s < KS; ++s)
      b.w[KS * m + s] = wb[(kBwdBase[L] + KS * m + s) * 64 + lane];
  return b;
}

// The arithmetic of a stage is laid out as one product at a time, each
// followed by the vector work that does not depend on it, between
// scheduling fences: a wave issues in order and a product holds the matrix
// pipeline for 32 cycles, so vector work placed after a product runs in its
// shadow, while two products back to back stall the wave for the second.

// Forward through layer L (1..4, W_{L+1}): `a` holds the pre-activation of
// its input a_L, which goes (after ReLU) to the stash and the B operands.
template <int L>
__device__ __forceinline__ void forward_layer(const FwdW& W, int h, char* row,
                                              bf16x8 ones, f32x16 a[2]) {
  bf16x8 nb[4];
  f32x16 c[2];
  if (L == 4) {
    // One out tile: two independent accumulations summed at the end (a
    // product that depends on the previous one waits for it to drain).
    c[0] = mfma(W.bias[0], ones, zero16());
    nb[0] = relu8(pack8(a[0], 0));
    sched_fence();
    nb[1] = relu8(pack8(a[0], 1));
    put_row(row, 0, h, nb[0], nb[1]);
    c[1] = mfma(W.w[0], nb[0], zero16());
    sched_fence();
    c[0] = mfma(W.w[1], nb[1], c[0]);
    nb[2] = relu8(pack8(a[1], 0));
    sched_fence();
    nb[3] = relu8(pack8(a[1], 1));
    put_row(row, 1, h, nb[2], nb[3]);
    c[1] = mfma(W.w[2], nb[2], c[1]);
    sched_fence();
    c[0] = mfma(W.w[3], nb[3], c[0]);
    sched_fence();
    a[0] = c[0] + c[1];
    return;
  }
  c[0] = mfma(W.bias[0], ones, zero16());
  nb[0] = relu8(pack8(a[0], 0));
  sched_fence();
  c[1] = mfma(W.bias[1], ones, zero16());
  nb[1] = relu8(pack8(a[0], 1));
  put_row(row, 0, h, nb[0], nb[1]);
  sched_fence();
  c[0] = mfma(W.w[0], nb[0], c[0]);
  nb[2] = relu8(pack8(a[1], 0));
  sched_fence();
  c[1] = mfma(W.w[4], nb[0], c[1]);
  nb[3] = relu8(pack8(a[1], 1));
  put_row(row, 1, h, nb[2], nb[3]);
  sched_fence();
#pragma unroll
  for (int s = 1; s < 4; ++s) {
#pragma unroll
    for (int m = 0; m < 2; ++m) {
      c[m] = mfma(W.w[4 * m + s], nb[s], c[m]);
      sched_fence();
    }
  }
  a[0] = c[0];
  a[1] = c[1];
}

// Backward through layer L (0-based weight index, 1..4): the delta of its
// output is in dd (packed, B-operand order; MO out tiles, only dd[0] for
// the 3-unit output layer); its input a_L is in `aslot`.  Adds dW_L (and
// the delta row sums to bs) and leaves the delta of a_L in dd.  After its
// own loads it issues the next stage's weights: those of layer L - 1 into
// `next`, or for L = 1 the next chunk's layer-1 blocks into `w1`.
template <int L, int MO>
__device__ __forceinline__ void backward_layer(
    const BwdW& W, const bf16x8* __restrict__ sw, int lane, int h,
    const char* aslot, int row_off, int trh, const Ident& I, u16x8 one,
    bf16x8 dd[4], f32x16* gacc, float* bs, BwdW* next, bf16x8* w1) {
  constexpr int KS = L < 4 ? 4 : 1;
  bf16x8 B[2][2], act[4];
#pragma unroll
  for (int s = 0; s < 2; ++s)
#pragma unroll
    for (int ni = 0; ni < 2; ++ni)
      B[s][ni] = tr_operand(aslot + trh + 16 * s * kRowH + 64 * ni, kRowH);
#pragma unroll
  for (int m = 0; m < 2; ++m)
    get_row(aslot + row_off, m, h, act[2 * m], act[2 * m + 1]);
  if (L > 1) {
    *next = load_bwd<L == 1 ? 1 : L - 1>(sw + kFwdBlocks * 64, lane);
  } else {
#pragma unroll
    for (int m = 0; m < 2; ++m) w1[m] = sw[(kFwdBase[0] + m) * 64 + lane];
  }
  sched_fence();
  // Transposes of the output delta (the second half is zero for L = 4).
  f32x16 T[MO];
#pragma unroll
  for (int mo = 0; mo < MO; ++mo) {
    T[mo] = mfma(dd[2 * mo], I.pos[0], zero16());
    sched_fence();
  }
  if (L < 4) {
#pragma unroll
    for (int mo = 0; mo < MO; ++mo) {
      T[mo] = mfma(dd[2 * mo + 1], I.pos[1], T[mo]);
      sched_fence();
    }
  }
  // d in = W d out, the two in tiles alternating; in their shadow the
  // delta operands (bfloat16 of the transposes) and their row sums.
  bf16x8 A[2][MO];
  f32x16 dn[2];
  int q = 0;
#pragma unroll
  for (int s = 0; s < KS; ++s) {
#pragma unroll
    for (int m = 0; m < 2; ++m) {
      dn[m] = mfma(W.w[KS * m + s], dd[s], s == 0 ? zero16() : dn[m]);
      if ((KS == 1 || s * 2 + m >= 2) && q < 2 * MO) {
        A[q / MO][q % MO] = pack8(T[q % MO], q / MO);
        bs[q % MO] = row_sum(A[q / MO][q % MO], bs[q % MO]);
        ++q;
      }
      sched_fence();
    }
  }
#pragma unroll
  for (; q < 2 * MO; ++q) {
    A[q / MO][q % MO] = pack8(T[q % MO], q / MO);
    bs[q % MO] = row_sum(A[q / MO][q % MO], bs[q % MO]);
  }
  // dW, each product followed by a quarter of the ReLU mask of d in.
  int i = 0, piece = 0;
#pragma unroll
  for (int s = 0; s < 2; ++s)
#pragma unroll
    for (int ni = 0; ni < 2; ++ni)
#pragma unroll
      for (int mo = 0; mo < MO; ++mo, ++i) {
        dw_mfma(gacc[tile_index(L, mo, ni)], A[s][mo], B[s][ni]);
        if (i > 0 && piece < 4) {
          const int m = piece >> 1, k = piece & 1;
          dd[2 * m + k] = mask8(pack8(dn[m], k), act[2 * m + k], one);
          ++piece;
        }
        sched_fence();
      }
#pragma unroll
  for (; piece < 4; ++piece) {
    const int m = piece >> 1, k = piece & 1;
    dd[2 * m + k] = mask8(pack8(dn[m], k), act[2 * m + k], one);
  }
}

__global__ __launch_bounds__(kBlock) void train_kernel(
    const bf16x8* __restrict__ packed, const float* __restrict__ x,
    const float* __restrict__ target, const float* __restrict__ mask,
    int64_t N, float scale, float* __restrict__ partials,
    const float* __restrict__ norm, int flags) {
  extern __shared__ __attribute__((aligned(16))) char smem[];
  bf16x8* sw = reinterpret_cast<bf16x8*>(smem);
  const bf16x8* wb = sw + kFwdBlocks * 64;
  float* nrm = reinterpret_cast<float*>(smem + kNormOff);
  {
    // All loads in flight before the first store (a load-store loop waits
    // out one L2 round trip per iteration).
    constexpr int kPer = (kPackedBlocks * 64 + kBlock - 1) / kBlock;
    bf16x8 v[kPer];
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
      const int i = threadIdx.x + k * kBlock;
      if (i < kPackedBlocks * 64) v[k] = ldg(packed + i);
    }
    float nv = 0.0f;
    if (threadIdx.x < 2 * kIn)
      nv = norm ? ldg(norm + threadIdx.x) : (threadIdx.x < kIn ? 0.0f : 1.0f);
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
      const int i = threadIdx.x + k * kBlock;
      if (i < kPackedBlocks * 64) sw[i] = v[k];
    }
    if (threadIdx.x < 2 * kIn) nrm[threadIdx.x] = nv;
  }

  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int h = lane >> 5, col = lane & 31;
  char* st = smem + kStashOff + wave * kWaveStash;
  // S[1..4]: the layer inputs a1..a4.
  char* S[5] = {nullptr, st, st + kSlotH, st + 2 * kSlotH, st + 3 * kSlotH};
  const int row_h = col * kRowH;
  __syncthreads();

  const bool modulate = (flags & DR_MLP_ALBEDO_MODULATE) != 0;
  const int trh = tr_offset(lane, kRowH);
  const Ident I = identities(lane);
  const u16x8 one = opaque_ones();
  const bf16x8 ones = ones_operand(h);
  // (x - shift) * scale for k-slots 8h..8h+7 (features past 11: 0 * 0).
  float shift[8], scl[8];
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    const int f = 8 * h + j;
    shift[j] = f < kIn ? nrm[f] : 0.0f;
    scl[j] = f < kIn ? nrm[kIn + f] : 0.0f;
  }

  f32x16 gacc[kTiles];
#pragma unroll
  for (int t = 0; t < kTiles; ++t) gacc[t] = zero16();
  float bs[4][2] = {{0.0f, 0.0f}, {0.0f, 0.0f}, {0.0f, 0.0f}, {0.0f, 0.0f}};
  float partial = 0.0f;

  // Chunk indices fit 32 bits (the host limits N to 2^35 pixels).
  const int chunks = int((N + 31) / 32);
  const int stride = int(gridDim.x) * kWaves;
  int c = int(blockIdx.x) * kWaves + wave;
  Fetch next = fetch(x, target, mask, int64_t(c) * 32 + col, N, h);
  bf16x8 w1[2];
#pragma unroll
  for (int m = 0; m < 2; ++m) w1[m] = sw[(kFwdBase[0] + m) * 64 + lane];
  for (; c < chunks; c += stride) {
    const int64_t p = int64_t(c) * 32 + col;
    const Fetch cur = next;
    next = fetch(x, target, mask, int64_t(c + stride) * 32 + col, N, h);
    // The previous chunk's stash reads come before this chunk's writes; the
    // empty statement keeps the weight reads from LDS inside the loop
    // (hoisted, they would pin 250 registers).
    wave_sync();
    __asm__ volatile("" ::: "memory");
    // ---- layer 1 (bias in the product); its input to the stash
    FwdW wn = load_fwd<1>(sw, lane);
    sched_fence();
    f32x16 a[2];
    bf16x8 b0;
    {
      const float raw[8] = {cur.xa.x, cur.xa.y, cur.xa.z, cur.xa.w,
                            cur.xb.x, cur.xb.y, cur.xb.z, cur.xb.w};
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        // Lane half 1: features 8..11, then the bias slots 12..14 = 1.
        float v = (raw[j] - shift[j]) * scl[j];
        if (j >= 4) v = h ? (j < 7 ? 1.0f : 0.0f) : v;
        b0[j] = static_cast<__bf16>(v);
      }
#pragma unroll
      for (int m = 0; m < 2; ++m) a[m] = mfma(w1[m], b0, zero16());
    }
    // ---- layers 2..5, each stage loading the next stage's weights
    FwdW wc = wn;
    wn = load_fwd<2>(sw, lane);
    sched_fence();
    forward_layer<1>(wc, h, S[1] + row_h, ones, a);
    wc = wn;
    wn = load_fwd<3>(sw, lane);
    sched_fence();
    forward_layer<2>(wc, h, S[2] + row_h, ones, a);
    wc = wn;
    wn = load_fwd<4>(sw, lane);
    sched_fence();
    forward_layer<3>(wc, h, S[3] + row_h, ones, a);
    wc = wn;
    BwdW bn = load_bwd<4>(wb, lane);
    sched_fence();
    forward_layer<4>(wc, h, S[4] + row_h, ones, a);
    const f32x16& y = a[0];
    // ---- loss and output delta (units 0..2: lane half 0, registers 0..2)
    f32x16 d5 = zero16();
    if (h == 0 && p < N) {
      const float alb[3] = {modulate ? cur.xc.y : 1.0f,
                            modulate ? cur.xc.z : 1.0f,
                            modulate ? cur.xc.w : 1.0f};
      const float t[3] = {cur.t0, cur.t1, cur.t2};
      const float mk = mask ? cur.mk : 1.0f;
#pragma unroll
      for (int r = 0; r < 3; ++r) {
        const float e = alb[r] * y[r] - t[r];
        partial += mk * (e * e) * scale;
        d5[r] = (((2.0f * e) * scale) * mk) * alb[r];
      }
    }
    bf16x8 dd[4];
    dd[0] = pack8(d5, 0);
    dd[1] = dd[2] = dd[3] = zero8();
    // The stash rows of a1..a4 are complete before other lanes read them.
    wave_sync();
    // ---- backward, each stage loading the next stage's weights
    BwdW bc = bn;
    backward_layer<4, 1>(bc, sw, lane, h, S[4], row_h, trh, I, one, dd, gacc,
                         bs[3], &bn, w1);
    bc = bn;
    backward_layer<3, 2>(bc, sw, lane, h, S[3], row_h, trh, I, one, dd, gacc,
                         bs[2], &bn, w1);
    bc = bn;
    backward_layer<2, 2>(bc, sw, lane, h, S[2], row_h, trh, I, one, dd, gacc,
                         bs[1], &bn, w1);
    bc = bn;
    backward_layer<1, 2>(bc, sw, lane, h, S[1], row_h, trh, I, one, dd, gacc,
                         bs[0], &bn, w1);
    // ---- dW1 (and b1's gradient, input slot 12): transposes of the delta
    // of a1 and of the input operand
    {
      f32x16 T[2], Ta;
      T[0] = mfma(dd[0], I.pos[0], zero16());
      sched_fence();
      T[1] = mfma(dd[2], I.pos[0], zero16());
      sched_fence();
      Ta = mfma(b0, I.pos[0], zero16());
      sched_fence();
      T[0] = mfma(dd[1], I.pos[1], T[0]);
      sched_fence();
      T[1] = mfma(dd[3], I.pos[1], T[1]);
      sched_fence();
#pragma unroll
      for (int s2 = 0; s2 < 2; ++s2) {
        const bf16x8 B = pack8(Ta, s2);
#pragma unroll
        for (int mo = 0; mo < 2; ++mo) {
          dw_mfma(gacc[tile_index(0, mo, 0)], pack8(T[mo], s2), B);
          sched_fence();
        }
      }
    }
    // The chunk's last accumulation leaves the matrix pipeline (an 8-pass
    // result needs up to 19 wait states before a VALU read) before any
    // code after the loop, in another basic block, reads the accumulators.
    __asm__ volatile("s_nop 7\n\ts_nop 7\n\ts_nop 7");
  }

  // ---- block sum, in two halves of eight tiles: every wave stores its
  // tiles in a region of its own, then all threads add the four regions in
  // wave order and write the block row.  Each accumulator is read once, in
  // straight-line code (behind a branch the compiler reads them all out at
  // the loop exit and spills).
  __syncthreads();
  float4* R4 = reinterpret_cast<float4*>(smem);
  float* Rb = reinterpret_cast<float*>(R4 + kWaves * kHalfTiles * 256);
  float4* out4 = reinterpret_cast<float4*>(partials + int64_t(blockIdx.x) * kRawStride);
#pragma unroll
  for (int half = 0; half < 2; ++half) {
#pragma unroll
    for (int t = 0; t < kHalfTiles; ++t) {
      const f32x16& g16 = gacc[kHalfTiles * half + t];
#pragma unroll
      for (int g = 0; g < 4; ++g)
        R4[(wave * kHalfTiles + t) * 256 + g * 64 + lane] =
            make_float4(g16[4 * g], g16[4 * g + 1], g16[4 * g + 2], g16[4 * g + 3]);
      sched_fence();
    }
    if (half == 1) {
#pragma unroll
      for (int l = 0; l < 4; ++l)
#pragma unroll
        for (int mo = 0; mo < 2; ++mo)
          Rb[wave * kRawTail + (2 * l + mo) * 64 + lane] =
              (l == 3 && mo == 1) ? 0.0f : bs[l][mo];
      Rb[wave * kRawTail + 512 + lane] = partial;
    }
    __syncthreads();
    for (int i = threadIdx.x; i < kHalfTiles * 256; i += kBlock) {
      float4 v = R4[i];
#pragma unroll
      for (int w = 1; w < kWaves; ++w) {
        const float4 o = R4[w * kHalfTiles * 256 + i];
        v.x += o.x;
        v.y += o.y;
        v.z += o.z;
        v.w += o.w;
      }
      // Streamed past the caches: the row is read back by another kernel.
      const f32x4 w = {v.x, v.y, v.z, v.w};
      __builtin_nontemporal_store(
          w, reinterpret_cast<f32x4*>(out4 + half * kHalfTiles * 256 + i));
    }
    __syncthreads();
  }
  float* out = partials + int64_t(blockIdx.x) * kRawStride;
  for (int i = threadIdx.x; i < kRawTail; i += kBlock) {
    float v = Rb[i];
#pragma unroll
    for (int w = 1; w < kWaves; ++w) v += Rb[w * kRawTail + i];
    out[kRawBias + i] = v;
  }
}

// Position P of a 64-position stash row -> unit.
__device__ __forceinline__ int unit_of(int P) {
  return c_unit(P >> 5, P & 15, (P >> 4) & 1);
}

// Weight-gradient raw index -> parameter index (-1: padding).
__device__ int tile_param(int e) {
  const int t = e >> 10, rem = e & 1023;
  const int l = (rem >> 2) & 63, r = 4 * (rem >> 8) + (rem & 3);
  const int ci = l & 31, ri = (r >> 2) * 8 + 4 * (l >> 5) + (r & 3);
  int layer, mo, ni;
  if (t < 2) {
    layer = 0; mo = t; ni = 0;
  } else if (t < 14) {
    layer = 1 + (t - 2) / 4; mo = ((t - 2) & 3) >> 1; ni = (t - 2) & 1;
  } else {
    layer = 4; mo = 0; ni = t - 14;
  }
  const int op = 32 * mo + ri, ip = 32 * ni + ci;
  if (layer == 0) {
    // Input columns: features 0..7 at 0..7, features 8..15 at 16..23.
    const int f = ip < 8 ? ip : (ip >= 16 && ip < 24 ? ip - 8 : -1);
    if (f >= 0 && f < kIn) return kW[0] + f * kHid + unit_of(op);
    return f == kBiasSlot ? kB[0] + unit_of(op) : -1;
  }
  if (layer == 4) return op < kOut ? kW[4] + unit_of(ip) * kOut + op : -1;
  return kW[layer] + unit_of(ip) * kHid + unit_of(op);
}


// Sum of the block rows of 64 raw columns per workgroup, one pass: 16
// thread groups each add every 16th row, then group 0 adds the 16 group sums
// in a fixed order and scatters the result to the parameters (a bias
// position's two lane halves, and the 64 per-lane losses, fall in the same
// 64 columns).  overwrite: grad = sum (else grad += sum), and loss likewise.
constexpr int kSumCols = 64, kSumGroups = 16;
static_assert(kRawBias % kSumCols == 0 && kRawLoss % kSumCols == 0 &&
                  kRawUsed % kSumCols == 0,
              "raw sections align with the column blocks");

__global__ __launch_bounds__(kSumCols * kSumGroups) void sum_kernel(
    const float* __restrict__ partials, int nblocks, float* __restrict__ grad,
    float* __restrict__ loss, int overwrite) {
  __shared__ float red[kSumGroups][kSumCols];
  const int c = threadIdx.x % kSumCols, grp = threadIdx.x / kSumCols;
  const int e = blockIdx.x * kSumCols + c;
  float t = 0.0f;
#pragma unroll 8
  for (int b = grp; b < nblocks; b += kSumGroups)
    t += partials[int64_t(b) * kRawStride + e];
  red[grp][c] = t;
  __syncthreads();
  if (grp != 0) return;
  t = red[0][c];
#pragma unroll
  for (int g = 1; g < kSumGroups; ++g) t += red[g][c];
  auto put = [&](float* dst, float v) { *dst = overwrite ? v : *dst + v; };
  if (e < kRawBias) {
    const int i = tile_param(e);
    if (i >= 0) put(grad + i, t);
  } else if (e < kRawLoss) {
    // Bias sums: lanes l and l + 32 hold the same position.
    red[0][c] = t;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    if (c >= 32) return;
    const int b = e - kRawBias;
    const int layer = 1 + (b >> 7), pos = 32 * ((b >> 6) & 1) + c;
    const float v = red[0][c] + red[0][c + 32];
    if (layer == 4) {
      if (pos < kOut) put(grad + kB[4] + pos, v);
    } else {
      put(grad + kB[layer] + unit_of(pos), v);
    }
  } else {
    // The 64 per-lane losses.
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) t += __shfl_xor(t, off);
    if (c == 0 && loss) put(loss, t);
  }
}

}  // namespace
}  // namespace mlp
}  // namespace dr

extern "C" {

size_t dr_mlp_train_workspace_size(int32_t num_blocks) {
  const int nb = num_blocks > 0 ? num_blocks : 256;
  return size_t(nb) * dr::mlp::kRawStride * sizeof(float);
}

int dr_mlp_train_step(const void* packed, const float* params, const float* x,
                      const float* target, const float* mask,
                      int64_t num_pixels, float loss_scale,
                      const float* input_norm, int32_t flags, float* grad,
                      float* loss, int32_t num_blocks, void* workspace,
                      size_t workspace_bytes, dr_stream_t stream) {
  namespace M = dr::mlp;
  DR_CHECK_ARG(num_pixels >= 0, "negative size");
  DR_CHECK_ARG(grad != nullptr, "null pointer");
  const int overwrite = (flags & DR_MLP_GRAD_OVERWRITE) ? 1 : 0;
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (num_pixels == 0) {
    if (overwrite) {  // a sum over no block rows: zeros
      M::sum_kernel<<<M::kRawUsed / M::kSumCols, M::kSumCols * M::kSumGroups,
                      0, s>>>(nullptr, 0, grad, loss, 1);
      DR_CHECK_LAUNCH("mlp train sum");
    }
    return DR_OK;
  }
  DR_CHECK_ARG(num_pixels <= (int64_t(1) << 35), "more than 2^35 pixels");
  DR_CHECK_ARG(packed && params && x && target, "null pointer");
  DR_CHECK_ARG(workspace != nullptr, "workspace is null");
  DR_CHECK_ARG((reinterpret_cast<uintptr_t>(x) & 15) == 0,
               "features must be 16-byte aligned");
  DR_CHECK_ARG((reinterpret_cast<uintptr_t>(workspace) & 15) == 0,
               "workspace must be 16-byte aligned");
  const int nb0 = num_blocks > 0 ? num_blocks : 256;
  if (workspace_bytes < dr_mlp_train_workspace_size(nb0))
    return dr::fail(dr::kWorkspaceTooSmall, "mlp train workspace too small");
  static bool lds_set = false;  // one attribute call per process
  if (!lds_set) {
    DR_CHECK_HIP(hipFuncSetAttribute(
        reinterpret_cast<const void*>(&M::train_kernel),
        hipFuncAttributeMaxDynamicSharedMemorySize, M::kLdsBytes));
    lds_set = true;
  }
  const int64_t chunks = (num_pixels + 31) / 32;
  const int nb = int(std::min<int64_t>(nb0, (chunks + M::kWaves - 1) / M::kWaves));
  float* partials = static_cast<float*>(workspace);
  M::train_kernel<<<unsigned(nb), M::kBlock, M::kLdsBytes, s>>>(
      static_cast<const M::bf16x8*>(packed), x, target, mask, num_pixels,
      loss_scale, partials, input_norm, flags);
  DR_CHECK_LAUNCH("mlp train step");
  M::sum_kernel<<<M::kRawUsed / M::kSumCols, M::kSumCols * M::kSumGroups, 0,
                  s>>>(partials, nb, grad, loss, overwrite);
  DR_CHECK_LAUNCH("mlp train sum");
  return DR_OK;
}

}  // extern "C"
