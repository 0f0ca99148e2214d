// Rasterize-then-splat accumulation for gfx950, forward and backward.
//
// Forward restates diffren/jax/internal/splat.py:26-175
// (splat_at_pixel_centers) in "pull" form: one thread per output pixel T
// gathers the 3x3 neighbourhood of source pixels S = T - (u, v) over all
// layers, so no atomics are needed and every output is written once.
//   weight  w = (mask * coef) * exp(-2 * |floor(xy) + 0.5 - xy + (u, v)|^2)
//           coef = 1.05 / sum_{u,v} exp(-2 (u^2 + v^2))            (:63-78)
//   surface index of (S, T): argmin_l |z_l(S) - z_0(T)| versus
//           argmin_l |z_l(T) - z_0(S)| (first index on ties)       (:131-145)
//   class   fg / surface / bg = surface > l / == l / < l            (:151-154)
//   acc[class] += (w * color, w);  n = acc / (max(acc_w - 1, 0) + 1) (:168-170)
//   out = over(n_fg, n_surf, n_bg)                 (composite.py:20-37)
// The loop order per target (layer, then u, then v) is the reference's.
//
// Backward (no reference counterpart: the reference differentiates this with
// XLA autodiff) runs in two passes:
//   1. per target T: recompute acc, back-propagate d_out through `over` and
//      the normalisation into G[T, class] = (d acc_color, d acc_weight);
//   2. per source S and layer l: pull G from the 9 targets and form
//      d color_l(S) = sum w * G_c  and
//      d xy_l(S)    = sum (G_c . color + G_w) * 4 w (floor(xy) + 0.5 - xy + (u,v)).
// z only selects layers (argmin) and gets no gradient, as in the reference.
#pragma clang fp contract(off)

#include <cmath>

#include "dr_common.h"
#include "../../include/diffren_amd.h"

namespace dr {
namespace {

constexpr int kBX = 64, kBY = 4;
// Frames of up to kMaxLayers layers keep a pixel's layer depths in
// registers; deeper frames (any L, as splat.py:26-175 takes) run a variant
// that reads them from memory where the surface index needs them.
constexpr int kMaxLayers = 8;

struct SplatFrame {
  const float* xyz;    // [L, H, W, 3] viewport x, y (pixels) and z in [0, 1]
  const float* color;  // [L, H, W, C]
  const float* mask;   // [L, H, W]
  int L, H, W;
  float coef;
};

// Surface index of the (source S, target T) pair; splat.py:131-145.  zS(l)
// and zT(l) give layer l's z of the source and the target pixel.
template <typename ZS, typename ZT>
__device__ __forceinline__ int surface_index(const ZS& zS, const ZT& zT,
                                             int L) {
  const float s0 = zS(0), t0 = zT(0);
  float bc = fabsf(s0 - t0);
  int ic = 0;
  float bt = fabsf(t0 - s0);
  int it = 0;
  for (int l = 1; l < L; ++l) {
    const float dc = fabsf(zS(l) - t0);
    if (dc < bc) { bc = dc; ic = l; }
    const float dt = fabsf(zT(l) - s0);
    if (dt < bt) { bt = dt; it = l; }
  }
  return bt < bc ? -it : ic;
}

// Layer l's z at pixel p, from memory.
__device__ __forceinline__ float frame_z(const SplatFrame& f, int l,
                                         int64_t p) {
  return f.xyz[3 * (l * int64_t(f.H) * f.W + p) + 2];
}

// Surface indices of pixel (x, y) with its 9 neighbours, tap order
// t = 3 (u + 1) + (v + 1).  As a target (kTarget) the neighbour at
// (x - u, y - v) is the source; otherwise (x, y) is the source and
// (x + u, y + v) the target.  Out-of-image neighbours get 0 (never read).
template <bool kAnyL, bool kTarget>
__device__ __forceinline__ void surface_indices(const SplatFrame& f, int x,
                                                int y, int surf[9]) {
  const int64_t p = int64_t(y) * f.W + x;
  float zp[kAnyL ? 1 : kMaxLayers], zn[kAnyL ? 1 : kMaxLayers];
  if constexpr (!kAnyL)
    for (int l = 0; l < f.L; ++l) zp[l] = frame_z(f, l, p);
#pragma unroll
  for (int t = 0; t < 9; ++t) {
    const int u = t / 3 - 1, v = t % 3 - 1;
    const int nx = kTarget ? x - u : x + u, ny = kTarget ? y - v : y + v;
    surf[t] = 0;
    if (nx < 0 || nx >= f.W || ny < 0 || ny >= f.H) continue;
    const int64_t n = int64_t(ny) * f.W + nx;
    if constexpr (kAnyL) {
      auto here = [&](int l) { return frame_z(f, l, p); };
      auto there = [&](int l) { return frame_z(f, l, n); };
      surf[t] = kTarget ? surface_index(there, here, f.L)
                        : surface_index(here, there, f.L);
    } else {
      for (int l = 0; l < f.L; ++l) zn[l] = frame_z(f, l, n);
      auto here = [&](int l) { return zp[l]; };
      auto there = [&](int l) { return zn[l]; };
      surf[t] = kTarget ? surface_index(there, here, f.L)
                        : surface_index(here, there, f.L);
    }
  }
}

__device__ __forceinline__ int layer_class(int surface, int l) {
  return surface > l ? 0 : (surface == l ? 1 : 2);
}

// Tap weight and offset d = floor(xy) + 0.5 - xy + (u, v).
__device__ __forceinline__ float tap_weight(float m, float coef, float sx,
                                            float sy, int u, int v, float& dx,
                                            float& dy) {
  dx = ((floorf(sx) + 0.5f) - sx) + float(u);
  dy = ((floorf(sy) + 0.5f) - sy) + float(v);
  const float d2 = dx * dx + dy * dy;
  return (m * coef) * expf(-2.0f * d2);
}

// acc[k][0..C) colour, acc[k][C] weight, for target (x, y).
template <int C, bool kAnyL>
__device__ __forceinline__ void accumulate_target(const SplatFrame& f, int x,
                                                  int y,
                                                  float acc[3][C + 1]) {
#pragma unroll
  for (int k = 0; k < 3; ++k)
#pragma unroll
    for (int c = 0; c <= C; ++c) acc[k][c] = 0.0f;
  const int64_t P = int64_t(f.H) * f.W;
  int surf[9];
  surface_indices<kAnyL, true>(f, x, y, surf);
  for (int l = 0; l < f.L; ++l) {
#pragma unroll
    for (int t = 0; t < 9; ++t) {
      const int u = t / 3 - 1, v = t % 3 - 1;
      const int sx = x - u, sy = y - v;
      if (sx < 0 || sx >= f.W || sy < 0 || sy >= f.H) continue;
      const int64_t q = l * P + int64_t(sy) * f.W + sx;
      const float m = f.mask[q];
      if (m == 0.0f) continue;
      float dx, dy;
      const float w =
          tap_weight(m, f.coef, f.xyz[3 * q], f.xyz[3 * q + 1], u, v, dx, dy);
      const int k = layer_class(surf[t], l);
      const float* col = f.color + q * C;
      // The reference's masked accumulation (splat.py:147-160): each class
      // adds the tap times its 0/1 mask, here fma(tap, mask, acc) -- one
      // rounding of acc + tap where the mask is 1 (tap * 1 is exact), acc
      // plus a zero where it is 0 (no bit changes: a sum that starts at +0
      // is never -0), and a non-finite tap reaches every class as tap * 0
      // does there.  Branch-free, so acc stays in registers.
      const float m0 = k == 0 ? 1.0f : 0.0f;
      const float m1 = k == 1 ? 1.0f : 0.0f;
      const float m2 = k == 2 ? 1.0f : 0.0f;
#pragma unroll
      for (int c = 0; c < C; ++c) {
        const float wc = w * col[c];
        acc[0][c] = fmaf(wc, m0, acc[0][c]);
        acc[1][c] = fmaf(wc, m1, acc[1][c]);
        acc[2][c] = fmaf(wc, m2, acc[2][c]);
      }
      acc[0][C] = fmaf(w, m0, acc[0][C]);
      acc[1][C] = fmaf(w, m1, acc[1][C]);
      acc[2][C] = fmaf(w, m2, acc[2][C]);
    }
  }
}

__device__ __forceinline__ float norm_scale(float w) {
  return 1.0f / (fmaxf(w - 1.0f, 0.0f) + 1.0f);
}

template <int C, bool kAnyL>
__global__ __launch_bounds__(kBX* kBY) void splat_fwd_kernel(
    SplatFrame f, float* __restrict__ out, float* __restrict__ accum,
    float* __restrict__ norm, float* __restrict__ accum_w) {
  const int2 tb = xcd_block_2d<16>();
  const int x = tb.x * kBX + threadIdx.x;
  const int y = tb.y * kBY + threadIdx.y;
  if (x >= f.W || y >= f.H) return;
  float acc[3][C + 1];
  accumulate_target<C, kAnyL>(f, x, y, acc);
  const int64_t P = int64_t(f.H) * f.W;
  const int64_t p = int64_t(y) * f.W + x;
  float n[3][C];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const float s = norm_scale(acc[k][C]);
#pragma unroll
    for (int c = 0; c < C; ++c) n[k][c] = acc[k][c] * s;
    if (accum) {
#pragma unroll
      for (int c = 0; c < C; ++c) accum[(k * P + p) * C + c] = acc[k][c];
    }
    if (norm) {
#pragma unroll
      for (int c = 0; c < C; ++c) norm[(k * P + p) * C + c] = n[k][c];
    }
    if (accum_w) accum_w[k * P + p] = acc[k][C];
  }
  // over(): out = n2; out = n1 + (1 - a1) out; out = n0 + (1 - a0) out.
  const float a1 = 1.0f - n[1][C - 1], a0 = 1.0f - n[0][C - 1];
#pragma unroll
  for (int c = 0; c < C; ++c) {
    float o = n[2][c];
    o = n[1][c] + a1 * o;
    o = n[0][c] + a0 * o;
    out[p * C + c] = o;
  }
}

// Pass 1 of the backward: G[p][k][0..C] = d acc (colour, weight).
template <int C, bool kAnyL>
__global__ __launch_bounds__(kBX* kBY) void splat_bwd_target_kernel(
    SplatFrame f, const float* __restrict__ d_out, float* __restrict__ G) {
  const int2 tb = xcd_block_2d<16>();
  const int x = tb.x * kBX + threadIdx.x;
  const int y = tb.y * kBY + threadIdx.y;
  if (x >= f.W || y >= f.H) return;
  float acc[3][C + 1];
  accumulate_target<C, kAnyL>(f, x, y, acc);
  const int64_t p = int64_t(y) * f.W + x;
  float s[3], n[3][C];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    s[k] = norm_scale(acc[k][C]);
#pragma unroll
    for (int c = 0; c < C; ++c) n[k][c] = acc[k][c] * s[k];
  }
  const float a0 = n[0][C - 1], a1 = n[1][C - 1];
  float g[C], dn[3][C];
  float da0 = 0.0f, da1 = 0.0f;
#pragma unroll
  for (int c = 0; c < C; ++c) {
    g[c] = d_out[p * C + c];
    const float o1 = n[1][c] + (1.0f - a1) * n[2][c];
    dn[0][c] = g[c];
    da0 -= g[c] * o1;
    const float go1 = (1.0f - a0) * g[c];
    dn[1][c] = go1;
    da1 -= go1 * n[2][c];
    dn[2][c] = (1.0f - a1) * go1;
  }
  dn[0][C - 1] += da0;
  dn[1][C - 1] += da1;
  float* gp = G + p * 3 * (C + 1);
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    // n = acc * s(acc_w); s' = -s^2 where acc_w > 1, else 0.
    const float ds = acc[k][C] > 1.0f ? -(s[k] * s[k]) : 0.0f;
    float dw = 0.0f;
#pragma unroll
    for (int c = 0; c < C; ++c) {
      gp[k * (C + 1) + c] = dn[k][c] * s[k];
      dw += dn[k][c] * acc[k][c];
    }
    gp[k * (C + 1) + C] = dw * ds;
  }
}

// Pass 2 of the backward: per source pixel, all layers.
template <int C, bool kAnyL>
__global__ __launch_bounds__(kBX* kBY) void splat_bwd_source_kernel(
    SplatFrame f, const float* __restrict__ G, float* __restrict__ d_color,
    float* __restrict__ d_xyz) {
  const int2 tb = xcd_block_2d<16>();
  const int x = tb.x * kBX + threadIdx.x;
  const int y = tb.y * kBY + threadIdx.y;
  if (x >= f.W || y >= f.H) return;
  const int64_t P = int64_t(f.H) * f.W;
  const int64_t p = int64_t(y) * f.W + x;
  int surf[9];
  surface_indices<kAnyL, false>(f, x, y, surf);
  for (int l = 0; l < f.L; ++l) {
    const int64_t q = l * P + p;
    const float m = f.mask[q];
    float dcol[C];
#pragma unroll
    for (int c = 0; c < C; ++c) dcol[c] = 0.0f;
    float gx = 0.0f, gy = 0.0f;
    if (m != 0.0f) {
      float col[C];
#pragma unroll
      for (int c = 0; c < C; ++c) col[c] = f.color[q * C + c];
      const float sx = f.xyz[3 * q], sy = f.xyz[3 * q + 1];
#pragma unroll
      for (int t = 0; t < 9; ++t) {
        const int u = t / 3 - 1, v = t % 3 - 1;
        const int tx = x + u, ty = y + v;
        if (tx < 0 || tx >= f.W || ty < 0 || ty >= f.H) continue;
        float dx, dy;
        const float w = tap_weight(m, f.coef, sx, sy, u, v, dx, dy);
        const int k = layer_class(surf[t], l);
        const float* gp =
            G + ((int64_t(ty) * f.W + tx) * 3 + k) * (C + 1);
        float dw = gp[C];
#pragma unroll
        for (int c = 0; c < C; ++c) {
          dcol[c] += w * gp[c];
          dw += gp[c] * col[c];
        }
        const float dw4 = dw * 4.0f * w;
        gx += dw4 * dx;
        gy += dw4 * dy;
      }
    }
#pragma unroll
    for (int c = 0; c < C; ++c) d_color[q * C + c] = dcol[c];
    d_xyz[3 * q] = gx;
    d_xyz[3 * q + 1] = gy;
    d_xyz[3 * q + 2] = 0.0f;
  }
}

float splat_coef() {
  double sum = 0.0;
  for (int u = -1; u <= 1; ++u)
    for (int v = -1; v <= 1; ++v) sum += std::exp(-2.0 * (u * u + v * v));
  return float(1.05 / sum);
}

template <int C>
int launch_fwd(const SplatFrame& f, float* out, float* accum, float* norm,
               float* accum_w, hipStream_t s) {
  dim3 grid((f.W + kBX - 1) / kBX, (f.H + kBY - 1) / kBY);
  if (f.L <= kMaxLayers)
    splat_fwd_kernel<C, false><<<grid, dim3(kBX, kBY), 0, s>>>(
        f, out, accum, norm, accum_w);
  else
    splat_fwd_kernel<C, true><<<grid, dim3(kBX, kBY), 0, s>>>(
        f, out, accum, norm, accum_w);
  DR_CHECK_LAUNCH("splat fwd");
  return kOk;
}

template <int C>
int launch_bwd(const SplatFrame& f, const float* d_out, float* G,
               float* d_color, float* d_xyz, hipStream_t s) {
  dim3 grid((f.W + kBX - 1) / kBX, (f.H + kBY - 1) / kBY);
  if (f.L <= kMaxLayers) {
    splat_bwd_target_kernel<C, false><<<grid, dim3(kBX, kBY), 0, s>>>(
        f, d_out, G);
    DR_CHECK_LAUNCH("splat bwd target");
    splat_bwd_source_kernel<C, false><<<grid, dim3(kBX, kBY), 0, s>>>(
        f, G, d_color, d_xyz);
  } else {
    splat_bwd_target_kernel<C, true><<<grid, dim3(kBX, kBY), 0, s>>>(
        f, d_out, G);
    DR_CHECK_LAUNCH("splat bwd target");
    splat_bwd_source_kernel<C, true><<<grid, dim3(kBX, kBY), 0, s>>>(
        f, G, d_color, d_xyz);
  }
  DR_CHECK_LAUNCH("splat bwd source");
  return kOk;
}

#define DR_SPLAT_SWITCH(C, CALL) \
  switch (C) {                   \
    case 1: return CALL(1);      \
    case 2: return CALL(2);      \
    case 3: return CALL(3);      \
    case 4: return CALL(4);      \
    case 5: return CALL(5);      \
    case 6: return CALL(6);      \
    case 7: return CALL(7);      \
    case 8: return CALL(8);      \
    case 9: return CALL(9);      \
    case 10: return CALL(10);    \
    case 11: return CALL(11);    \
    case 12: return CALL(12);    \
    case 13: return CALL(13);    \
    case 14: return CALL(14);    \
    case 15: return CALL(15);    \
    case 16: return CALL(16);    \
    case 20: return CALL(20);    \
    case 24: return CALL(24);    \
    case 28: return CALL(28);    \
    case 32: return CALL(32);    \
  }

// Channel counts the kernels are instantiated for; callers pad other counts
// up to one of them (or split them), see splat.py.
bool splat_channels_ok(int c) {
  return (c >= 1 && c <= 16) || c == 20 || c == 24 || c == 28 || c == 32;
}

int check_frame(const SplatFrame& f, int C) {
  DR_CHECK_ARG(f.L >= 1, "num_layers must be >= 1");
  DR_CHECK_ARG(f.H > 0 && f.W > 0, "image size must be > 0");
  DR_CHECK_ARG(splat_channels_ok(C),
               "splat channel count must be 1-16, 20, 24, 28 or 32");
  DR_CHECK_ARG(f.xyz && f.color && f.mask, "null input pointer");
  return kOk;
}

}  // namespace
}  // namespace dr

extern "C" {

float dr_splat_coefficient(void) { return dr::splat_coef(); }

size_t dr_splat_bwd_workspace_size(int32_t height, int32_t width,
                                   int32_t channels) {
  return size_t(height) * width * 3 * (channels + 1) * sizeof(float);
}

int dr_splat_fwd(const float* xyz, const float* color, const float* mask,
                 int32_t num_layers, int32_t height, int32_t width,
                 int32_t channels, float* out, float* accum, float* norm,
                 float* accum_weights, dr_stream_t stream) {
  dr::SplatFrame f{xyz, color, mask, num_layers, height, width,
                   dr::splat_coef()};
  const int rc = dr::check_frame(f, channels);
  if (rc != DR_OK) return rc;
  DR_CHECK_ARG(out != nullptr, "out is null");
  hipStream_t s = static_cast<hipStream_t>(stream);
#define DR_CALL(CC) dr::launch_fwd<CC>(f, out, accum, norm, accum_weights, s)
  DR_SPLAT_SWITCH(channels, DR_CALL)
#undef DR_CALL
  return dr::fail(dr::kUnsupported, "unsupported channel count");
}

int dr_splat_bwd(const float* xyz, const float* color, const float* mask,
                 int32_t num_layers, int32_t height, int32_t width,
                 int32_t channels, const float* d_out, float* d_color,
                 float* d_xyz, void* workspace, size_t workspace_bytes,
                 dr_stream_t stream) {
  dr::SplatFrame f{xyz, color, mask, num_layers, height, width,
                   dr::splat_coef()};
  const int rc = dr::check_frame(f, channels);
  if (rc != DR_OK) return rc;
  DR_CHECK_ARG(d_out && d_color && d_xyz, "null gradient pointer");
  DR_CHECK_ARG(workspace != nullptr, "workspace is null");
  if (workspace_bytes < dr_splat_bwd_workspace_size(height, width, channels))
    return dr::fail(dr::kWorkspaceTooSmall, "splat bwd workspace too small");
  hipStream_t s = static_cast<hipStream_t>(stream);
  float* G = static_cast<float*>(workspace);
#define DR_CALL(CC) dr::launch_bwd<CC>(f, d_out, G, d_color, d_xyz, s)
  DR_SPLAT_SWITCH(channels, DR_CALL)
#undef DR_CALL
  return dr::fail(dr::kUnsupported, "unsupported channel count");
}

}  // extern "C"
