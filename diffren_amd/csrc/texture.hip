// Bilinear texture resampling for gfx950, forward and backward:
// image.bilinear_resample (diffren/jax/utils/image.py:21-118), the gather at
// the heart of shaders.texture_map (diffren/jax/utils/shaders.py:47-72).
//
// image [H, W, C] (HWC, float32), sample locations [N, 2] as (x, y) in image
// units, texel (i, j) centred at (j + 0.5, i + 0.5).  Per sample, as the
// reference evaluates it on a 1-texel padded image (constant 0 or edge
// replicate, :66-67):
//   l = (y - 0.5 + 1, x - 0.5 + 1)
//   f = clip(floor(l), 0, (H, W))                  (:69-71)
//   c = clip(l - f, 0, 1),  w = 1 - c              (:73-76)
//   out = (((P00 w0) w1 + (P10 c0) w1) + (P11 c0) c1) + (P01 w0) c1
// with Pab the padded texel (f0 + a, f1 + b) (:86-92).  One thread per
// sample; channels are contiguous, so a sample reads four C-float runs.
//
// Backward: d image[texel] += d out * corner weight (float atomics; padded
// texels map to nothing for constant padding, to the clamped edge texel for
// edge padding); d (x, y) from the corner-weight derivatives times the
// derivative of the clip, which follows JAX's differentiation of
// clip = minimum(maximum(t, 0), 1): 1 inside (0, 1), 1/2 where t equals a
// bound (tie), 0 outside.
#pragma clang fp contract(off)

#include "dr_common.h"
#include "../../include/diffren_amd.h"

namespace dr {
namespace {

constexpr int kBlock = 256;

struct Footprint {
  int f0, f1;        // padded-image corner indices
  float c0, c1;      // ceil weights
  float t0, t1;      // l - f before clipping
};

__device__ __forceinline__ Footprint footprint(float x, float y, int H,
                                               int W) {
  Footprint fp;
  const float l0 = (y - 0.5f) + 1.0f;
  const float l1 = (x - 0.5f) + 1.0f;
  // fmaxf maps NaN to the lower bound, keeping every access in bounds.
  const float g0 = fminf(fmaxf(floorf(l0), 0.0f), float(H));
  const float g1 = fminf(fmaxf(floorf(l1), 0.0f), float(W));
  fp.f0 = int(g0);
  fp.f1 = int(g1);
  fp.t0 = l0 - g0;
  fp.t1 = l1 - g1;
  fp.c0 = fminf(fmaxf(fp.t0, 0.0f), 1.0f);
  fp.c1 = fminf(fmaxf(fp.t1, 0.0f), 1.0f);
  return fp;
}

// Texel of the padded image [H + 2, W + 2] at (i, j), as an index into the
// unpadded image, or -1 for a constant-padding zero.
__device__ __forceinline__ int64_t texel(int i, int j, int H, int W,
                                         bool edge) {
  int a = i - 1, b = j - 1;
  if (edge) {
    a = min(max(a, 0), H - 1);
    b = min(max(b, 0), W - 1);
  } else if (a < 0 || a >= H || b < 0 || b >= W) {
    return -1;
  }
  return int64_t(a) * W + b;
}

__device__ __forceinline__ float fetch(const float* __restrict__ img,
                                       int64_t t, int C, int c) {
  return t < 0 ? 0.0f : img[t * C + c];
}

__device__ __forceinline__ float clip_grad(float t) {
  if (t > 0.0f && t < 1.0f) return 1.0f;
  if (t == 0.0f || t == 1.0f) return 0.5f;
  return 0.0f;
}

__global__ __launch_bounds__(kBlock) void resample_fwd_kernel(
    const float* __restrict__ img, int H, int W, int C,
    const float2* __restrict__ loc, int64_t N, bool edge,
    float* __restrict__ out) {
  const int64_t n = int64_t(blockIdx.x) * kBlock + threadIdx.x;
  if (n >= N) return;
  const float2 xy = loc[n];
  const Footprint fp = footprint(xy.x, xy.y, H, W);
  const float w0 = 1.0f - fp.c0, w1 = 1.0f - fp.c1;
  const int64_t t00 = texel(fp.f0, fp.f1, H, W, edge);
  const int64_t t10 = texel(fp.f0 + 1, fp.f1, H, W, edge);
  const int64_t t11 = texel(fp.f0 + 1, fp.f1 + 1, H, W, edge);
  const int64_t t01 = texel(fp.f0, fp.f1 + 1, H, W, edge);
  float* o = out + n * C;
  for (int c = 0; c < C; ++c) {
    const float ll = (fetch(img, t00, C, c) * w0) * w1;
    const float ul = (fetch(img, t10, C, c) * fp.c0) * w1;
    const float ur = (fetch(img, t11, C, c) * fp.c0) * fp.c1;
    const float lr = (fetch(img, t01, C, c) * w0) * fp.c1;
    o[c] = ((ll + ul) + ur) + lr;
  }
}

__global__ __launch_bounds__(kBlock) void resample_bwd_kernel(
    const float* __restrict__ img, int H, int W, int C,
    const float2* __restrict__ loc, int64_t N, bool edge,
    const float* __restrict__ dout, float* __restrict__ dimg,
    float2* __restrict__ dloc) {
  const int64_t n = int64_t(blockIdx.x) * kBlock + threadIdx.x;
  if (n >= N) return;
  const float2 xy = loc[n];
  const Footprint fp = footprint(xy.x, xy.y, H, W);
  const float w0 = 1.0f - fp.c0, w1 = 1.0f - fp.c1;
  const int64_t t00 = texel(fp.f0, fp.f1, H, W, edge);
  const int64_t t10 = texel(fp.f0 + 1, fp.f1, H, W, edge);
  const int64_t t11 = texel(fp.f0 + 1, fp.f1 + 1, H, W, edge);
  const int64_t t01 = texel(fp.f0, fp.f1 + 1, H, W, edge);
  const float* g = dout + n * C;
  float dc0 = 0.0f, dc1 = 0.0f;
  for (int c = 0; c < C; ++c) {
    const float gc = g[c];
    if (dimg != nullptr && gc != 0.0f) {
      if (t00 >= 0) unsafeAtomicAdd(dimg + t00 * C + c, gc * (w0 * w1));
      if (t10 >= 0) unsafeAtomicAdd(dimg + t10 * C + c, gc * (fp.c0 * w1));
      if (t11 >= 0) unsafeAtomicAdd(dimg + t11 * C + c, gc * (fp.c0 * fp.c1));
      if (t01 >= 0) unsafeAtomicAdd(dimg + t01 * C + c, gc * (w0 * fp.c1));
    }
    const float p00 = fetch(img, t00, C, c), p10 = fetch(img, t10, C, c);
    const float p11 = fetch(img, t11, C, c), p01 = fetch(img, t01, C, c);
    // d out / d c0 and d out / d c1 (w = 1 - c).
    dc0 += gc * (((p10 - p00) * w1) + ((p11 - p01) * fp.c1));
    dc1 += gc * (((p01 - p00) * w0) + ((p11 - p10) * fp.c0));
  }
  if (dloc != nullptr)
    dloc[n] = make_float2(dc1 * clip_grad(fp.t1), dc0 * clip_grad(fp.t0));
}

inline unsigned blocks_for(int64_t n) {
  return unsigned((n + kBlock - 1) / kBlock);
}

int check_args(const float* image, int H, int W, int C, const float* loc,
               int64_t N, int pad_mode) {
  DR_CHECK_ARG(H > 0 && W > 0 && C > 0, "image must be non-empty [H, W, C]");
  DR_CHECK_ARG(N >= 0, "negative sample count");
  DR_CHECK_ARG(pad_mode == DR_PAD_CONSTANT || pad_mode == DR_PAD_EDGE,
               "pad_mode must be DR_PAD_CONSTANT or DR_PAD_EDGE");
  DR_CHECK_ARG(N == 0 || (image && loc), "null pointer");
  DR_CHECK_ARG((reinterpret_cast<uintptr_t>(loc) & 7) == 0,
               "locations must be 8-byte aligned [N, 2] float32");
  return DR_OK;
}

}  // namespace
}  // namespace dr

extern "C" {

int dr_bilinear_resample_fwd(const float* image, int32_t height,
                             int32_t width, int32_t channels,
                             const float* locations, int64_t num_samples,
                             int32_t pad_mode, float* out,
                             dr_stream_t stream) {
  const int rc = dr::check_args(image, height, width, channels, locations,
                                num_samples, pad_mode);
  if (rc != DR_OK) return rc;
  if (num_samples == 0) return DR_OK;
  DR_CHECK_ARG(out != nullptr, "null output");
  dr::resample_fwd_kernel<<<dr::blocks_for(num_samples), dr::kBlock, 0,
                            static_cast<hipStream_t>(stream)>>>(
      image, height, width, channels,
      reinterpret_cast<const float2*>(locations), num_samples,
      pad_mode == DR_PAD_EDGE, out);
  DR_CHECK_LAUNCH("bilinear resample fwd");
  return DR_OK;
}

int dr_bilinear_resample_bwd(const float* image, int32_t height,
                             int32_t width, int32_t channels,
                             const float* locations, int64_t num_samples,
                             int32_t pad_mode, const float* d_out,
                             float* d_image, float* d_locations,
                             dr_stream_t stream) {
  const int rc = dr::check_args(image, height, width, channels, locations,
                                num_samples, pad_mode);
  if (rc != DR_OK) return rc;
  if (num_samples == 0) return DR_OK;
  DR_CHECK_ARG(d_out != nullptr, "null output gradient");
  DR_CHECK_ARG((reinterpret_cast<uintptr_t>(d_locations) & 7) == 0,
               "d_locations must be 8-byte aligned");
  if (d_image == nullptr && d_locations == nullptr) return DR_OK;
  dr::resample_bwd_kernel<<<dr::blocks_for(num_samples), dr::kBlock, 0,
                            static_cast<hipStream_t>(stream)>>>(
      image, height, width, channels,
      reinterpret_cast<const float2*>(locations), num_samples,
      pad_mode == DR_PAD_EDGE, d_out, d_image,
      reinterpret_cast<float2*>(d_locations));
  DR_CHECK_LAUNCH("bilinear resample bwd");
  return DR_OK;
}

}  // extern "C"
