// Barycentric attribute interpolation for gfx950, forward and backward, plus
// the homogeneous vertex transform used to reach clip space.
//
// Forward (diffren/jax/internal/interpolate.py:24-63, vmapped over layers at
// :118-145):  for every pixel-layer p with triangle id t = ids[p],
//   out[p, c] = (b0 * A[v0, c] + b1 * A[v1, c]) + b2 * A[v2, c]   if t >= 0
//   out[p, c] = background[c]                                      otherwise
// with (v0, v1, v2) = triangles[t].  The reference evaluates
// mask * sum + (1 - mask) * background, which is the same for finite inputs.
// The vertex ids are gathered here from the triangle id, so no vertex_ids
// frame (rasterize.py:111-113) is ever written.
//
// Backward: dA[v_k, c] += b_k * dout[p, c] for foreground p (the transpose of
// the gather; float atomics, so the summation order is not fixed).
#pragma clang fp contract(off)

#include "dr_common.h"
#include "../../include/diffren_amd.h"

namespace dr {
namespace {

constexpr int kBlock = 256;

__device__ __forceinline__ bool pixel_vertices(const int32_t* __restrict__ ids,
                                               const int32_t* __restrict__ tris,
                                               int32_t T, int32_t V, int64_t p,
                                               int& v0, int& v1, int& v2) {
  const int t = ids[p];
  if (t < 0 || t >= T) return false;
  v0 = tris[3 * int64_t(t) + 0];
  v1 = tris[3 * int64_t(t) + 1];
  v2 = tris[3 * int64_t(t) + 2];
  return unsigned(v0) < unsigned(V) && unsigned(v1) < unsigned(V) &&
         unsigned(v2) < unsigned(V);
}

// One thread per pixel-layer; C4 float4 channels when the rows are aligned.
template <bool kVec4>
__global__ __launch_bounds__(kBlock) void interp_fwd_kernel(
    const int32_t* __restrict__ ids, const float* __restrict__ barys,
    const int32_t* __restrict__ tris, int32_t T, const float* __restrict__ A,
    int32_t V, int32_t C, const float* __restrict__ bg, int64_t N,
    float* __restrict__ out) {
  const int64_t p = int64_t(blockIdx.x) * kBlock + threadIdx.x;
  if (p >= N) return;
  int v0, v1, v2;
  float* o = out + p * C;
  if (!pixel_vertices(ids, tris, T, V, p, v0, v1, v2)) {
    for (int c = 0; c < C; ++c) o[c] = bg ? bg[c] : 0.0f;
    return;
  }
  const float b0 = barys[3 * p], b1 = barys[3 * p + 1], b2 = barys[3 * p + 2];
  if (kVec4) {
    const float4* a0 = reinterpret_cast<const float4*>(A + int64_t(v0) * C);
    const float4* a1 = reinterpret_cast<const float4*>(A + int64_t(v1) * C);
    const float4* a2 = reinterpret_cast<const float4*>(A + int64_t(v2) * C);
    float4* o4 = reinterpret_cast<float4*>(o);
    for (int c = 0; c < C / 4; ++c) {
      const float4 x = a0[c], y = a1[c], z = a2[c];
      float4 r;
      r.x = (b0 * x.x + b1 * y.x) + b2 * z.x;
      r.y = (b0 * x.y + b1 * y.y) + b2 * z.y;
      r.z = (b0 * x.z + b1 * y.z) + b2 * z.z;
      r.w = (b0 * x.w + b1 * y.w) + b2 * z.w;
      o4[c] = r;
    }
  } else {
    const float* a0 = A + int64_t(v0) * C;
    const float* a1 = A + int64_t(v1) * C;
    const float* a2 = A + int64_t(v2) * C;
    for (int c = 0; c < C; ++c) o[c] = (b0 * a0[c] + b1 * a1[c]) + b2 * a2[c];
  }
}

// When C divides 64 (C = 1, 2, 4, 8, 16, 32, 64) the wave's 64 x 3
// (pixel, vertex) records are flushed C lanes per record: in round r the
// lanes take records r * (64 / C) ... in lane order, each lane one channel,
// so one atomic instruction touches 64 / C contiguous C-float rows instead
// of 64 scattered addresses (grouped float atomics run several times faster
// than scattered ones).  Same atomics, same values; only their grouping
// differs.  Other C: one pixel per lane.
__global__ __launch_bounds__(kBlock) void interp_bwd_kernel(
    const int32_t* __restrict__ ids, const float* __restrict__ barys,
    const int32_t* __restrict__ tris, int32_t T, const float* __restrict__ dout,
    int32_t V, int32_t C, int64_t N, float* __restrict__ dA) {
  const int64_t p = int64_t(blockIdx.x) * kBlock + threadIdx.x;
  int v[3] = {0, 0, 0};
  float b[3] = {0.0f, 0.0f, 0.0f};
  const bool ok = p < N && pixel_vertices(ids, tris, T, V, p, v[0], v[1], v[2]);
  if (ok) {
#pragma unroll
    for (int k = 0; k < 3; ++k) b[k] = barys[3 * p + k];
  }
  if (64 % C == 0) {
    // Whole waves take part (no lane has left): the shuffles read every lane.
    const int lane = __lane_id();
    const int per = 64 / C, c = lane % C;
    const int64_t wave_base = p - lane;
    for (int r = 0; r < 3 * C; ++r) {
      const int pair = r * per + lane / C;  // record: vertex k of pixel src
      const int k = (r * per) >> 6;         // the same for the whole wave
      const int src = pair & 63;
      const int vk = __shfl(k == 0 ? v[0] : (k == 1 ? v[1] : v[2]), src);
      const float bk = __shfl(k == 0 ? b[0] : (k == 1 ? b[1] : b[2]), src);
      if (!__shfl(int(ok), src)) continue;
      const float gc = dout[(wave_base + src) * C + c];
      if (gc == 0.0f) continue;
      gatomic_add(dA + int64_t(vk) * C + c, bk * gc);
    }
    return;
  }
  if (!ok) return;
  const float* g = dout + p * C;
  for (int c = 0; c < C; ++c) {
    const float gc = g[c];
    if (gc == 0.0f) continue;
#pragma unroll
    for (int k = 0; k < 3; ++k)
      gatomic_add(dA + int64_t(v[k]) * C + c, b[k] * gc);
  }
}

// d_barys[p, k] = sum_c dout[p, c] A[v_k, c] for foreground p, 0 elsewhere:
// the barycentric half of the interpolation VJP, needed when the
// barycentrics are themselves differentiable (rasterize.py:120-121).
__global__ __launch_bounds__(kBlock) void interp_bwd_barys_kernel(
    const int32_t* __restrict__ ids, const int32_t* __restrict__ tris,
    int32_t T, const float* __restrict__ A, int32_t V, int32_t C,
    const float* __restrict__ dout, int64_t N, float* __restrict__ dbarys) {
  const int64_t p = int64_t(blockIdx.x) * kBlock + threadIdx.x;
  if (p >= N) return;
  int v0, v1, v2;
  float* d = dbarys + 3 * p;
  if (!pixel_vertices(ids, tris, T, V, p, v0, v1, v2)) {
    d[0] = d[1] = d[2] = 0.0f;
    return;
  }
  const float* g = dout + p * C;
  const float* a0 = A + int64_t(v0) * C;
  const float* a1 = A + int64_t(v1) * C;
  const float* a2 = A + int64_t(v2) * C;
  float s0 = 0.0f, s1 = 0.0f, s2 = 0.0f;
  for (int c = 0; c < C; ++c) {
    const float gc = g[c];
    s0 += gc * a0[c];
    s1 += gc * a1[c];
    s2 += gc * a2[c];
  }
  d[0] = s0;
  d[1] = s1;
  d[2] = s2;
}

// clip[v, i] = ((M[i,0] x + M[i,1] y) + M[i,2] z) + M[i,3] w  (w = 1 for xyz
// input): transforms.py:40-78 for a [4, 4] matrix, one fixed summation order.
__global__ __launch_bounds__(kBlock) void transform_kernel(
    const float* __restrict__ M, const float* __restrict__ in, int32_t D,
    int64_t V, float* __restrict__ out) {
  const int64_t v = int64_t(blockIdx.x) * kBlock + threadIdx.x;
  if (v >= V) return;
  const float x = in[v * D], y = in[v * D + 1], z = in[v * D + 2];
  const float w = D == 4 ? in[v * D + 3] : 1.0f;
  float4 r;
  float* rp = &r.x;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const float* m = M + 4 * i;
    rp[i] = ((m[0] * x + m[1] * y) + m[2] * z) + m[3] * w;
  }
  reinterpret_cast<float4*>(out)[v] = r;
}

// d_in[v, j] = sum_i M[i, j] * d_clip[v, i] for j < D (d w dropped for D=3).
__global__ __launch_bounds__(kBlock) void transform_bwd_kernel(
    const float* __restrict__ M, const float* __restrict__ dclip, int32_t D,
    int64_t V, float* __restrict__ din) {
  const int64_t v = int64_t(blockIdx.x) * kBlock + threadIdx.x;
  if (v >= V) return;
  const float4 g = reinterpret_cast<const float4*>(dclip)[v];
  for (int j = 0; j < D; ++j)
    din[v * D + j] = ((M[j] * g.x + M[4 + j] * g.y) + M[8 + j] * g.z) +
                     M[12 + j] * g.w;
}

// One mesh through `views` matrices [views][4][4]: out [views][V][4].
__global__ __launch_bounds__(kBlock) void transform_multi_kernel(
    const float* __restrict__ Ms, const float* __restrict__ in, int32_t D,
    int64_t V, float* __restrict__ out) {
  const int64_t v = int64_t(blockIdx.x) * kBlock + threadIdx.x;
  if (v >= V) return;
  const float* M = Ms + 16 * blockIdx.y;
  const float x = in[v * D], y = in[v * D + 1], z = in[v * D + 2];
  const float w = D == 4 ? in[v * D + 3] : 1.0f;
  float4 r;
  float* rp = &r.x;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const float* m = M + 4 * i;
    rp[i] = ((m[0] * x + m[1] * y) + m[2] * z) + m[3] * w;
  }
  reinterpret_cast<float4*>(out)[int64_t(blockIdx.y) * V + v] = r;
}

inline unsigned blocks_for(int64_t n) {
  return unsigned((n + kBlock - 1) / kBlock);
}

}  // namespace
}  // namespace dr

extern "C" {

int dr_interpolate_fwd(const int32_t* triangle_ids, const float* barycentrics,
                       const int32_t* triangles, int32_t num_triangles,
                       const float* attributes, int32_t num_vertices,
                       int32_t channels, const float* background,
                       int64_t num_pixels, float* out, dr_stream_t stream) {
  DR_CHECK_ARG(channels > 0, "channels must be > 0");
  DR_CHECK_ARG(num_pixels >= 0, "num_pixels must be >= 0");
  DR_CHECK_ARG(num_triangles >= 0 && num_vertices >= 0, "negative mesh size");
  if (num_pixels == 0) return DR_OK;
  DR_CHECK_ARG(triangle_ids && barycentrics && out, "null frame pointer");
  DR_CHECK_ARG(num_triangles == 0 || (triangles && attributes),
               "null mesh pointer");
  hipStream_t s = static_cast<hipStream_t>(stream);
  const bool vec4 = channels % 4 == 0 &&
                    (reinterpret_cast<uintptr_t>(attributes) & 15) == 0 &&
                    (reinterpret_cast<uintptr_t>(out) & 15) == 0;
  if (vec4)
    dr::interp_fwd_kernel<true><<<dr::blocks_for(num_pixels), dr::kBlock, 0,
                                  s>>>(triangle_ids, barycentrics, triangles,
                                       num_triangles, attributes, num_vertices,
                                       channels, background, num_pixels, out);
  else
    dr::interp_fwd_kernel<false><<<dr::blocks_for(num_pixels), dr::kBlock, 0,
                                   s>>>(triangle_ids, barycentrics, triangles,
                                        num_triangles, attributes,
                                        num_vertices, channels, background,
                                        num_pixels, out);
  DR_CHECK_LAUNCH("interpolate fwd");
  return DR_OK;
}

int dr_interpolate_bwd(const int32_t* triangle_ids, const float* barycentrics,
                       const int32_t* triangles, int32_t num_triangles,
                       const float* d_out, int32_t num_vertices,
                       int32_t channels, int64_t num_pixels,
                       float* d_attributes, dr_stream_t stream) {
  DR_CHECK_ARG(channels > 0, "channels must be > 0");
  DR_CHECK_ARG(num_pixels >= 0, "num_pixels must be >= 0");
  if (num_pixels == 0 || num_triangles == 0) return DR_OK;
  DR_CHECK_ARG(triangle_ids && barycentrics && d_out && d_attributes &&
                   triangles,
               "null pointer");
  dr::interp_bwd_kernel<<<dr::blocks_for(num_pixels), dr::kBlock, 0,
                          static_cast<hipStream_t>(stream)>>>(
      triangle_ids, barycentrics, triangles, num_triangles, d_out,
      num_vertices, channels, num_pixels, d_attributes);
  DR_CHECK_LAUNCH("interpolate bwd");
  return DR_OK;
}

int dr_interpolate_bwd_barycentrics(const int32_t* triangle_ids,
                                    const int32_t* triangles,
                                    int32_t num_triangles,
                                    const float* attributes,
                                    int32_t num_vertices, int32_t channels,
                                    const float* d_out, int64_t num_pixels,
                                    float* d_barycentrics,
                                    dr_stream_t stream) {
  DR_CHECK_ARG(channels > 0, "channels must be > 0");
  DR_CHECK_ARG(num_pixels >= 0, "num_pixels must be >= 0");
  DR_CHECK_ARG(num_triangles >= 0 && num_vertices >= 0, "negative mesh size");
  if (num_pixels == 0) return DR_OK;
  DR_CHECK_ARG(triangle_ids && d_out && d_barycentrics, "null pointer");
  DR_CHECK_ARG(num_triangles == 0 || (triangles && attributes),
               "null mesh pointer");
  dr::interp_bwd_barys_kernel<<<dr::blocks_for(num_pixels), dr::kBlock, 0,
                                static_cast<hipStream_t>(stream)>>>(
      triangle_ids, triangles, num_triangles, attributes, num_vertices,
      channels, d_out, num_pixels, d_barycentrics);
  DR_CHECK_LAUNCH("interpolate bwd barycentrics");
  return DR_OK;
}

int dr_transform_homogeneous(const float* matrix, const float* vertices,
                             int32_t dims, int64_t num_vertices, float* out,
                             dr_stream_t stream) {
  DR_CHECK_ARG(dims == 3 || dims == 4, "vertices must be xyz or xyzw");
  DR_CHECK_ARG(num_vertices >= 0, "num_vertices must be >= 0");
  if (num_vertices == 0) return DR_OK;
  DR_CHECK_ARG(matrix && vertices && out, "null pointer");
  DR_CHECK_ARG((reinterpret_cast<uintptr_t>(out) & 15) == 0,
               "out must be 16-byte aligned");
  dr::transform_kernel<<<dr::blocks_for(num_vertices), dr::kBlock, 0,
                         static_cast<hipStream_t>(stream)>>>(
      matrix, vertices, dims, num_vertices, out);
  DR_CHECK_LAUNCH("transform");
  return DR_OK;
}

int dr_transform_homogeneous_multi(const float* matrices, int32_t views,
                                   const float* vertices, int32_t dims,
                                   int64_t num_vertices, float* out,
                                   dr_stream_t stream) {
  DR_CHECK_ARG(dims == 3 || dims == 4, "vertices must be xyz or xyzw");
  DR_CHECK_ARG(num_vertices >= 0 && views >= 0 && views <= 65535,
               "bad sizes");
  if (num_vertices == 0 || views == 0) return DR_OK;
  DR_CHECK_ARG(matrices && vertices && out, "null pointer");
  DR_CHECK_ARG((reinterpret_cast<uintptr_t>(out) & 15) == 0,
               "out must be 16-byte aligned");
  dr::transform_multi_kernel<<<dim3(dr::blocks_for(num_vertices), views),
                               dr::kBlock, 0,
                               static_cast<hipStream_t>(stream)>>>(
      matrices, vertices, dims, num_vertices, out);
  DR_CHECK_LAUNCH("transform multi");
  return DR_OK;
}

int dr_transform_homogeneous_bwd(const float* matrix, const float* d_clip,
                                 int32_t dims, int64_t num_vertices,
                                 float* d_vertices, dr_stream_t stream) {
  DR_CHECK_ARG(dims == 3 || dims == 4, "vertices must be xyz or xyzw");
  DR_CHECK_ARG(num_vertices >= 0, "num_vertices must be >= 0");
  if (num_vertices == 0) return DR_OK;
  DR_CHECK_ARG(matrix && d_clip && d_vertices, "null pointer");
  DR_CHECK_ARG((reinterpret_cast<uintptr_t>(d_clip) & 15) == 0,
               "d_clip must be 16-byte aligned");
  dr::transform_bwd_kernel<<<dr::blocks_for(num_vertices), dr::kBlock, 0,
                             static_cast<hipStream_t>(stream)>>>(
      matrix, d_clip, dims, num_vertices, d_vertices);
  DR_CHECK_LAUNCH("transform bwd");
  return DR_OK;
}

}  // extern "C"
