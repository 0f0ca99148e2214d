// Differentiable barycentric coordinates for gfx950, forward and backward:
// the compute_diff_barys path of diffren/jax/internal/rasterize.py:120-256
// (differentiable_barycentrics, normalized_pixel_coordinates,
// compute_triangle_matrix, compute_barycentric_coordinates).
//
// For every pixel-layer p = (l, y, x) with triangle id t = ids[p] >= 0 and
// clip vertices V_k = (x_k, y_k, w_k) of triangles[t]:
//   px = (2x + 1) / W - 1,  py = (2y + 1) / H - 1      (:184-189)
//   d_k = V_{k+1} x V_{k+2} (the adjugate rows d_k1, d_k2, d_k3, :209-217)
//   det = (x0 d_01 + y0 d_02) + w0 d_03,  sign = sign(det), 1 if det == 0
//   r_k = (sign d_k1 px + sign d_k2 py) + sign d_k3          (:243-245)
//   b_k = nan_to_num(r_k / ((r_0 + r_1) + r_2))              (:252-254)
// and b = 0 on background (the reference multiplies by the foreground mask,
// :158-159).  One thread per pixel-layer recomputes the 3x3 matrix from the
// three vertex loads instead of storing a [T, 3, 3] table: the same float
// expressions in the same order, no extra pass.  Float expressions follow
// the jnp source left to right with contraction off.
//
// Backward (the reference differentiates this with XLA autodiff): with
// q_k = r_k / s, s = sum r, and db the incoming gradient,
//   dq_k = db_k where q_k is finite (nan_to_num passes finite values only),
//   dr_k = (dq_k - sum_i dq_i q_i) / s,      g_k = sign dr_k,
//   r_k = sign P . (V_a x V_b)  (a = k+1, b = k+2, P = (px, py, 1)), so
//   dV_a += g_k (V_b x P),  dV_b += g_k (P x V_a),
// scattered to d_clip[v][0, 1, 3] (z gets no gradient) with float atomics.
#pragma clang fp contract(off)

#include <cfloat>

#include "dr_common.h"
#include "../../include/diffren_amd.h"

namespace dr {
namespace {

constexpr int kBlock = 256;

struct Vec3 {
  float x, y, w;
};

__device__ __forceinline__ Vec3 cross(const Vec3& a, const Vec3& b) {
  return {a.y * b.w - a.w * b.y, a.w * b.x - a.x * b.w, a.x * b.y - a.y * b.x};
}

__device__ __forceinline__ float nan_to_num(float v) {
  if (v != v) return 0.0f;
  if (v == INFINITY) return FLT_MAX;
  if (v == -INFINITY) return -FLT_MAX;
  return v;
}

__device__ __forceinline__ bool finite(float v) {
  return v == v && v != INFINITY && v != -INFINITY;
}

struct PixelTriangle {
  int v[3];
  Vec3 P[3];
};

__device__ __forceinline__ bool load_triangle(
    const int32_t* __restrict__ ids, const int32_t* __restrict__ tris,
    int32_t T, const float4* __restrict__ clip, int64_t V, int64_t p,
    PixelTriangle& pt) {
  const int t = ids[p];
  if (t < 0 || t >= T) return false;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    pt.v[k] = tris[3 * int64_t(t) + k];
    if (pt.v[k] < 0 || pt.v[k] >= V) return false;
  }
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const float4 c = clip[pt.v[k]];
    pt.P[k] = {c.x, c.y, c.w};
  }
  return true;
}

// Pixel centre in the reference's normalized coordinates: an int32 ramp
// divided by the size in float32, minus 1.
__device__ __forceinline__ float ndc_center(int i, int n) {
  return static_cast<float>(2 * i + 1) / static_cast<float>(n) - 1.0f;
}

// Sign-corrected adjugate rows M_k = sign * (V_{k+1} x V_{k+2}) and the
// unnormalised coordinates r_k, exactly as compute_triangle_matrix /
// compute_barycentric_coordinates evaluate them.
__device__ __forceinline__ void raw_coordinates(const PixelTriangle& pt,
                                                float px, float py,
                                                float& sign, float r[3]) {
  const Vec3 d0 = cross(pt.P[1], pt.P[2]);
  const Vec3 d1 = cross(pt.P[2], pt.P[0]);
  const Vec3 d2 = cross(pt.P[0], pt.P[1]);
  // determinant = xyw[0,0] d11 + xyw[1,0] d12 + xyw[2,0] d13
  const float det = (pt.P[0].x * d0.x + pt.P[0].y * d0.y) + pt.P[0].w * d0.w;
  // jnp.sign(det) + (det == 0): +1, -1, 1 for zero, NaN for NaN.
  sign = det > 0.0f ? 1.0f : (det < 0.0f ? -1.0f : (det == 0.0f ? 1.0f : det));
  const Vec3 d[3] = {d0, d1, d2};
#pragma unroll
  for (int k = 0; k < 3; ++k)
    r[k] = ((sign * d[k].x) * px + (sign * d[k].y) * py) + sign * d[k].w;
}

__global__ __launch_bounds__(kBlock) void diffbary_fwd_kernel(
    const float4* __restrict__ clip, int64_t V,
    const int32_t* __restrict__ tris, int32_t T,
    const int32_t* __restrict__ ids, int32_t W, int32_t H, int64_t N,
    float* __restrict__ barys) {
  const int64_t p = int64_t(blockIdx.x) * kBlock + threadIdx.x;
  if (p >= N) return;
  float* out = barys + 3 * p;
  PixelTriangle pt;
  if (!load_triangle(ids, tris, T, clip, V, p, pt)) {
    out[0] = out[1] = out[2] = 0.0f;
    return;
  }
  const int x = int(p % W), y = int((p / W) % H);
  float sign, r[3];
  raw_coordinates(pt, ndc_center(x, W), ndc_center(y, H), sign, r);
  const float s = (r[0] + r[1]) + r[2];
#pragma unroll
  for (int k = 0; k < 3; ++k) out[k] = nan_to_num(r[k] / s);
}

__global__ __launch_bounds__(kBlock) void diffbary_bwd_kernel(
    const float4* __restrict__ clip, int64_t V,
    const int32_t* __restrict__ tris, int32_t T,
    const int32_t* __restrict__ ids, int32_t W, int32_t H, int64_t N,
    const float* __restrict__ dbarys, float* __restrict__ dclip) {
  const int64_t p = int64_t(blockIdx.x) * kBlock + threadIdx.x;
  if (p >= N) return;
  const float db[3] = {dbarys[3 * p], dbarys[3 * p + 1], dbarys[3 * p + 2]};
  if (db[0] == 0.0f && db[1] == 0.0f && db[2] == 0.0f) return;
  PixelTriangle pt;
  if (!load_triangle(ids, tris, T, clip, V, p, pt)) return;
  const int x = int(p % W), y = int((p / W) % H);
  const float px = ndc_center(x, W), py = ndc_center(y, H);
  float sign, r[3];
  raw_coordinates(pt, px, py, sign, r);
  const float s = (r[0] + r[1]) + r[2];
  if (s == 0.0f) return;  // no finite coordinate, nothing flows back
  float dq[3], q[3], dot = 0.0f;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    q[k] = r[k] / s;
    dq[k] = finite(q[k]) ? db[k] : 0.0f;
    dot += dq[k] * (finite(q[k]) ? q[k] : 0.0f);
  }
  const Vec3 P = {px, py, 1.0f};
  Vec3 dv[3] = {{0.0f, 0.0f, 0.0f}, {0.0f, 0.0f, 0.0f}, {0.0f, 0.0f, 0.0f}};
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const float g = sign * ((dq[k] - dot) / s);
    const int a = (k + 1) % 3, b = (k + 2) % 3;
    const Vec3 ca = cross(pt.P[b], P);   // d r_k / d V_a
    const Vec3 cb = cross(P, pt.P[a]);   // d r_k / d V_b
    dv[a].x += g * ca.x; dv[a].y += g * ca.y; dv[a].w += g * ca.w;
    dv[b].x += g * cb.x; dv[b].y += g * cb.y; dv[b].w += g * cb.w;
  }
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    float* d = dclip + 4 * int64_t(pt.v[k]);
    unsafeAtomicAdd(d + 0, dv[k].x);
    unsafeAtomicAdd(d + 1, dv[k].y);
    unsafeAtomicAdd(d + 3, dv[k].w);
  }
}

inline unsigned blocks_for(int64_t n) {
  return unsigned((n + kBlock - 1) / kBlock);
}

int check_common(const float* clip, int64_t V, const int32_t* tris,
                 int32_t T, const int32_t* ids, int32_t W, int32_t H,
                 int32_t L) {
  DR_CHECK_ARG(W > 0 && H > 0 && L > 0, "image size and layers must be > 0");
  DR_CHECK_ARG(V >= 0 && T >= 0, "negative mesh size");
  DR_CHECK_ARG(ids != nullptr, "null triangle ids");
  DR_CHECK_ARG(T == 0 || (clip && tris), "null mesh pointer");
  DR_CHECK_ARG((reinterpret_cast<uintptr_t>(clip) & 15) == 0,
               "clip vertices must be 16-byte aligned [V, 4] float32");
  return DR_OK;
}

}  // namespace
}  // namespace dr

extern "C" {

int dr_diff_barycentrics_fwd(const float* clip_vertices, int64_t num_vertices,
                             const int32_t* triangles, int32_t num_triangles,
                             const int32_t* triangle_ids, int32_t width,
                             int32_t height, int32_t num_layers,
                             float* barycentrics, dr_stream_t stream) {
  const int rc = dr::check_common(clip_vertices, num_vertices, triangles,
                                  num_triangles, triangle_ids, width, height,
                                  num_layers);
  if (rc != DR_OK) return rc;
  DR_CHECK_ARG(barycentrics != nullptr, "null barycentrics");
  const int64_t n = int64_t(num_layers) * height * width;
  dr::diffbary_fwd_kernel<<<dr::blocks_for(n), dr::kBlock, 0,
                            static_cast<hipStream_t>(stream)>>>(
      reinterpret_cast<const float4*>(clip_vertices), num_vertices, triangles,
      num_triangles, triangle_ids, width, height, n, barycentrics);
  DR_CHECK_LAUNCH("diff barycentrics fwd");
  return DR_OK;
}

int dr_diff_barycentrics_bwd(const float* clip_vertices, int64_t num_vertices,
                             const int32_t* triangles, int32_t num_triangles,
                             const int32_t* triangle_ids, int32_t width,
                             int32_t height, int32_t num_layers,
                             const float* d_barycentrics,
                             float* d_clip_vertices, dr_stream_t stream) {
  const int rc = dr::check_common(clip_vertices, num_vertices, triangles,
                                  num_triangles, triangle_ids, width, height,
                                  num_layers);
  if (rc != DR_OK) return rc;
  if (num_triangles == 0) return DR_OK;
  DR_CHECK_ARG(d_barycentrics && d_clip_vertices, "null gradient pointer");
  const int64_t n = int64_t(num_layers) * height * width;
  dr::diffbary_bwd_kernel<<<dr::blocks_for(n), dr::kBlock, 0,
                            static_cast<hipStream_t>(stream)>>>(
      reinterpret_cast<const float4*>(clip_vertices), num_vertices, triangles,
      num_triangles, triangle_ids, width, height, n, d_barycentrics,
      d_clip_vertices);
  DR_CHECK_LAUNCH("diff barycentrics bwd");
  return DR_OK;
}

}  // extern "C"
