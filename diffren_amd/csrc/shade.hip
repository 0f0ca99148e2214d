// Deferred-shading inputs and the many-light Phong shader for gfx950.
//
// G-buffer (12 channels per pixel-layer): world position (3), unit normal
// (3), unit view direction eye - position (3), albedo (3); zeros on
// background.  Position, normal and albedo are barycentric interpolations of
// per-vertex attributes exactly as interpolate.hip computes them
// ((b0 a0 + b1 a1) + b2 a2); the normal is then normalised (the reference's
// transforms.l2_normalize, x * rsqrt(sum x^2 + eps), transforms.py:26-37).
//
// Phong with K point lights, the multi-light form of shaders.point_light
// (diffren/jax/utils/shaders.py:100-173), ambient counted once:
//   out = albedo * (ka + sum_i I_i * decay_i * (kd * nol_i + ks * noh_i^ns))
//   decay_i = 1 / (c + l d_i + q d_i^2), d_i = |light_i - position|
//   nol = clip(n . L, 0, 1), noh = clip(n . normalize(L + V), 0, 1)
#pragma clang fp contract(off)

#include "dr_common.h"
#include "../../include/diffren_amd.h"

namespace dr {
namespace {

constexpr int kBlock = 256;
constexpr float kNormEps = 1e-12f;

__device__ __forceinline__ float3 normalize3(float x, float y, float z) {
  const float s = rsqrtf((x * x + y * y) + z * z + kNormEps);
  return make_float3(x * s, y * s, z * s);
}

__global__ __launch_bounds__(kBlock) void gbuffer_kernel(
    const int32_t* __restrict__ ids, const float* __restrict__ barys,
    const int32_t* __restrict__ tris, int32_t T,
    const float* __restrict__ positions, const float* __restrict__ normals,
    const float* __restrict__ albedo, int32_t V, const float* __restrict__ eye,
    int64_t N, float* __restrict__ out) {
  const int64_t p = int64_t(blockIdx.x) * kBlock + threadIdx.x;
  if (p >= N) return;
  float g[12];
#pragma unroll
  for (int c = 0; c < 12; ++c) g[c] = 0.0f;
  const int t = ids[p];
  if (t >= 0 && t < T) {
    const int v0 = tris[3 * int64_t(t)], v1 = tris[3 * int64_t(t) + 1],
              v2 = tris[3 * int64_t(t) + 2];
    if (unsigned(v0) < unsigned(V) && unsigned(v1) < unsigned(V) &&
        unsigned(v2) < unsigned(V)) {
      const float b0 = barys[3 * p], b1 = barys[3 * p + 1],
                  b2 = barys[3 * p + 2];
      float pos[3], nrm[3], alb[3];
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        pos[c] = (b0 * positions[3 * v0 + c] + b1 * positions[3 * v1 + c]) +
                 b2 * positions[3 * v2 + c];
        nrm[c] = (b0 * normals[3 * v0 + c] + b1 * normals[3 * v1 + c]) +
                 b2 * normals[3 * v2 + c];
        alb[c] = (b0 * albedo[3 * v0 + c] + b1 * albedo[3 * v1 + c]) +
                 b2 * albedo[3 * v2 + c];
      }
      const float3 n = normalize3(nrm[0], nrm[1], nrm[2]);
      const float3 vd =
          normalize3(eye[0] - pos[0], eye[1] - pos[1], eye[2] - pos[2]);
      g[0] = pos[0]; g[1] = pos[1]; g[2] = pos[2];
      g[3] = n.x; g[4] = n.y; g[5] = n.z;
      g[6] = vd.x; g[7] = vd.y; g[8] = vd.z;
      g[9] = alb[0]; g[10] = alb[1]; g[11] = alb[2];
    }
  }
  float4* o = reinterpret_cast<float4*>(out + 12 * p);
  o[0] = make_float4(g[0], g[1], g[2], g[3]);
  o[1] = make_float4(g[4], g[5], g[6], g[7]);
  o[2] = make_float4(g[8], g[9], g[10], g[11]);
}

// Lights staged in LDS: [K][4] = position xyz, intensity.
__global__ __launch_bounds__(kBlock) void phong_kernel(
    const float* __restrict__ gbuf, int64_t N, const float* __restrict__ lights,
    int K, dr_phong_params prm, float* __restrict__ out) {
  extern __shared__ float4 sl[];
  for (int i = threadIdx.x; i < K; i += kBlock)
    sl[i] = reinterpret_cast<const float4*>(lights)[i];
  __syncthreads();
  const int64_t p = int64_t(blockIdx.x) * kBlock + threadIdx.x;
  if (p >= N) return;
  const float4* g4 = reinterpret_cast<const float4*>(gbuf + 12 * p);
  const float4 a = g4[0], b = g4[1], c = g4[2];
  const float px = a.x, py = a.y, pz = a.z;
  const float nx = a.w, ny = b.x, nz = b.y;
  const float vx = b.z, vy = b.w, vz = c.x;
  float lum = 0.0f;
  for (int i = 0; i < K; ++i) {
    const float4 L = sl[i];
    const float lx = L.x - px, ly = L.y - py, lz = L.z - pz;
    const float d2 = (lx * lx + ly * ly) + lz * lz;
    const float d = sqrtf(d2);
    const float inv = rsqrtf(d2 + kNormEps);
    const float ux = lx * inv, uy = ly * inv, uz = lz * inv;
    const float nol =
        fminf(fmaxf((nx * ux + ny * uy) + nz * uz, 0.0f), 1.0f);
    const float3 h = normalize3(ux + vx, uy + vy, uz + vz);
    const float noh =
        fminf(fmaxf((nx * h.x + ny * h.y) + nz * h.z, 0.0f), 1.0f);
    const float spec = prm.ks * powf(noh, prm.ns);
    const float decay =
        1.0f / ((prm.attenuation_constant + prm.attenuation_linear * d) +
                prm.attenuation_quadratic * d2);
    lum += L.w * decay * (prm.kd * nol + spec);
  }
  const float s = prm.ka + lum;
  float* o = out + 3 * p;
  o[0] = c.y * s;
  o[1] = c.z * s;
  o[2] = c.w * s;
}

}  // namespace
}  // namespace dr

extern "C" {

int dr_gbuffer(const int32_t* triangle_ids, const float* barycentrics,
               const int32_t* triangles, int32_t num_triangles,
               const float* positions, const float* normals,
               const float* albedo, int32_t num_vertices, const float* eye,
               int64_t num_pixels, float* gbuffer, dr_stream_t stream) {
  DR_CHECK_ARG(num_pixels >= 0 && num_triangles >= 0 && num_vertices >= 0,
               "negative size");
  if (num_pixels == 0) return DR_OK;
  DR_CHECK_ARG(triangle_ids && barycentrics && gbuffer && eye, "null pointer");
  DR_CHECK_ARG(num_triangles == 0 ||
                   (triangles && positions && normals && albedo),
               "null mesh pointer");
  DR_CHECK_ARG((reinterpret_cast<uintptr_t>(gbuffer) & 15) == 0,
               "gbuffer must be 16-byte aligned");
  dr::gbuffer_kernel<<<unsigned((num_pixels + dr::kBlock - 1) / dr::kBlock),
                       dr::kBlock, 0, static_cast<hipStream_t>(stream)>>>(
      triangle_ids, barycentrics, triangles, num_triangles, positions,
      normals, albedo, num_vertices, eye, num_pixels, gbuffer);
  DR_CHECK_LAUNCH("gbuffer");
  return DR_OK;
}

int dr_phong_shade(const float* gbuffer, int64_t num_pixels,
                   const float* lights, int32_t num_lights,
                   const dr_phong_params* params, float* out,
                   dr_stream_t stream) {
  DR_CHECK_ARG(num_pixels >= 0, "negative size");
  DR_CHECK_ARG(num_lights >= 0 && num_lights <= 2048,
               "num_lights must be in [0, 2048]");
  DR_CHECK_ARG(params != nullptr, "params is null");
  if (num_pixels == 0) return DR_OK;
  DR_CHECK_ARG(gbuffer && out && (num_lights == 0 || lights), "null pointer");
  DR_CHECK_ARG((reinterpret_cast<uintptr_t>(gbuffer) & 15) == 0 &&
                   (reinterpret_cast<uintptr_t>(lights) & 15) == 0,
               "gbuffer and lights must be 16-byte aligned");
  dr::phong_kernel<<<unsigned((num_pixels + dr::kBlock - 1) / dr::kBlock),
                     dr::kBlock, size_t(num_lights) * 16 + 16,
                     static_cast<hipStream_t>(stream)>>>(
      gbuffer, num_pixels, lights, num_lights, *params, out);
  DR_CHECK_LAUNCH("phong");
  return DR_OK;
}

}  // extern "C"
