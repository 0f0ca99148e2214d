// Fused, batched rasterize-then-splat render for gfx950: forward, L2 loss
// gradient and backward over many (mesh, camera, resolution) jobs per launch.
//
// Per job the inputs are the rasterizer's outputs (triangle ids, z, and
// barycentrics, layer-major [L][H][W]), the clip-space vertices, the
// triangles and the colours to splat: per-vertex colours under an identity
// shader (render.py:123-132 with shading_function = lambda a: a["colors"]),
// or the shaded pixels [L][H][W][C] of any shading function
// (splat_shaded_pixels' shaded_pixels, splat.py:178-229).  What the
// reference computes as separate XLA passes -- interpolate the colours
// (interpolate.py:24-63), interpolate clip xyzw with background (0,0,1,1)
// and map to the viewport (splat.py:209-221), splat_at_pixel_centers
// (splat.py:26-175) -- is done here without writing any intermediate frame:
//
//  target pass (one block per 32x8 tile of output pixels of one job):
//    stage the tile plus a 1-pixel halo, all layers, in LDS: fg mask,
//    viewport x, y, z and the interpolated colour, computed from ids,
//    barycentrics and vertex gathers; then pull-splat every output pixel from
//    its 3x3 neighbourhood (same arithmetic and order as splat.hip), write the
//    image and, when a gradient is wanted, back-propagate d image (given, or
//    2 (image - target) * loss_scale for an L2 loss) through `over` and the
//    normalisation into G[pixel][class] = (d acc colour, d acc weight).
//  source pass (one thread per source pixel, all layers): re-interpolate the
//    pixel's colour and position, pull G from the 9 pixels it splats onto,
//    form d colour and d viewport xy, map d xy to d clip (x, y, w), and
//    scatter b_k-weighted gradients to the triangle's three vertices --
//    d colours[v] and d vertices[v] = M^T d clip (or d clip itself when no
//    camera is given); with shaded pixels d colour is the pixel-layer's
//    d shaded, written once, and only positions are scattered.  Float
//    atomics: summation order is not fixed (a fixed-point mode fixes it).
//
// Viewport z is taken from the rasterizer's z buffer, ((z + 1) * 1) * 0.5:
// the rasterizer computes z = ((b0 z0 + b1 z1) + b2 z2) / ((b0 w0 + b1 w1) +
// b2 w2) (rasterize_triangles_impl.cc:87-89), the same expression the
// interpolated clip buffer divides out (splat.py:218), so it is the same
// value bit for bit.
#pragma clang fp contract(off)

#include <cmath>

#include "dr_common.h"
#include "../../include/diffren_amd.h"

namespace dr {
namespace {

// Output tiles of both passes: 32x8 pixels, 256 threads.  A target tile
// stages its 34x10 region (tile plus a 1-pixel halo) -- 1.33 staged pixels
// per output pixel, against 1.55 for 64x4 tiles, whose halo splats were
// gathered and interpolated by two blocks each; a source block's four 16x4
// wave patches form a 2x2 square, so its waves share more vertices.
constexpr int kBX = 32, kBY = 8;
constexpr int kRX = kBX + 2, kRY = kBY + 2, kRegion = kRX * kRY;
constexpr int kMaxLayers = 8;

// Per-pixel gradient scratch written by the target pass for the source pass,
// kRecord floats per pixel in two planes of the job's buffer:
//  * hot plane [P][kHot]: 16 bytes holding the pixel's 9 surface indices
//    (one signed byte per tap, tap order t = 3 (u + 1) + (v + 1)), then the
//    "surface" class (1) -- what nearly every fragment reads;
//  * cold plane [P][kCold] after it: classes 0 and 2, written only where the
//    class received splat weight (only then does a source read it).
// A class is kStride floats: d acc colour [C], d acc weight, padding to a
// float4 multiple.
template <int C>
struct Scratch {
  static constexpr int kStride = ((C + 1) + 3) & ~3;
  static constexpr int kHot = 4 + kStride;
  static constexpr int kCold = 2 * kStride;
  static constexpr int kRecord = kHot + kCold;  // floats
};

// Class k of pixel p's record in a job's scratch of P pixels.
template <int C>
__device__ __forceinline__ float* scratch_class(float* base, int64_t P,
                                                int64_t p, int k) {
  using S = Scratch<C>;
  return k == 1 ? base + p * S::kHot + 4
                : base + P * S::kHot + p * S::kCold + (k == 0 ? 0 : S::kStride);
}

// Orders a wave's LDS writes before its lanes' later LDS reads (release /
// acquire at wavefront scope around the wave barrier).
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ int find_job(const dr_splat_job* jobs, int n,
                                        int64_t tile) {
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (jobs[mid].tile_begin <= tile) lo = mid; else hi = mid - 1;
  }
  return lo;
}

struct Corner {
  int v[3];
  float b[3];
};

// Triangle corners of pixel-layer q; false on background or bad indices.
// The three indices and three barycentrics are loaded together (global-
// memory loads; see ldg) before any of them is checked.
__device__ __forceinline__ bool corners(const dr_splat_job& J, int64_t q,
                                        Corner& k) {
  const int t = ldg(J.triangle_ids + q);
  if (t < 0 || t >= J.num_triangles) return false;
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    k.v[i] = ldg(J.triangles + 3 * int64_t(t) + i);
    k.b[i] = ldg(J.barycentrics + 3 * q + i);
  }
  const unsigned nv = unsigned(J.num_vertices);
  return unsigned(k.v[0]) < nv && unsigned(k.v[1]) < nv &&
         unsigned(k.v[2]) < nv;
}

__device__ __forceinline__ float interp3(const Corner& k, float a0, float a1,
                                         float a2) {
  return (k.b[0] * a0 + k.b[1] * a1) + k.b[2] * a2;
}

// Offset from a splat to the centre of its own pixel, floor(x) + 0.5 - x,
// where x = ((c / w + 1) * size) * 0.5 is its viewport coordinate
// (splat.py:214-221, 103-104).  The forward (target pass) forms it exactly
// as the reference does, in float32 and in that order, so the image is the
// reference's float32 image.
__device__ __forceinline__ float reference_offset(float c, float w, int size) {
  const float x = ((c / w + 1.0f) * float(size)) * 0.5f;
  return (floorf(x) + 0.5f) - x;
}

// The same offset for the backward.  In rasterize-then-splat x sits on the
// pixel centre up to rounding, so reference_offset is pure rounding noise of
// ~6e-5 px at x ~ 1000 -- yet it is what the tap weights' position
// gradients are evaluated at.  The source pass computes it as
//   -(a c + k w) / w,  a = size / 2, k = a - floor(x) - 0.5 (both exact),
// with the product k w split exactly by an fma (two-product), so the
// cancellation happens without rounding and the offset is accurate to a few
// ulps of itself: the gradient is then the derivative at the exact splat
// position, as the float64 oracle evaluates it.
__device__ __forceinline__ float center_offset(float c, float w, int size) {
  const float a = 0.5f * float(size);
  const float x = ((c / w + 1.0f) * float(size)) * 0.5f;
  const float k = (a - floorf(x)) - 0.5f;
  const float p = k * w;
  const float e = fmaf(k, w, -p);
  const float num = fmaf(a, c, p) + e;
  return -(num / w);
}

// Interpolated clip x, y, w and the splat's pixel-centre offsets.
// Clip-space vertex v: the stored clip position, or (without one) the world
// position transformed as dr_transform_homogeneous does, rows x, y, w only
// (the splat takes z from the rasterizer's z buffer).
__device__ __forceinline__ float4 clip_vertex(const dr_splat_job& J, int v) {
  if (J.clip_vertices)
    return ldg4(J.clip_vertices + 4 * int64_t(v));
  const float* p = J.positions + 3 * int64_t(v);
  const float x = ldg(p), y = ldg(p + 1), z = ldg(p + 2);
  const float* M = J.camera;
  float4 r;
  r.x = ((ldg(M) * x + ldg(M + 1) * y) + ldg(M + 2) * z) + ldg(M + 3) * 1.0f;
  r.y = ((ldg(M + 4) * x + ldg(M + 5) * y) + ldg(M + 6) * z) +
        ldg(M + 7) * 1.0f;
  r.z = 0.0f;
  r.w = ((ldg(M + 12) * x + ldg(M + 13) * y) + ldg(M + 14) * z) +
        ldg(M + 15) * 1.0f;
  return r;
}

// kForward: the reference's float32 offsets (target pass); otherwise the
// compensated ones (source pass).
template <bool kForward>
__device__ __forceinline__ void splat_position(const dr_splat_job& J,
                                               const Corner& k, float& ox,
                                               float& oy, float& cx,
                                               float& cy, float& cw) {
  const float4 p0 = clip_vertex(J, k.v[0]), p1 = clip_vertex(J, k.v[1]),
               p2 = clip_vertex(J, k.v[2]);
  cx = interp3(k, p0.x, p1.x, p2.x);
  cy = interp3(k, p0.y, p1.y, p2.y);
  cw = interp3(k, p0.w, p1.w, p2.w);
  if (kForward) {
    ox = reference_offset(cx, cw, J.width);
    oy = reference_offset(cy, cw, J.height);
  } else {
    ox = center_offset(cx, cw, J.width);
    oy = center_offset(cy, cw, J.height);
  }
}

__device__ __forceinline__ float viewport_z(float z) {
  return ((z + 1.0f) * 1.0f) * 0.5f;
}

// Viewport z of pixel-layer q; background is the clear depth's, 1.  The z
// buffer is read only under a triangle (the rasterizer's sparse mode leaves
// background z unwritten).
__device__ __forceinline__ float layer_z(const dr_splat_job& J, int64_t q) {
  return ldg(J.triangle_ids + q) >= 0 ? viewport_z(ldg(J.z_buffer + q)) : 1.0f;
}

__device__ __forceinline__ int surface_index(const float* zS, const float* zT,
                                             int L) {
  float bc = fabsf(zS[0] - zT[0]);
  int ic = 0;
  float bt = fabsf(zT[0] - zS[0]);
  int it = 0;
  for (int l = 1; l < L; ++l) {
    const float dc = fabsf(zS[l] - zT[0]);
    if (dc < bc) { bc = dc; ic = l; }
    const float dt = fabsf(zT[l] - zS[0]);
    if (dt < bt) { bt = dt; it = l; }
  }
  return bt < bc ? -it : ic;
}

__device__ __forceinline__ int layer_class(int surface, int l) {
  return surface > l ? 0 : (surface == l ? 1 : 2);
}

// Tap weight for offset d = (ox + u, oy + v) (splat.py:106-115; the
// foreground mask is 1 for every staged splat).
__device__ __forceinline__ float tap_weight(float coef, float ox, float oy,
                                            int u, int v, float& dx,
                                            float& dy) {
  dx = ox + float(u);
  dy = oy + float(v);
  const float d2 = dx * dx + dy * dy;
  return (1.0f * coef) * expf(-2.0f * d2);
}

__device__ __forceinline__ float norm_scale(float w) {
  return 1.0f / (fmaxf(w - 1.0f, 0.0f) + 1.0f);
}

// Stores a C-channel row to global memory (16-byte stores when C % 4 == 0:
// rows are then 16-byte aligned); v == nullptr stores zeros.
template <int C>
__device__ __forceinline__ void store_row(float* dst, const float* v) {
#pragma unroll
  for (int c = 0; c < C; ++c) {
    if constexpr (C % 4 == 0) {
      if (c % 4 == 0)
        gst4(dst + c, v ? make_float4(v[c], v[c + 1], v[c + 2], v[c + 3])
                        : make_float4(0.0f, 0.0f, 0.0f, 0.0f));
    } else {
      gst(dst + c, v ? v[c] : 0.0f);
    }
  }
}

// Loads a C-channel row from global memory (16-byte loads when C % 4 == 0).
template <int C>
__device__ __forceinline__ void load_row(const float* src, float* v) {
#pragma unroll
  for (int c = 0; c < C; ++c) {
    if constexpr (C % 4 == 0) {
      if (c % 4 == 0) {
        const float4 q = ldg4(src + c);
        v[c] = q.x; v[c + 1] = q.y; v[c + 2] = q.z; v[c + 3] = q.w;
      }
    } else {
      v[c] = ldg(src + c);
    }
  }
}

// LDS: for all layers, the region's triangle ids st[L][kRegion] and viewport
// z sz[L][kRegion] (the surface indices need every layer's z at once); then,
// one layer at a time, pixel-centre offsets x, y and colour[C] of that layer
// (x = NaN marks background / outside the image: the fg mask is 1 exactly
// where the rasterizer stored a triangle).  Staging a single layer's splats
// keeps the block's LDS small enough for seven resident blocks per CU.
// Six resident waves per SIMD fit without spilling for up to 4 channels.
template <int C, bool kAccum>
__global__ __launch_bounds__(kBX* kBY)
__attribute__((amdgpu_waves_per_eu(C <= 4 ? 6 : (C <= 12 ? 4 : (C <= 16 ? 3 : 2)))))
void target_kernel(
    const dr_splat_job* __restrict__ jobs, int num_jobs, int L, float coef,
    int grad_mode, float loss_scale, float* __restrict__ loss_partials) {
  // float4-aligned, so a C % 4 == 0 colour row is read and written as
  // 16-byte LDS accesses (lanes at a 16-byte stride with 4-byte accesses
  // meet the same banks several times per instruction).
  extern __shared__ float4 lds4[];
  float* lds = reinterpret_cast<float*>(lds4);
  const int R = kRegion;
  int* st = reinterpret_cast<int*>(lds);
  float* sz = lds + L * R;
  float* sx = sz + L * R;
  float* sy = sx + R;
  float* sc = sy + R;  // [R][C]
  __shared__ int layer_live[kMaxLayers];

  const int64_t tile = xcd_block<32>();
  const dr_splat_job& J = jobs[find_job(jobs, num_jobs, tile)];
  const int W = J.width, H = J.height;
  const int ntx = (W + kBX - 1) / kBX;
  const int local = int(tile - J.tile_begin);
  const int x0 = (local % ntx) * kBX, y0 = (local / ntx) * kBY;
  const int64_t P = int64_t(W) * H;
  const int tid = threadIdx.y * kBX + threadIdx.x;
  constexpr int kThreads = kBX * kBY;
  // A job without the inputs of the requested gradient (target for mode 1,
  // d_image for mode 2, grad_scratch for both) gets its image only.
  const int mode =
      grad_mode == 1 ? (J.target && J.grad_scratch ? 1 : 0)
                     : (grad_mode == 2 && J.d_image && J.grad_scratch ? 2 : 0);

  // Ids and z of every layer (z buffer read only under a triangle).  A
  // pixel's layers are a prefix of fragments (the rasterizer fills them
  // front to back), so layer l's id is read only where layer l - 1 holds a
  // fragment: background pixels cost one load.
  if (tid < kMaxLayers) layer_live[tid] = 0;
  __syncthreads();
  for (int r = tid; r < R; r += kThreads) {
    const int px = x0 - 1 + r % kRX, py = y0 - 1 + r / kRX;
    bool more = px >= 0 && px < W && py >= 0 && py < H;
    const int64_t pix = int64_t(py) * W + px;
    for (int l = 0; l < L; ++l) {
      int t = -1;
      float z = 1.0f;
      if (more) {
        const int64_t q = l * P + pix;
        t = ldg(J.triangle_ids + q);
        if (t >= 0 && t < J.num_triangles) z = viewport_z(ldg(J.z_buffer + q));
        else t = -1;
        more = t >= 0;
      }
      st[l * R + r] = t;
      sz[l * R + r] = z;
      if (t >= 0) layer_live[l] = 1;
    }
  }
  __syncthreads();

  const int x = x0 + threadIdx.x, y = y0 + threadIdx.y;
  const bool in_image = x < W && y < H;
  const int rT = (threadIdx.y + 1) * kRX + threadIdx.x + 1;
  // The 9 surface indices, one signed byte each (tap t in byte t % 4 of
  // word t / 4): the same bytes the gradient scratch records.  A region
  // with no fragment (most of a frame) splats nothing and writes no
  // gradient record, so it needs none.
  // Live layers form a prefix (a pixel has a layer-l fragment only if it
  // has one at every layer before l).  Past the first empty layer every
  // region pixel has z = 1 at every layer, and further such layers can never
  // win the surface index's strict-< argmins against it, so the index is
  // computed over the live layers plus one empty layer.
  int live = 0;
  for (int l = 0; l < L; ++l) live += layer_live[l] != 0;
  const bool region_live = live > 0;
  const int Ls = min(L, live + 1);
  uint32_t sw[3] = {0u, 0u, 0u};
  if (in_image && region_live) {
    float zT[kMaxLayers], zS[kMaxLayers];
    for (int l = 0; l < Ls; ++l) zT[l] = sz[l * R + rT];
#pragma unroll
    for (int t = 0; t < 9; ++t) {
      const int u = t / 3 - 1, v = t % 3 - 1;
      const int rS = rT - v * kRX - u;
      for (int l = 0; l < Ls; ++l) zS[l] = sz[l * R + rS];
      sw[t >> 2] |= (uint32_t(surface_index(zS, zT, Ls)) & 0xffu)
                    << (8 * (t & 3));
    }
  }
  float acc[3][C + 1];
#pragma unroll
  for (int k = 0; k < 3; ++k)
#pragma unroll
    for (int c = 0; c <= C; ++c) acc[k][c] = 0.0f;

  for (int l = 0; l < L; ++l) {
    if (!layer_live[l]) continue;  // block-uniform
    // Stage layer l's splats; each thread issues the loads of its items
    // stage by stage (triangles + barycentrics, then vertex gathers) so its
    // dependent load chains are in flight together.
    constexpr int kChunk = (kRegion + kThreads - 1) / kThreads;
    Corner k[kChunk];
    bool ok[kChunk];
#pragma unroll
    for (int j = 0; j < kChunk; ++j) {
      const int i = tid + j * kThreads;
      ok[j] = false;
      if (i < R) {
        const int t = st[l * R + i];
        if (t >= 0) {
          const int px = x0 - 1 + i % kRX, py = y0 - 1 + i / kRX;
          const int64_t q = l * P + int64_t(py) * W + px;
          ok[j] = true;
#pragma unroll
          for (int e = 0; e < 3; ++e) {
            k[j].v[e] = ldg(J.triangles + 3 * int64_t(t) + e);
            k[j].b[e] = ldg(J.barycentrics + 3 * q + e);
          }
        }
      }
    }
#pragma unroll
    for (int j = 0; j < kChunk; ++j) {
      const int i = tid + j * kThreads;
      if (i >= R) break;
      float px = __int_as_float(0x7fc00000), py = 0.0f;
      float col[C];
#pragma unroll
      for (int c = 0; c < C; ++c) col[c] = 0.0f;
      bool good = ok[j];
#pragma unroll
      for (int e = 0; e < 3; ++e)
        good = good && unsigned(k[j].v[e]) < unsigned(J.num_vertices);
      if (good) {
        float cx, cy, cw;
        splat_position<true>(J, k[j], px, py, cx, cy, cw);
        if (J.shaded) {
          // The shading function's output for this pixel-layer.
          const int qx = x0 - 1 + i % kRX, qy = y0 - 1 + i / kRX;
          const float* sp = J.shaded + (l * P + int64_t(qy) * W + qx) * C;
#pragma unroll
          for (int c = 0; c < C; ++c) {
            if constexpr (C % 4 == 0) {
              if (c % 4 == 0) {
                const float4 q4 = ldg4(sp + c);
                col[c] = q4.x; col[c + 1] = q4.y; col[c + 2] = q4.z;
                col[c + 3] = q4.w;
              }
            } else {
              col[c] = ldg(sp + c);
            }
          }
        } else {
          const float* a0 = J.colors + int64_t(k[j].v[0]) * C;
          const float* a1 = J.colors + int64_t(k[j].v[1]) * C;
          const float* a2 = J.colors + int64_t(k[j].v[2]) * C;
#pragma unroll
          for (int c = 0; c < C; ++c)
            col[c] = interp3(k[j], ldg(a0 + c), ldg(a1 + c), ldg(a2 + c));
        }
      }
      sx[i] = px;
      sy[i] = py;
#pragma unroll
      for (int c = 0; c < C; ++c) {
        if constexpr (C % 4 == 0) {
          if (c % 4 == 0)
            *reinterpret_cast<float4*>(sc + i * C + c) =
                make_float4(col[c], col[c + 1], col[c + 2], col[c + 3]);
        } else {
          sc[i * C + c] = col[c];
        }
      }
    }
    __syncthreads();
    if (in_image) {
#pragma unroll
      for (int t = 0; t < 9; ++t) {
        const int u = t / 3 - 1, v = t % 3 - 1;
        const int i = rT - v * kRX - u;
        // Background or outside the image (x NaN): weight 0 -- a select,
        // not a branch (its colour row is 0, so the adds change nothing).
        const float px = sx[i];
        float dx, dy;
        const float w0 = tap_weight(coef, px, sy[i], u, v, dx, dy);
        const float w = px == px ? w0 : 0.0f;
        const int kc = layer_class(
            int(int8_t(sw[t >> 2] >> (8 * (t & 3)))), l);
        float cv[C];
#pragma unroll
        for (int c = 0; c < C; ++c) {
          if constexpr (C % 4 == 0) {
            if (c % 4 == 0) {
              const float4 q =
                  *reinterpret_cast<const float4*>(sc + i * C + c);
              cv[c] = q.x; cv[c + 1] = q.y; cv[c + 2] = q.z; cv[c + 3] = q.w;
            }
          } else {
            cv[c] = sc[i * C + c];
          }
        }
        // The reference's masked accumulation (splat.py:147-160) as
        // fma(tap, 0/1 class mask, acc): the bits of an add of the tap into
        // its class and of a zero into the others (sums start at +0 and are
        // never -0), in one instruction per class and channel.
        const float m0 = kc == 0 ? 1.0f : 0.0f;
        const float m1 = kc == 1 ? 1.0f : 0.0f;
        const float m2 = kc == 2 ? 1.0f : 0.0f;
#pragma unroll
        for (int c = 0; c < C; ++c) {
          const float wc = w * cv[c];
          acc[0][c] = fmaf(wc, m0, acc[0][c]);
          acc[1][c] = fmaf(wc, m1, acc[1][c]);
          acc[2][c] = fmaf(wc, m2, acc[2][c]);
        }
        acc[0][C] = fmaf(w, m0, acc[0][C]);
        acc[1][C] = fmaf(w, m1, acc[1][C]);
        acc[2][C] = fmaf(w, m2, acc[2][C]);
      }
    }
    __syncthreads();  // the next layer's staging overwrites sx, sy, sc
  }

  float partial = 0.0f;
  if (in_image) {
    float s[3], n[3][C];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      s[k] = norm_scale(acc[k][C]);
#pragma unroll
      for (int c = 0; c < C; ++c) n[k][c] = acc[k][c] * s[k];
    }
    const float a0 = n[0][C - 1], a1 = n[1][C - 1];
    const int64_t p = int64_t(y) * W + x;
    if (kAccum && (J.accum || J.norm || J.accum_weights)) {
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        if (J.accum_weights) gst(J.accum_weights + k * P + p, acc[k][C]);
        if (J.accum) store_row<C>(J.accum + (k * P + p) * C, acc[k]);
        if (J.norm) store_row<C>(J.norm + (k * P + p) * C, n[k]);
      }
    }
    float o[C], g[C];
#pragma unroll
    for (int c = 0; c < C; ++c) {
      o[c] = n[2][c];
      o[c] = n[1][c] + (1.0f - a1) * o[c];
      o[c] = n[0][c] + (1.0f - a0) * o[c];
    }
    if (J.image) store_row<C>(J.image + p * C, o);
    if (mode == 1) {
      float t[C];
      load_row<C>(J.target + p * C, t);
#pragma unroll
      for (int c = 0; c < C; ++c) {
        const float d = o[c] - t[c];
        partial += d * d;
        g[c] = (2.0f * d) * loss_scale;
      }
    } else if (mode == 2) {
      load_row<C>(J.d_image + p * C, g);
    }
    // A pixel that received no splat weight is never read by the source
    // pass (every foreground source adds a positive weight to each of its
    // 9 targets), so its gradient scratch is left unwritten.
    if (mode != 0 && (acc[0][C] + acc[1][C]) + acc[2][C] > 0.0f) {
      float dn[3][C];
      float da0 = 0.0f, da1 = 0.0f;
#pragma unroll
      for (int c = 0; c < C; ++c) {
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
      constexpr int GS = Scratch<C>::kStride;
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        if (k != 1 && !(acc[k][C] > 0.0f)) continue;  // never read
        float* gp = scratch_class<C>(J.grad_scratch, P, p, k);
        const float ds = acc[k][C] > 1.0f ? -(s[k] * s[k]) : 0.0f;
        float rec[GS];
        float dw = 0.0f;
#pragma unroll
        for (int c = 0; c < C; ++c) {
          rec[c] = dn[k][c] * s[k];
          dw += dn[k][c] * acc[k][c];
        }
        rec[C] = dw * ds;
#pragma unroll
        for (int c = C + 1; c < GS; ++c) rec[c] = 0.0f;
#pragma unroll
        for (int c = 0; c < GS; c += 4)
          gst4(gp + c, make_float4(rec[c], rec[c + 1], rec[c + 2],
                                   rec[c + 3]));
      }
      gstu4(reinterpret_cast<uint32_t*>(J.grad_scratch +
                                        p * Scratch<C>::kHot),
            make_uint4(sw[0], sw[1], sw[2], 0u));
    }
  }
  if (grad_mode == 1) {
    // Tile partial of the loss: wave shuffles, then the block's waves.
    __shared__ float wave_sum[kBX * kBY / 64];
    for (int off = 32; off > 0; off >>= 1) partial += __shfl_xor(partial, off);
    if (__lane_id() == 0) wave_sum[tid >> 6] = partial;
    __syncthreads();
    if (tid == 0) {
      float t = 0.0f;
      for (int w = 0; w < kBX * kBY / 64; ++w) t += wave_sum[w];
      loss_partials[tile] = t * loss_scale;
    }
  }
}

// Sums the per-tile loss partials in a fixed order (one block) into *loss.
__global__ __launch_bounds__(1024) void loss_reduce_kernel(
    const float* __restrict__ partials, int64_t n, float* __restrict__ loss) {
  __shared__ float red[1024];
  float t = 0.0f;
  for (int64_t i = threadIdx.x; i < n; i += 1024) t += partials[i];
  red[threadIdx.x] = t;
  __syncthreads();
  for (int d = 512; d > 0; d >>= 1) {
    if (int(threadIdx.x) < d) red[threadIdx.x] += red[threadIdx.x + d];
    __syncthreads();
  }
  // An atomic add, because job groups on concurrent streams share *loss;
  // with two groups the result does not depend on their order.
  if (threadIdx.x == 0) unsafeAtomicAdd(loss, red[0]);
}

// Gradient record of one vertex: [R] floats, R = record_size(C):
// cols 0..3 d position (x, y, z, 0) -- or d clip (x, y, 0, w) when the job
// has no camera -- and cols 4 .. 4 + C - 1 d colour.  Jobs with shaded
// pixels scatter positions only: R = 4.  kBase is the number of leading
// columns a fragment contributes to.
template <int C, bool kShaded = false>
struct Record {
  static constexpr int kBase = kShaded ? 4 : 4 + C;
  static constexpr int kSize =
      kShaded ? 4 : ((4 + C) <= 8 ? 8 : ((4 + C) <= 16 ? 16 : 32));
};



// Source pass.  Float atomics to the vertex gradient records are the cost of
// this pass (they run at a fixed rate of grouped operations), so it issues as
// few as it can:
//  * each wave covers a 16x4 patch of pixels, whose fragments reference each
//    mesh vertex ~4 times on average (neighbouring pixels' triangles share
//    vertices);
//  * neighbouring pixels of a row often lie on the same triangle (an active
//    wave's ~62 fragments lie on ~36 triangles at 1M triangles / 1024^2):
//    the corner records of a run of such lanes are summed into the run's
//    first lane by a segmented sum over the row (DPP), and only that lane
//    goes on -- this also takes away the same-address LDS atomics of lanes
//    on one triangle;
//  * the wave finds repeated vertices among the remaining three vertex
//    records per lane with a small LDS hash table, keeps one LDS row per
//    distinct vertex (the first occurrence writes it, repeats add into it)
//    and flushes only those rows;
//  * the flush puts R consecutive lanes on one record, because grouped
//    atomics run ~7x faster than one scattered address per lane.
constexpr int kHashSlots = 256;   // > 3 x 64 entries per wave and layer
// Distinct vertices a wave keeps in LDS per layer (its 192 entries reference
// about a quarter as many; further distinct vertices are flushed directly).
constexpr int kRows = 96;

// A fragment's record for its corner vertex v is b * base: base holds
// (d position or d clip [4], d colour [C]); columns past kBase are zero.
// Without the run merge (wide records) it is formed where it is used, so
// the three corners' records are never all live in registers.
template <int kBase>
__device__ __forceinline__ float record_entry(float b, const float* base,
                                              int c) {
  return c < kBase ? b * base[c] : 0.0f;
}

// x of the lane kD places to the right in this lane's row of 16 lanes (DPP
// row_shl:kD), or `fallback` past the row's end.
template <int kD>
__device__ __forceinline__ int row_from_right(int x, int fallback) {
  return __builtin_amdgcn_update_dpp(fallback, x, 0x100 + kD, 0xf, 0xf,
                                     false);
}
template <int kD>
__device__ __forceinline__ float row_from_right(float x) {
  return __int_as_float(row_from_right<kD>(__float_as_int(x), 0));
}

// One doubling step of a segmented suffix sum over rows of 16 lanes: a lane
// whose run (seg: the lane index where the run starts) continues kD lanes to
// the right adds that lane's partial sums, which then cover up to 2 kD lanes
// of the run.  Returns false when no lane of the wave takes anything (then
// no run is longer than kD and later steps would take nothing either).
template <int kD, int N>
__device__ __forceinline__ bool run_sum_step(int seg, float (&x)[3][N]) {
  const bool take = row_from_right<kD>(seg, -1) == seg;
  if (__ballot(take) == 0) return false;
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int c = 0; c < N; ++c) {
      const float r = row_from_right<kD>(x[i][c]);
      x[i][c] += take ? r : 0.0f;
    }
  return true;
}

// Adds v to a gradient entry: a float atomic, or -- in the deterministic
// mode (fixed_shift > 0) -- an exact 64-bit integer atomic of v * 2^shift
// rounded to the nearest integer, so the sum does not depend on the order
// in which the additions arrive (the entry's memory holds an int64).
__device__ __forceinline__ void grad_add(float* p, int64_t i, float v,
                                         int fixed_shift, float fixed_scale) {
  if (fixed_shift > 0) {
    const long long q = llrintf(v * fixed_scale);
    gatomic_add_u64(reinterpret_cast<unsigned long long*>(p) + i,
                    static_cast<unsigned long long>(q));
  } else {
    gatomic_add(p + i, v);
  }
}

// Adds a corner's record (entry(c), c < R) straight to vertex v's gradient.
template <int R, typename Entry>
__device__ __forceinline__ void flush_record(float* grads, int v,
                                             const Entry& entry,
                                             int fixed_shift,
                                             float fixed_scale) {
  if (grads == nullptr) return;
#pragma unroll
  for (int c = 0; c < R; ++c) {
    const float e = entry(c);
    if (e != 0.0f)
      grad_add(grads, int64_t(v) * R + c, e, fixed_shift, fixed_scale);
  }
}

// The wave's 16x4 sources splat onto an 18x6 patch of targets.  Their
// gradient-scratch records are staged in LDS once, coalesced, before any
// dependent load: the 9 surface bytes of every record, and (for C <= 4) the
// "surface" class of G, which nearly every layer-0 fragment reads; other
// classes and layers read G from global memory.  The stage is dead once
// layer 0 has been differentiated, so the per-vertex rows reuse its space.
constexpr int kPatchX = 18, kPatchY = 6, kPatch = kPatchX * kPatchY;

template <int C, bool kShaded>
struct SourceLds {
  static constexpr int R = Record<C, kShaded>::kSize;
  static constexpr int GS = Scratch<C>::kStride;
  static constexpr bool kStageG = GS <= 8;
  static constexpr int kStageBytes =
      kPatch * 16 + (kStageG ? kPatch * GS * 4 : 0);
  // Rows are R + 1 floats apart: lanes adding into different rows then
  // meet different LDS banks (a stride of R = 8 floats puts every lane of a
  // column on the same few banks).
  static constexpr int RS = R + 1;
  static constexpr int kRowBytes = kRows * RS * 4 + kRows * 4;
  static constexpr int kBytes =
      ((kStageBytes > kRowBytes ? kStageBytes : kRowBytes) + 15) & ~15;
};

// One wave per block: the waves of a tile share no LDS and no barrier, and a
// wave with no fragment (most of a frame) then frees its LDS and its slot at
// once instead of holding them until the other waves of its tile finish
// (166 -> 147 us per 4-view launch against 256-thread blocks).
constexpr int kSourcePatches = (kBX / 16) * (kBY / 4);  // 16x4 patches per tile

template <int C, bool kShaded>
__global__ __launch_bounds__(64) void source_kernel(
    const dr_splat_job* __restrict__ jobs, int num_jobs, int L, float coef) {
  constexpr int R = Record<C, kShaded>::kSize;
  constexpr int NB = Record<C, kShaded>::kBase;
  using S = SourceLds<C, kShaded>;
  __shared__ float4 wave_lds[S::kBytes / 16];
  __shared__ int hkey[kHashSlots];
  __shared__ int hrow[kHashSlots];
  // Block b is patch b % kSourcePatches of tile b / kSourcePatches, with the
  // XCD runs of 32 tiles the target pass uses.
  const int64_t item = xcd_block<kSourcePatches * 32>();
  const int64_t tile = item / kSourcePatches;
  const int wave = int(item % kSourcePatches);
  const dr_splat_job& J = jobs[find_job(jobs, num_jobs, tile)];
  const int W = J.width, H = J.height;
  const int ntx = (W + kBX - 1) / kBX;
  const int local = int(tile - J.tile_begin);
  const int lane = __lane_id();
  // Patch w of the tile, patches row by row.
  constexpr int kPatchesX = kBX / 16;
  const int x = (local % ntx) * kBX + 16 * (wave % kPatchesX) + (lane & 15);
  const int y = (local / ntx) * kBY + 4 * (wave / kPatchesX) + (lane >> 4);
  constexpr int RS = S::RS;
  float* val = reinterpret_cast<float*>(wave_lds);  // [kRows][RS]
  int* vtx = reinterpret_cast<int*>(wave_lds) + kRows * RS;
  const uint4* stage_surf = reinterpret_cast<const uint4*>(wave_lds);
  const float4* stage_g = wave_lds + kPatch;
  // Nothing to differentiate into, no gradient scratch to read, or a job
  // of the other colour source (this launch handles jobs with shaded pixels
  // only, or jobs without): the job contributes nothing here (one job per
  // block, so the whole block leaves).
  float* const d_shaded = kShaded ? J.d_shaded : nullptr;
  if ((J.d_records == nullptr && J.d_camera == nullptr &&
       d_shaded == nullptr) ||
      J.grad_scratch == nullptr || (J.shaded != nullptr) != kShaded)
    return;
  const int fshift = J.fixed_shift;
  const float fscale = fshift > 0 ? ldexpf(1.0f, fshift) : 1.0f;
  const bool inside = x < W && y < H;
  const int64_t P = int64_t(W) * H;
  const int64_t p = int64_t(y) * W + x;

  constexpr int GS = Scratch<C>::kStride;
  // A wave whose 64 pixels hold no fragment has nothing to differentiate
  // (most waves of a frame): it leaves before staging its patch, whose
  // records it would fetch from memory for nothing (the block is this one
  // wave, so the exit frees its LDS at once).
  // Layers are a prefix of fragments, so layer 0 tells.
  const bool any = inside && ldg(J.triangle_ids + p) >= 0;
  if (__ballot(any) == 0) {
    // d shaded is 0 wherever no fragment is (the fg mask zeroes the taps).
    if (d_shaded && inside)
      for (int l = 0; l < L; ++l)
        store_row<C>(d_shaded + (l * P + p) * C, nullptr);
    return;
  }
  // Layer 0's corners are loaded now, so their loads are in flight together
  // with the patch staging below instead of after it.
  Corner k0;
  const bool frag0 = any && corners(J, p, k0);
  // Stage the patch's records (out-of-image targets as zeros).  Records of
  // targets that received no splat weight were never written; they are
  // only read for taps of foreground sources, whose targets all were.
  {
    const int px0 = x - (lane & 15) - 1, py0 = y - (lane >> 4) - 1;
    uint4* ss = reinterpret_cast<uint4*>(wave_lds);
    float4* sg = wave_lds + kPatch;
    for (int e = lane; e < kPatch; e += 64) {
      const int tx = px0 + e % kPatchX, ty = py0 + e / kPatchX;
      uint4 sv = make_uint4(0u, 0u, 0u, 0u);
      float4 g4[S::kStageG ? GS / 4 : 1];
#pragma unroll
      for (int i = 0; i < (S::kStageG ? GS / 4 : 1); ++i)
        g4[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
      if (tx >= 0 && tx < W && ty >= 0 && ty < H) {
        const float* gp =
            J.grad_scratch + (int64_t(ty) * W + tx) * Scratch<C>::kHot;
        sv = ldgu4(reinterpret_cast<const uint32_t*>(gp));
        if (S::kStageG) {
#pragma unroll
          for (int i = 0; i < GS / 4; ++i)
            g4[i] = ldg4(gp + 4 + 4 * i);
        }
      }
      ss[e] = sv;
      if (S::kStageG) {
#pragma unroll
        for (int i = 0; i < GS / 4; ++i) sg[i * kPatch + e] = g4[i];
      }
    }
  }
  wave_sync();

  // Surface indices of this source with its 9 targets, as the target pass
  // computed them: byte t of each target's record.  One signed byte per
  // tap, tap t in byte t % 4 of word t / 4.
  auto patch_index = [&](int t) {
    const int u = t / 3 - 1, v = t % 3 - 1;
    return (1 + (lane >> 4) + v) * kPatchX + 1 + (lane & 15) + u;
  };
  uint32_t surf[3] = {0u, 0u, 0u};
  if (any) {
#pragma unroll
    for (int t = 0; t < 9; ++t) {
      const uint4 sv = stage_surf[patch_index(t)];
      const uint32_t word = t < 4 ? sv.x : (t < 8 ? sv.y : sv.z);
      surf[t >> 2] |= ((word >> (8 * (t & 3))) & 0xffu) << (8 * (t & 3));
    }
  }
  // Every lane of the wave runs every layer iteration it visits (ballots
  // below); the wave stops at its first layer without a fragment, since the
  // deeper ones are empty too.
  bool prev = any;  // this lane had a fragment in the previous layer
  int l = 0;
  for (; l < L; ++l) {
    Corner k = k0;
    const bool frag = l == 0 ? frag0 : prev && corners(J, l * P + p, k);
    prev = frag;
    if (l > 0 && __ballot(frag) == 0) break;
    float base[NB];
    float dclip[3] = {0.0f, 0.0f, 0.0f};  // d clip x, y, w
    float dcol[C];
#pragma unroll
    for (int c = 0; c < C; ++c) dcol[c] = 0.0f;
    if (frag) {
      float sx, sy, cx, cy, cw;
      splat_position<false>(J, k, sx, sy, cx, cy, cw);
      float col[C];
      if constexpr (kShaded) {
        const float* sp = J.shaded + (l * P + p) * C;
#pragma unroll
        for (int c = 0; c < C; ++c) {
          if constexpr (C % 4 == 0) {
            if (c % 4 == 0) {
              const float4 q4 = ldg4(sp + c);
              col[c] = q4.x; col[c + 1] = q4.y; col[c + 2] = q4.z;
              col[c + 3] = q4.w;
            }
          } else {
            col[c] = ldg(sp + c);
          }
        }
      } else {
        const float* a0 = J.colors + int64_t(k.v[0]) * C;
        const float* a1 = J.colors + int64_t(k.v[1]) * C;
        const float* a2 = J.colors + int64_t(k.v[2]) * C;
#pragma unroll
        for (int c = 0; c < C; ++c)
          col[c] = interp3(k, ldg(a0 + c), ldg(a1 + c), ldg(a2 + c));
      }
      float gx = 0.0f, gy = 0.0f;
#pragma unroll
      for (int t = 0; t < 9; ++t) {
        const int u = t / 3 - 1, v = t % 3 - 1;
        const int tx = x + u, ty = y + v;
        if (tx < 0 || tx >= W || ty < 0 || ty >= H) continue;
        float dx, dy;
        const float w = tap_weight(coef, sx, sy, u, v, dx, dy);
        const int cls =
            layer_class(int(int8_t(surf[t >> 2] >> (8 * (t & 3)))), l);
        float gv[GS];
        if (S::kStageG && l == 0 && cls == 1) {
          // One float4 plane per quarter of the class record: lanes
          // reading neighbouring patch entries meet distinct LDS banks.
          const float4* gs = stage_g + patch_index(t);
#pragma unroll
          for (int c = 0; c < GS; c += 4) {
            const float4 q4 = gs[(c / 4) * kPatch];
            gv[c] = q4.x; gv[c + 1] = q4.y; gv[c + 2] = q4.z;
            gv[c + 3] = q4.w;
          }
        } else {
          const float* gp = scratch_class<C>(J.grad_scratch, P,
                                             int64_t(ty) * W + tx, cls);
#pragma unroll
          for (int c = 0; c < GS; c += 4) {
            const float4 q4 = ldg4(gp + c);
            gv[c] = q4.x; gv[c + 1] = q4.y; gv[c + 2] = q4.z;
            gv[c + 3] = q4.w;
          }
        }
        float dw = gv[C];
#pragma unroll
        for (int c = 0; c < C; ++c) {
          dcol[c] += w * gv[c];
          dw += gv[c] * col[c];
        }
        const float dw4 = dw * 4.0f * w;
        gx += dw4 * dx;
        gy += dw4 * dy;
      }
      // The offsets are -(x - centre) with x = ((cx / cw + 1) W) / 2, so
      // d cx = gx W / (2 cw), d cw = -(gx W cx + gy H cy) / (2 cw^2) with
      // (gx, gy) the gradient with respect to x, y.
      const float hx = gx * (0.5f * float(W)), hy = gy * (0.5f * float(H));
      const float inv_w = 1.0f / cw;
      const float dcx = hx * inv_w, dcy = hy * inv_w;
      const float dcw = -(dcx * cx + dcy * cy) * inv_w;
      float dp[4];
      if (J.camera) {
        // d v_j = sum_i M[i][j] d clip_i over i in {x, y, w} (z gets none).
        const float* M = J.camera;
#pragma unroll
        for (int j = 0; j < 3; ++j)
          dp[j] = (ldg(M + j) * dcx + ldg(M + 4 + j) * dcy) +
                  ldg(M + 12 + j) * dcw;
        dp[3] = 0.0f;
      } else {
        dp[0] = dcx; dp[1] = dcy; dp[2] = 0.0f; dp[3] = dcw;
      }
#pragma unroll
      for (int c = 0; c < 4; ++c) base[c] = dp[c];
      if constexpr (!kShaded) {
#pragma unroll
        for (int c = 0; c < C; ++c) base[4 + c] = dcol[c];
      }
      dclip[0] = dcx;
      dclip[1] = dcy;
      dclip[2] = dcw;
    }
    if (J.d_camera) {
      // clip = M [p, 1] with p the fragment's interpolated world point (the
      // barycentrics interpolate clip coordinates linearly), so
      // d M[i][j] = sum over fragments of d clip_i [p, 1]_j, rows x, y, w:
      // a wave sum per layer, then one atomic per entry.
      float ph[4] = {0.0f, 0.0f, 0.0f, 0.0f};
      if (frag) {
#pragma unroll
        for (int j = 0; j < 3; ++j)
          ph[j] = interp3(k, ldg(J.positions + 3 * int64_t(k.v[0]) + j),
                          ldg(J.positions + 3 * int64_t(k.v[1]) + j),
                          ldg(J.positions + 3 * int64_t(k.v[2]) + j));
        ph[3] = (k.b[0] + k.b[1]) + k.b[2];
      }
#pragma unroll
      for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          float d = dclip[i] * ph[j];
          for (int off = 32; off > 0; off >>= 1) d += __shfl_xor(d, off);
          if (lane == 0 && d != 0.0f)
            grad_add(J.d_camera, (i < 2 ? 4 * i : 12) + j, d, fshift,
                     fscale);
        }
    }
    // With shaded pixels the colour gradient belongs to the pixel-layer:
    // written once, 0 where no fragment is.
    if (d_shaded && inside)
      store_row<C>(d_shaded + (l * P + p) * C, frag ? dcol : nullptr);
    // Wave-local dedupe into compact LDS rows (one per distinct vertex, in
    // the order first occurrences claim them), then a grouped flush.
    const uint64_t act = __ballot(frag);
    if (act == 0) continue;
    // Corner records, summed over runs of lanes on one triangle when they
    // are narrow enough to keep all three in registers (the runs' first
    // lanes then carry the entries; the other lanes drop out).  Below four
    // channels the kernel holds five waves per SIMD without the merge and
    // four with it, so it is left out there.
    constexpr bool kMergeRuns = NB <= 8 && C >= 4;
    float rec[3][kMergeRuns ? NB : 1];
    bool ent = frag;
    if constexpr (kMergeRuns) {
#pragma unroll
      for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int c = 0; c < NB; ++c)
          rec[i][c] = frag ? k.b[i] * base[c] : 0.0f;
      // Lanes without a fragment get distinct negative keys: runs of one.
      const int tri = frag ? ldg(J.triangle_ids + l * P + p) : -1 - lane;
      const int left = __builtin_amdgcn_update_dpp(
          int(0x80000000), tri, 0x111, 0xf, 0xf, false);  // row_shr:1
      const bool starts = (lane & 15) == 0 || left != tri;
      const uint64_t upto = ~uint64_t(0) >> (63 - lane);
      const int seg = 63 - __builtin_clzll(__ballot(starts) & upto);
      ent = frag && starts;
      if (run_sum_step<1>(seg, rec))
        if (run_sum_step<2>(seg, rec))
          if (run_sum_step<4>(seg, rec)) run_sum_step<8>(seg, rec);
    }
    // Column c of corner i's record.
    auto entry = [&](int i, int c) -> float {
      if constexpr (kMergeRuns)
        return c < NB ? rec[i][c] : 0.0f;
      else
        return record_entry<NB>(k.b[i], base, c);
    };
    for (int i = lane; i < kHashSlots; i += 64) hkey[i] = -1;
    wave_sync();
    float* grads = J.d_records;
    int slot[3] = {0, 0, 0};
    bool first[3] = {false, false, false};
    if (ent) {
      // Linear probing; the table has more slots than a wave has entries,
      // so every probe sequence ends at a free slot or a match.  The three
      // vertices' first probes are issued back to back and collisions walk
      // on afterwards: slots are never freed, so each vertex still ends in
      // exactly one slot whatever the order of the probes.
      uint32_t h[3];
      int old[3];
#pragma unroll
      for (int i = 0; i < 3; ++i) {
        h[i] = (uint32_t(k.v[i]) * 2654435761u) >> 24;
        old[i] = atomicCAS(&hkey[h[i]], -1, k.v[i]);
      }
#pragma unroll
      for (int i = 0; i < 3; ++i) {
        for (int probe = 1;
             probe < kHashSlots && old[i] != -1 && old[i] != k.v[i];
             ++probe) {
          h[i] = (h[i] + 1) & (kHashSlots - 1);
          old[i] = atomicCAS(&hkey[h[i]], -1, k.v[i]);
        }
        first[i] = old[i] == -1;
        slot[i] = int(h[i]);
      }
    }
    // First occurrences take rows in lane order, vertex by vertex: ballots
    // rank them, so no LDS counter (and no atomic round trip) is needed.
    int nr = 0;  // rows taken so far (wave-uniform)
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      const bool f = ent && first[i];
      const uint64_t fm = __ballot(f);
      if (f) {
        const int r = nr + __popcll(fm & lanemask_lt());
        hrow[slot[i]] = r < kRows ? r : -1;
        if (r < kRows) {
          vtx[r] = k.v[i];
#pragma unroll
          for (int c = 0; c < R; ++c)
            val[r * RS + c] = entry(i, c);
        } else {
          flush_record<R>(grads, k.v[i], [&](int c) { return entry(i, c); },
                          fshift, fscale);
        }
      }
      nr += __popcll(fm);
    }
    wave_sync();
    if (ent) {
#pragma unroll
      for (int i = 0; i < 3; ++i) {
        if (first[i]) continue;
        const int r = hrow[slot[i]];
        if (r < 0) {
          flush_record<R>(grads, k.v[i], [&](int c) { return entry(i, c); },
                          fshift, fscale);
          continue;
        }
#pragma unroll
        for (int c = 0; c < R; ++c) {
          const float e = entry(i, c);
          if (e != 0.0f) atomicAdd(&val[r * RS + c], e);
        }
      }
    }
    wave_sync();
    const int n = min(nr, kRows);
    if (grads) {
      for (int e0 = 0; e0 < n; e0 += 64 / R) {
        const int e = e0 + lane / R, c = lane % R;
        if (e < n) {
          const float g = val[e * RS + c];
          if (g != 0.0f)
            grad_add(grads, int64_t(vtx[e]) * R + c, g, fshift, fscale);
        }
      }
    }
    wave_sync();
  }
  // The layers the wave did not visit hold no fragment: d shaded is 0.
  if (d_shaded && inside)
    for (; l < L; ++l) store_row<C>(d_shaded + (l * P + p) * C, nullptr);
}

__global__ void fixed_to_float_kernel(const int64_t* __restrict__ fixed,
                                      int64_t n, double scale,
                                      float* __restrict__ out) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) out[i] = float(double(fixed[i]) * scale);
}

size_t target_lds_bytes(int L, int C) {
  return (size_t(L) * 2 + 2 + C) * kRegion * sizeof(float);
}

float splat_coef() {
  double sum = 0.0;
  for (int u = -1; u <= 1; ++u)
    for (int v = -1; v <= 1; ++v) sum += std::exp(-2.0 * (u * u + v * v));
  return float(1.05 / sum);
}

template <int C, bool kAccum>
int launch_target_kernel(const dr_splat_job* jobs, int n, int64_t tiles,
                         int L, int grad_mode, float loss_scale,
                         float* partials, hipStream_t s) {
  const size_t lds = target_lds_bytes(L, C);
  if (lds > 65536) {
    DR_CHECK_HIP(hipFuncSetAttribute(
        reinterpret_cast<const void*>(&target_kernel<C, kAccum>),
        hipFuncAttributeMaxDynamicSharedMemorySize, int(lds)));
  }
  target_kernel<C, kAccum><<<unsigned(tiles), dim3(kBX, kBY), lds, s>>>(
      jobs, n, L, splat_coef(), grad_mode, loss_scale, partials);
  return kOk;
}

template <int C>
int launch_target(const dr_splat_job* jobs, int n, int64_t tiles, int L,
                  int grad_mode, bool accum, float loss_scale, float* loss,
                  float* partials, hipStream_t s) {
  const int rc =
      accum ? launch_target_kernel<C, true>(jobs, n, tiles, L, grad_mode,
                                            loss_scale, partials, s)
            : launch_target_kernel<C, false>(jobs, n, tiles, L, grad_mode,
                                             loss_scale, partials, s);
  if (rc != kOk) return rc;
  DR_CHECK_LAUNCH("render splat target pass");
  if (grad_mode == 1 && loss) {
    loss_reduce_kernel<<<1, 1024, 0, s>>>(partials, tiles, loss);
    DR_CHECK_LAUNCH("render splat loss reduce");
  }
  return kOk;
}

template <int C>
int launch_source(const dr_splat_job* jobs, int n, int64_t tiles, int L,
                  bool shaded, hipStream_t s) {
  DR_CHECK_ARG(tiles * kSourcePatches < (int64_t(1) << 32),
               "too many tiles for one source-pass launch");
  const unsigned blocks = unsigned(tiles * kSourcePatches);
  if (shaded)
    source_kernel<C, true><<<blocks, 64, 0, s>>>(jobs, n, L, splat_coef());
  else
    source_kernel<C, false><<<blocks, 64, 0, s>>>(jobs, n, L, splat_coef());
  DR_CHECK_LAUNCH("render splat source pass");
  return kOk;
}

// Channel counts the fused kernels are instantiated for: 1-16, 20, 24, 28
// (a vertex gradient record of 4 + C floats fits 32).
#define DR_RENDER_SWITCH(C, CALL) \
  switch (C) {                    \
    case 1: return CALL(1);       \
    case 2: return CALL(2);       \
    case 3: return CALL(3);       \
    case 4: return CALL(4);       \
    case 5: return CALL(5);       \
    case 6: return CALL(6);       \
    case 7: return CALL(7);       \
    case 8: return CALL(8);       \
    case 9: return CALL(9);       \
    case 10: return CALL(10);     \
    case 11: return CALL(11);     \
    case 12: return CALL(12);     \
    case 13: return CALL(13);     \
    case 14: return CALL(14);     \
    case 15: return CALL(15);     \
    case 16: return CALL(16);     \
    case 20: return CALL(20);     \
    case 24: return CALL(24);     \
    case 28: return CALL(28);     \
  }

bool render_channels_ok(int c) {
  return (c >= 1 && c <= 16) || c == 20 || c == 24 || c == 28;
}

int check(int n, int64_t tiles, int L, int C) {
  DR_CHECK_ARG(n > 0, "need at least one job");
  DR_CHECK_ARG(tiles > 0 && tiles < (int64_t(1) << 31), "bad tile count");
  DR_CHECK_ARG(L >= 1 && L <= kMaxLayers, "num_layers must be in [1, 8]");
  DR_CHECK_ARG(render_channels_ok(C),
               "fused render supports 1-16, 20, 24 or 28 channels");
  DR_CHECK_ARG(target_lds_bytes(L, C) <= 160 * 1024,
               "layers x channels exceed the LDS stage");
  return kOk;
}

}  // namespace
}  // namespace dr


extern "C" {

int dr_splat_jobs_prepare(dr_splat_job* jobs, int32_t n, int32_t channels,
                          int64_t* total_tiles) {
  DR_CHECK_ARG(jobs != nullptr && n > 0, "need at least one job");
  DR_CHECK_ARG(dr::render_channels_ok(channels),
               "unsupported channel count (1-16, 20, 24 or 28)");
  int64_t tiles = 0;
  for (int i = 0; i < n; ++i) {
    dr_splat_job& j = jobs[i];
    DR_CHECK_ARG(j.width > 0 && j.height > 0, "image size must be > 0");
    DR_CHECK_ARG(j.num_vertices >= 0 && j.num_triangles >= 0,
                 "negative mesh size");
    DR_CHECK_ARG(j.triangle_ids && j.z_buffer && j.barycentrics,
                 "null rasterizer output");
    DR_CHECK_ARG(j.num_triangles == 0 ||
                     ((j.clip_vertices || (j.positions && j.camera)) &&
                      j.triangles && (j.colors || j.shaded)),
                 "null mesh pointer (clip_vertices, or positions + camera; "
                 "colors or shaded)");
    DR_CHECK_ARG(j.d_shaded == nullptr || j.shaded != nullptr,
                 "d_shaded needs shaded");
    DR_CHECK_ARG((reinterpret_cast<uintptr_t>(j.shaded) & 15) == 0 &&
                     (reinterpret_cast<uintptr_t>(j.d_shaded) & 15) == 0,
                 "shaded and d_shaded must be 16-byte aligned");
    DR_CHECK_ARG((reinterpret_cast<uintptr_t>(j.clip_vertices) & 15) == 0,
                 "clip_vertices must be 16-byte aligned");
    DR_CHECK_ARG((reinterpret_cast<uintptr_t>(j.grad_scratch) & 15) == 0,
                 "grad_scratch must be 16-byte aligned");
    const auto a16 = [](const void* q) {
      return (reinterpret_cast<uintptr_t>(q) & 15) == 0;
    };
    DR_CHECK_ARG(channels % 4 != 0 ||
                     (a16(j.image) && a16(j.target) && a16(j.d_image) &&
                      a16(j.accum) && a16(j.norm)),
                 "image, target, d_image, accum and norm must be 16-byte "
                 "aligned when the channel count is a multiple of 4");
    DR_CHECK_ARG(j.d_camera == nullptr || (j.camera && j.positions),
                 "d_camera needs camera and positions");
    DR_CHECK_ARG(j.fixed_shift >= 0 && j.fixed_shift <= 62,
                 "fixed_shift must be in [0, 62]");
    DR_CHECK_ARG(j.fixed_shift == 0 ||
                     ((reinterpret_cast<uintptr_t>(j.d_records) & 7) == 0 &&
                      (reinterpret_cast<uintptr_t>(j.d_camera) & 7) == 0),
                 "fixed-point gradients need 8-byte aligned buffers");
    j.tile_begin = tiles;
    tiles += int64_t((j.width + dr::kBX - 1) / dr::kBX) *
             ((j.height + dr::kBY - 1) / dr::kBY);
  }
  if (total_tiles) *total_tiles = tiles;
  return DR_OK;
}

int32_t dr_render_scratch_floats(int32_t channels) {
  if (!dr::render_channels_ok(channels)) return 0;
  return 3 * (((channels + 1) + 3) & ~3) + 4;
}

int32_t dr_render_record_size(int32_t channels) {
  if (!dr::render_channels_ok(channels)) return 0;
  return (4 + channels) <= 8 ? 8 : ((4 + channels) <= 16 ? 16 : 32);
}

size_t dr_render_splat_workspace_size(int64_t total_tiles) {
  return size_t(total_tiles > 0 ? total_tiles : 1) * sizeof(float);
}

int dr_render_splat_forward(const dr_splat_job* jobs_device, int32_t num_jobs,
                            int64_t total_tiles, int32_t num_layers,
                            int32_t channels, int32_t grad_mode,
                            float loss_scale, float* loss, void* workspace,
                            size_t workspace_bytes, dr_stream_t stream) {
  int rc = dr::check(num_jobs, total_tiles, num_layers, channels);
  if (rc != DR_OK) return rc;
  DR_CHECK_ARG(jobs_device != nullptr, "jobs is null");
  const bool accum = (grad_mode & DR_RENDER_ACCUM_BUFFERS) != 0;
  grad_mode &= ~DR_RENDER_ACCUM_BUFFERS;
  DR_CHECK_ARG(grad_mode >= 0 && grad_mode <= 2,
               "grad_mode must be 0, 1, 2 (| DR_RENDER_ACCUM_BUFFERS)");
  if (grad_mode == 1) {
    DR_CHECK_ARG(workspace != nullptr, "workspace is null");
    if (workspace_bytes < dr_render_splat_workspace_size(total_tiles))
      return dr::fail(dr::kWorkspaceTooSmall, "render workspace too small");
  }
  hipStream_t s = static_cast<hipStream_t>(stream);
  float* partials = static_cast<float*>(workspace);
#define DR_CALL(CC)                                                     \
  dr::launch_target<CC>(jobs_device, num_jobs, total_tiles, num_layers, \
                        grad_mode, accum, loss_scale, loss, partials, s)
  DR_RENDER_SWITCH(channels, DR_CALL)
#undef DR_CALL
  return dr::fail(dr::kUnsupported, "unsupported channel count");
}

int dr_fixed_to_float(const int64_t* fixed, int64_t n, int32_t fixed_shift,
                      float* out, dr_stream_t stream) {
  DR_CHECK_ARG(n >= 0, "negative count");
  DR_CHECK_ARG(fixed_shift > 0 && fixed_shift <= 62,
               "fixed_shift must be in [1, 62]");
  if (n == 0) return DR_OK;
  DR_CHECK_ARG(fixed && out, "null pointer");
  const unsigned blocks = unsigned((n + 255) / 256);
  dr::fixed_to_float_kernel<<<blocks, 256, 0,
                              static_cast<hipStream_t>(stream)>>>(
      fixed, n, ldexp(1.0, -fixed_shift), out);
  DR_CHECK_LAUNCH("fixed to float");
  return DR_OK;
}

int dr_render_splat_backward(const dr_splat_job* jobs_device,
                             int32_t num_jobs, int64_t total_tiles,
                             int32_t num_layers, int32_t channels,
                             int32_t flags, dr_stream_t stream) {
  int rc = dr::check(num_jobs, total_tiles, num_layers, channels);
  if (rc != DR_OK) return rc;
  DR_CHECK_ARG(jobs_device != nullptr, "jobs is null");
  DR_CHECK_ARG((flags & ~DR_RENDER_SHADED) == 0,
               "flags must be 0 or DR_RENDER_SHADED");
  const bool shaded = (flags & DR_RENDER_SHADED) != 0;
  hipStream_t s = static_cast<hipStream_t>(stream);
#define DR_CALL(CC)                                                        \
  dr::launch_source<CC>(jobs_device, num_jobs, total_tiles, num_layers, \
                        shaded, s)
  DR_RENDER_SWITCH(channels, DR_CALL)
#undef DR_CALL
  return dr::fail(dr::kUnsupported, "unsupported channel count");
}

}  // extern "C"
