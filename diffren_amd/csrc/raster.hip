// Layered z-buffer rasterizer for gfx950: screen-tile binning, per-tile
// coverage and L-deep depth test in LDS, and an in-tile resolve that writes
// triangle ids, z and perspective-correct barycentrics.
//
// Results are bit-identical to diffren's CPU rasterizer
// (diffren/common/kernels/rasterize_triangles_impl.cc:27-118 with the helpers
// of rasterize_utils.h:51-248):
//   * coverage uses the reference's exact int64 16.16 fixed-point edge
//     functions, evaluated with two's-complement wrap;
//   * FP32 math is evaluated without contraction (pragma below, and the file is
//     built with -ffp-contract=off) with correctly rounded division;
//   * float->int conversions reproduce x86-64 cvttss2si (the platform the
//     reference CPU kernel is defined on): NaN / out of range -> INT_MIN;
//   * each pixel keeps the L smallest (z, triangle id) pairs, z compared as
//     floats (so -0 == +0) with ties going to the lower id, which is what the
//     reference's id-ordered insertion sort with strict '<' produces
//     (rasterize_triangles_impl.cc:95-114).  Fragments with z outside
//     [-1, 1] (impl.cc:92), z == 1 (never below the clear depth) or NaN z are
//     never stored.
//
// Pipeline per call (all jobs of a batch in each launch):
//   clear -> bin count -> scan (2 kernels) -> bin fill -> tile raster.
#pragma clang fp contract(off)

#include <algorithm>
#include <cstring>
#include <mutex>

#include "dr_common.h"
#include "../../include/diffren_amd.h"

namespace dr {
namespace {

// Threads per block of the per-triangle and scan kernels: two waves, so
// with two job groups on concurrent streams they interleave with the other
// group's kernels.  The tile pass takes 256 or 128 (dr_raster_config
// tile_threads).
constexpr int kBlock = 128;
// Depth layers per pixel; the tile's L-deep key lists must also fit the
// 64 KiB LDS budget (L x tile pixels x 8 bytes).
constexpr int kMaxLayers = 64;
constexpr uint64_t kEmpty = ~uint64_t(0);
// Clipped-bbox area above which a triangle is rasterized by a whole wave
// instead of a single lane.
constexpr int kWaveRasterArea = 24;

// ---------------------------------------------------------------------------
// Exact reference arithmetic.

__device__ __forceinline__ int f2i_x86(float f) {
  return (f >= -2147483648.0f && f < 2147483648.0f) ? static_cast<int>(f)
                                                     : INT32_MIN;
}

__device__ __forceinline__ int64_t f2l_x86(float f) {
  return (f >= -9223372036854775808.0f && f < 9223372036854775808.0f)
             ? static_cast<int64_t>(f)
             : INT64_MIN;
}

__device__ __forceinline__ int64_t to_fixed(float f) {
  return f2l_x86(f * 65536.0f);
}

__device__ __forceinline__ int64_t wmul(int64_t a, int64_t b) {
  return static_cast<int64_t>(static_cast<uint64_t>(a) *
                              static_cast<uint64_t>(b));
}
__device__ __forceinline__ int64_t wadd(int64_t a, int64_t b) {
  return static_cast<int64_t>(static_cast<uint64_t>(a) +
                              static_cast<uint64_t>(b));
}
__device__ __forceinline__ int64_t wsub(int64_t a, int64_t b) {
  return static_cast<int64_t>(static_cast<uint64_t>(a) -
                              static_cast<uint64_t>(b));
}

struct Box {
  int left, right, bottom, top;
};

__device__ __forceinline__ int clampi(int v, int lo, int hi) {
  return v < lo ? lo : (v > hi ? hi : v);
}

// rasterize_utils.h:75-84
__device__ __forceinline__ void grow(float x, float y, float w, int W, int H,
                                     Box& b) {
  const float sx = 0.5f * (x / w + 1.0f) * static_cast<float>(W);
  const float sy = 0.5f * (y / w + 1.0f) * static_cast<float>(H);
  b.left = clampi(f2i_x86(floorf(fminf(static_cast<float>(b.left), sx))), 0, W);
  b.right = clampi(f2i_x86(ceilf(fmaxf(static_cast<float>(b.right), sx))), 0, W);
  b.bottom =
      clampi(f2i_x86(floorf(fminf(static_cast<float>(b.bottom), sy))), 0, H);
  b.top = clampi(f2i_x86(ceilf(fmaxf(static_cast<float>(b.top), sy))), 0, H);
}

// rasterize_utils.h:86-99
__device__ __forceinline__ void grow_near(const float4& p, const float4& q,
                                          int W, int H, Box& b) {
  const float den = p.z - q.z + p.w - q.w;
  if (den != 0.0f) {
    const float t = (p.z + p.w) / den;
    grow(p.x + t * (q.x - p.x), p.y + t * (q.y - p.y), p.w + t * (q.w - p.w),
         W, H, b);
  }
}

__device__ __forceinline__ bool in_front(const float4& v) {
  return v.w > 0.0f && v.z >= -v.w;
}

// rasterize_utils.h:111-161
__device__ __forceinline__ bool triangle_bbox(const float4 v[3], int W, int H,
                                              Box& b) {
  b.left = W;
  b.right = 0;
  b.bottom = H;
  b.top = 0;
  const bool vis0 = in_front(v[0]), vis1 = in_front(v[1]),
             vis2 = in_front(v[2]);
  if (vis0) {
    grow(v[0].x, v[0].y, v[0].w, W, H, b);
    if (!vis1) grow_near(v[0], v[1], W, H, b);
    if (!vis2) grow_near(v[0], v[2], W, H, b);
  }
  if (vis1) {
    grow(v[1].x, v[1].y, v[1].w, W, H, b);
    if (!vis2) grow_near(v[1], v[2], W, H, b);
    if (!vis0) grow_near(v[1], v[0], W, H, b);
  }
  if (vis2) {
    grow(v[2].x, v[2].y, v[2].w, W, H, b);
    if (!vis0) grow_near(v[2], v[0], W, H, b);
    if (!vis1) grow_near(v[2], v[1], W, H, b);
  }
  return b.right > b.left && b.top > b.bottom;
}

// Fixed-point matrix [[x0 x1 x2] [y0 y1 y2] [w0 w1 w2]] (row-major).
__device__ __forceinline__ void fixed_matrix(const float4 v[3], int64_t m[9]) {
  m[0] = to_fixed(v[0].x); m[1] = to_fixed(v[1].x); m[2] = to_fixed(v[2].x);
  m[3] = to_fixed(v[0].y); m[4] = to_fixed(v[1].y); m[5] = to_fixed(v[2].y);
  m[6] = to_fixed(v[0].w); m[7] = to_fixed(v[1].w); m[8] = to_fixed(v[2].w);
}

// rasterize_utils.h:176-206.  For kNone/kFront the reference multiplies each
// int64 entry by a float, i.e. converts the entry to float, multiplies and
// truncates back; that rounding is reproduced.
__device__ __forceinline__ void adjugate(const int64_t m[9], int cull,
                                         int64_t r[9]) {
  r[0] = wsub(wmul(m[4], m[8]), wmul(m[7], m[5]));
  r[1] = wsub(wmul(m[2], m[7]), wmul(m[8], m[1]));
  r[2] = wsub(wmul(m[1], m[5]), wmul(m[4], m[2]));
  r[3] = wsub(wmul(m[5], m[6]), wmul(m[8], m[3]));
  r[4] = wsub(wmul(m[0], m[8]), wmul(m[6], m[2]));
  r[5] = wsub(wmul(m[2], m[3]), wmul(m[5], m[0]));
  r[6] = wsub(wmul(m[3], m[7]), wmul(m[6], m[4]));
  r[7] = wsub(wmul(m[1], m[6]), wmul(m[7], m[0]));
  r[8] = wsub(wmul(m[0], m[4]), wmul(m[3], m[1]));
  if (cull == DR_CULL_NONE || cull == DR_CULL_FRONT) {
    const int64_t det_i =
        wadd(wadd(wmul(m[0], r[0]), wmul(m[1], r[3])), wmul(m[2], r[6]));
    const float det = static_cast<float>(det_i);
    const float s = (cull == DR_CULL_NONE) ? copysignf(1.0f, det) : -1.0f;
#pragma unroll
    for (int i = 0; i < 9; ++i) r[i] = f2l_x86(static_cast<float>(r[i]) * s);
  }
}

// Fixed-point matrix and adjugate of a triangle.  When every coordinate is
// below 2^15 in magnitude (the common case; NaN fails the test) its
// fixed-point value fits in 32 bits and equals to_fixed's, and the
// adjugate's products are formed from sign-extended 32-bit operands, which
// are cheaper; the wrapping 64-bit arithmetic is the same either way.
__device__ __forceinline__ bool small_coords(const float4 v[3]) {
  const float lim = 32768.0f;
  return fabsf(v[0].x) < lim && fabsf(v[1].x) < lim && fabsf(v[2].x) < lim &&
         fabsf(v[0].y) < lim && fabsf(v[1].y) < lim && fabsf(v[2].y) < lim &&
         fabsf(v[0].w) < lim && fabsf(v[1].w) < lim && fabsf(v[2].w) < lim;
}

// fixed_matrix for small_coords triangles: 32-bit conversions.
__device__ __forceinline__ void fixed_matrix_small(const float4 v[3],
                                                   int64_t m[9]) {
  const float c[9] = {v[0].x, v[1].x, v[2].x, v[0].y, v[1].y,
                      v[2].y, v[0].w, v[1].w, v[2].w};
#pragma unroll
  for (int i = 0; i < 9; ++i) m[i] = static_cast<int32_t>(c[i] * 65536.0f);
}

__device__ __forceinline__ void edge_setup(const float4 v[3], int cull,
                                           int64_t r[9]) {
  int64_t m[9];
  if (small_coords(v)) {
    fixed_matrix_small(v, m);
    adjugate(m, cull, r);
  } else {
    fixed_matrix(v, m);
    adjugate(m, cull, r);
  }
}

// Conservative early rejection of triangles that cannot cover any pixel
// because of culling.  Exact: it only fires when every fixed-point entry is
// below 2^20 in magnitude (so det and all edge values are computed without
// int64 wrap) and all three fixed-point w are positive.  Then, with
// p = (px, py, 2^16), sum_i w_i * e_i(p) = det * 2^16 for the unrounded
// adjugate, so det <= 0 leaves no pixel with all e_i >= 0 and one > 0
// (kBack).  For kFront the adjugate entries were rounded through float;
// the rejection additionally requires det to exceed the largest possible
// effect of that rounding on sum_i w_i * e_i(p) over the image.
__device__ __forceinline__ bool cannot_cover(const int64_t m[9], int cull) {
  if (cull == DR_CULL_NONE) return false;
  const int64_t lim = int64_t(1) << 20;
#pragma unroll
  for (int i = 0; i < 9; ++i)
    if (m[i] >= lim || m[i] <= -lim) return false;
  if (m[6] <= 0 || m[7] <= 0 || m[8] <= 0) return false;
  // Every entry is below 2^20 here: products of their 32-bit values.
  auto mul = [](int64_t a, int64_t b) {
    return int64_t(int32_t(a)) * int64_t(int32_t(b));
  };
  const int64_t r0 = mul(m[4], m[8]) - mul(m[7], m[5]);
  const int64_t r3 = mul(m[5], m[6]) - mul(m[8], m[3]);
  const int64_t r6 = mul(m[3], m[7]) - mul(m[6], m[4]);
  const int64_t det = m[0] * r0 + m[1] * r3 + m[2] * r6;
  if (cull == DR_CULL_BACK) return det <= 0;
  // kFront: culled when det > 0 by more than the float rounding can move.
  if (det <= 0) return false;
  int64_t r[9];
  r[0] = r0; r[3] = r3; r[6] = r6;
  r[1] = mul(m[2], m[7]) - mul(m[8], m[1]);
  r[2] = mul(m[1], m[5]) - mul(m[4], m[2]);
  r[4] = mul(m[0], m[8]) - mul(m[6], m[2]);
  r[5] = mul(m[2], m[3]) - mul(m[5], m[0]);
  r[7] = mul(m[1], m[6]) - mul(m[7], m[0]);
  r[8] = mul(m[0], m[4]) - mul(m[3], m[1]);
  double bound = 0.0;
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    const double row = fabs(double(r[3 * i])) + fabs(double(r[3 * i + 1])) +
                       fabs(double(r[3 * i + 2]));
    bound += double(m[6 + i]) * (row * 0x1p-23 + 3.0);
  }
  return double(det) > 2.0 * bound + 1.0;
}

struct Fragment {
  float b0, b1, b2, z;
};

// Edge value r0 * px + r1 * py + r2 * 2^16 of one adjugate row at the
// fixed-point pixel centre (px, py), in wrapping int64 arithmetic.
__device__ __forceinline__ int64_t edge_at(const int64_t r[3], int32_t px,
                                           int32_t py) {
  return wadd(wadd(wmul(r[0], px), wmul(r[1], py)),
              static_cast<int64_t>(static_cast<uint64_t>(r[2]) << 16));
}
__device__ __forceinline__ int64_t edge_sum(int64_t a, int64_t b, int64_t c) {
  return wadd(wadd(a, b), c);
}

// The same edge values in double precision, from the row (r0, r1, r2 * 2^16)
// as doubles.  When every adjugate entry is below 2^33 in magnitude (true of
// every triangle a few pixels across), each product and partial sum below is
// an integer under 2^51, so the two FMAs are exact and give the int64 value
// (which then does not wrap), and so is the sum of three edges.  A 64-bit
// integer product takes several quarter-rate multiplies; a double FMA is one
// full-rate instruction, and its conversion to float one more.  Neither FMA
// returns -0 (the constant term is an integer converted to double, never -0),
// so a zero edge converts to +0 as in the integer path.
__device__ __forceinline__ double edge_at(const double r[3], int32_t px,
                                          int32_t py) {
  return fma(r[0], static_cast<double>(px),
             fma(r[1], static_cast<double>(py), r[2]));
}
__device__ __forceinline__ double edge_sum(double a, double b, double c) {
  return (a + b) + c;
}

// Double rows of the adjugate r for edge_at; false (and rows unused) when
// some entry is not below 2^33 in magnitude.
__device__ __forceinline__ bool edges_f64(const int64_t r[9], double d[9]) {
  bool exact = true;
#pragma unroll
  for (int i = 0; i < 9; ++i) {
    exact &= static_cast<uint64_t>(r[i]) + (uint64_t(1) << 33) <
             (uint64_t(1) << 34);
    d[i] = static_cast<double>(r[i]) * (i % 3 == 2 ? 65536.0 : 1.0);
  }
  return exact;
}

// Calls fn(rows) with the adjugate's exact double rows when edges_f64
// accepts them (every triangle a few pixels across), else with the int64
// adjugate; shade_pixel gives the same bits with either.
template <typename Fn>
__device__ __forceinline__ void with_edges(const int64_t r[9], Fn&& fn) {
  double d[9];
  if (edges_f64(r, d)) fn(d); else fn(r);
}

// Edge test at pixel (ix, iy); rasterize_utils.h:220-248 and
// rasterize_triangles_impl.cc:73-92.  Returns true when the pixel is inside
// and its z is storable.  `r` holds the adjugate as int64 or, when
// edges_f64 accepted it, its exact double rows; both give the same bits.
template <typename E>
__device__ __forceinline__ bool shade_pixel(int ix, int iy, float sx2,
                                            float sy2, const E r[9],
                                            const float4 v[3], Fragment& f) {
  const float px = sx2 * (static_cast<float>(ix) + 0.5f) - 1.0f;
  const float py = sy2 * (static_cast<float>(iy) + 0.5f) - 1.0f;
  // Pixel centres of the image lie in (-1, 1), so px * 2^16 is within 32-bit
  // range and its 32-bit truncation is to_fixed's value exactly; the
  // sign-extended 32-bit operand makes the edge products cheaper.
  const int32_t pxi = static_cast<int32_t>(px * 65536.0f);
  const int32_t pyi = static_cast<int32_t>(py * 65536.0f);
  E e[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) e[i] = edge_at(r + 3 * i, pxi, pyi);
  if (!((e[0] >= 0 && e[1] >= 0 && e[2] >= 0) &&
        (e[0] > 0 || e[1] > 0 || e[2] > 0)))
    return false;
  const float sum = static_cast<float>(edge_sum(e[0], e[1], e[2]));
  f.b0 = static_cast<float>(e[0]) / sum;
  f.b1 = static_cast<float>(e[1]) / sum;
  f.b2 = static_cast<float>(e[2]) / sum;
  const float cz = f.b0 * v[0].z + f.b1 * v[1].z + f.b2 * v[2].z;
  const float cw = f.b0 * v[0].w + f.b1 * v[1].w + f.b2 * v[2].w;
  f.z = cz / cw;
  // Storable iff -1 <= z < 1 (z == 1 never beats the clear depth; NaN never
  // compares below anything).
  return f.z >= -1.0f && f.z < 1.0f;
}

// Order-preserving 32-bit image of z in [-1, 1) with -0 folded onto +0.
__device__ __forceinline__ uint32_t z_order(float z) {
  const uint32_t b = __float_as_uint(z == 0.0f ? 0.0f : z);
  return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}

// ---------------------------------------------------------------------------
// Jobs.

// All kernels read jobs from a table in global memory (the single-job entry
// point stores its one job into the workspace first), so a job is never
// selected between a kernel-argument copy and a table entry.
struct Jobs {
  const dr_raster_job* table;  // device table of `count` jobs
  int32_t count;
};

// Largest j with begin(j) <= flat.
template <bool kTiles>
__device__ __forceinline__ int find_job(const Jobs& J, int64_t flat) {
  int lo = 0, hi = J.count - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    const int64_t b =
        kTiles ? J.table[mid].tile_begin : J.table[mid].triangle_begin;
    if (b <= flat) lo = mid; else hi = mid - 1;
  }
  return lo;
}

__global__ void store_job_kernel(dr_raster_job job,
                                 dr_raster_job* __restrict__ slot) {
  *slot = job;
}

// Clip-space position of world point (x, y, z, 1) under the row-major 4x4
// matrix M: clip_i = ((M[i,0] x + M[i,1] y) + M[i,2] z) + M[i,3], the
// expression dr_transform_homogeneous evaluates (so results are identical to
// rasterizing its output).
__device__ __forceinline__ float4 transform_point(const float* __restrict__ M,
                                                  float x, float y, float z) {
  float4 r;
  float* rp = &r.x;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const float* m = M + 4 * i;
    rp[i] = ((ldg(m) * x + ldg(m + 1) * y) + ldg(m + 2) * z) +
            ldg(m + 3) * 1.0f;
  }
  return r;
}

// Loads triangle `t` of `job` in clip space; false if an index is out of
// range.  The three indices are loaded together, then the three vertices
// together (two dependent round trips, not one per vertex), as global-memory
// loads: the job table hands out generic pointers, whose flat loads would
// also hold up every wait on the LDS counter.
__device__ __forceinline__ bool load_triangle(const dr_raster_job& job,
                                              int t, float4 v[3]) {
  const int64_t b = 3 * int64_t(t);
  const int32_t i0 = ldg(job.triangles + b), i1 = ldg(job.triangles + b + 1),
                i2 = ldg(job.triangles + b + 2);
  const uint32_t nv = uint32_t(job.num_vertices);
  if (uint32_t(i0) >= nv || uint32_t(i1) >= nv || uint32_t(i2) >= nv)
    return false;
  const int32_t idx[3] = {i0, i1, i2};
  if (job.camera) {
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      const float* p = job.vertices + 3 * int64_t(idx[k]);
      v[k] = transform_point(job.camera, ldg(p), ldg(p + 1), ldg(p + 2));
    }
  } else {
#pragma unroll
    for (int k = 0; k < 3; ++k) v[k] = ldg4(job.vertices + 4 * int64_t(idx[k]));
  }
  return true;
}

__device__ __forceinline__ int tiles_x(const dr_raster_job& job, int tw) {
  return (job.width + tw - 1) / tw;
}

// ---------------------------------------------------------------------------
// Kernels.

__global__ void clear_kernel(uint32_t* p, int64_t n) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}


// Binning.  The count pass visits every triangle once (one lane each):
// gathers its vertices, clips its bounding box, and rejects it when culling
// leaves it no pixel.  A triangle whose clipped box is at most 4 x 4 pixels
// (nearly every triangle of a dense mesh) is rasterized right there: its
// fixed-point set-up and edge tests run once, and when it covers at most six
// pixel centres its fragments -- pixel and z, the same arithmetic as the
// tile pass -- are kept in a compact per-triangle record; a triangle that
// covers none leaves no entry at all.  Larger triangles (or small ones with
// more fragments) are binned as whole triangles into every tile their box
// touches, and the tile pass walks their pixels.
//
// Per-triangle record, two planes:
//   head[t] = (w0, w1):
//     w0 == 0            no entry;
//     w0 & kHeadFrag     n = w0 & 7 fragments (1..6) with 4-bit offsets
//                        (dy << 2 | dx) from the box origin in bits 3..26,
//                        w1 = origin x | origin y << 16;
//     w0 & kHeadTri      nx = w0 & 255 by ny = (w0 >> 8) & 255 tiles from
//                        first tile w1 (flat over all jobs), tiles per row
//                        (w0 >> 16) & 8191; kHeadRecompute: recompute them.
//   frag_z[t][8]: the fragments' z (float bits), written for fragments only.
// Each (tile, entry) of the count pass adds one to counter[tile]; after the
// scan the fill pass replays the same entries from the records (no vertex
// gathers) and appends them to their tiles' bins: a 64-bit key
// (z order << 32 | triangle) with the pixel's index in the tile, or
// (kTriMarker << 32 | triangle) for a whole triangle.  Lanes of a wave that
// hit the same tile share one atomic (ballot, then ranks from popcounts).
constexpr uint32_t kHeadFrag = 0x80000000u;
constexpr uint32_t kHeadTri = 0x40000000u;
constexpr uint32_t kHeadRecompute = 0x20000000u;
constexpr int kInlineFrags = 6;
constexpr int kSmallSide = 4;
// High word of a whole-triangle bin entry: no storable z has this order key
// (z < 1 maps below 0xbf800000).
constexpr uint32_t kTriMarker = 0xffffffffu;

// Entries of one lane's triangle.
struct LaneEntries {
  uint32_t w0 = 0, w1 = 0;  // record head
  int n = 0;                // entries
  int t = 0;                // triangle index in its job
  // fragments: the job's first tile and tiles per row
  int64_t tile_begin = 0;
  int ntx = 1;
  // whole triangle: first tile, nx tiles per tile row, tiles per image row
  uint32_t tile0 = 0;
  int nx = 1;
  uint32_t tpr = 1;
};

template <int TW, int TH>
__device__ __forceinline__ uint32_t entry_tile(const LaneEntries& e, int k,
                                               int& pix) {
  if (e.w0 & kHeadFrag) {
    const uint32_t o = (e.w0 >> (3 + 4 * k)) & 15u;
    const int px = int(e.w1 & 0xffffu) + int(o & 3u);
    const int py = int(e.w1 >> 16) + int(o >> 2);
    const int tx = px / TW, ty = py / TH;
    pix = (py - ty * TH) * TW + (px - tx * TW);
    return uint32_t(e.tile_begin + int64_t(ty) * e.ntx + tx);
  }
  pix = 0;
  return e.tile0 + uint32_t(k / e.nx) * e.tpr + uint32_t(k % e.nx);
}

// Tile range of a whole triangle with clipped box b.
template <int TW, int TH>
__device__ __forceinline__ void whole_triangle(const dr_raster_job& job,
                                               const Box& b, LaneEntries& e) {
  const int tx0 = b.left / TW, ty0 = b.bottom / TH;
  const int ny = (b.top - 1) / TH - ty0 + 1;
  e.nx = (b.right - 1) / TW - tx0 + 1;
  e.n = e.nx * ny;
  e.tpr = uint32_t(tiles_x(job, TW));
  e.tile0 = uint32_t(job.tile_begin + int64_t(ty0) * e.tpr + tx0);
  e.w1 = e.tile0;
  e.w0 = (e.nx < 256 && ny < 256 && e.tpr < 8192)
             ? kHeadTri | uint32_t(e.nx) | (uint32_t(ny) << 8) |
                   (e.tpr << 16)
             : kHeadTri | kHeadRecompute;
}

// Whole-triangle entries from a record head (recomputed from the vertices
// when the range did not fit the head).
template <int TW, int TH>
__device__ __forceinline__ void whole_from_head(const Jobs& J, int64_t flat,
                                                LaneEntries& e) {
  if (e.w0 & kHeadRecompute) {
    const dr_raster_job& job = J.table[find_job<false>(J, flat)];
    float4 v[3];
    Box b;
    e.n = 0;
    if (load_triangle(job, e.t, v) &&
        triangle_bbox(v, job.width, job.height, b))
      whole_triangle<TW, TH>(job, b, e);
    return;
  }
  e.nx = int(e.w0 & 255u);
  e.n = e.nx * int((e.w0 >> 8) & 255u);
  e.tpr = (e.w0 >> 16) & 8191u;
  e.tile0 = e.w1;
}

// Groups the lanes of a wave by tile with ballots alone (no memory
// traffic): each lane with `has` learns the lowest lane of its group, its
// rank in the group and the group size.
__device__ __forceinline__ void group_by_tile(bool has, uint32_t tile,
                                              int& leader, int& rank,
                                              int& size) {
  uint64_t todo = __ballot(has);
  leader = rank = size = 0;
  while (todo) {
    const int first = __ffsll(static_cast<unsigned long long>(todo)) - 1;
    const uint32_t lt = __builtin_amdgcn_readlane(tile, first);
    const bool mine = has && tile == lt;
    const uint64_t peers = __ballot(mine);
    if (mine) {
      leader = first;
      rank = __popcll(peers & lanemask_lt());
      size = __popcll(peers);
    }
    todo &= ~peers;
  }
}

// Count-pass work of one loaded triangle.  Instantiated separately for
// small_coords triangles, whose fixed-point entries are sign-extended 32-bit
// values: their cull test and adjugate then use 32 x 32 -> 64-bit products
// (a merged matrix would make every product a full 64-bit one).
template <int TW, int TH, bool kSmall>
__device__ __forceinline__ void count_triangle(const dr_raster_job& job,
                                               const float4 v[3], int cull,
                                               int64_t flat, LaneEntries& e,
                                               uint4* __restrict__ frag_z) {
  int64_t m[9];
  if (kSmall) fixed_matrix_small(v, m); else fixed_matrix(v, m);
  // The exact cull first: it needs only the fixed-point matrix, and the
  // triangles it rejects (most back faces) then skip the clipped box and
  // its six divisions.  Either test alone means no entries.
  Box b;
  if (cannot_cover(m, cull) || !triangle_bbox(v, job.width, job.height, b))
    return;
  if (b.right - b.left <= kSmallSide && b.top - b.bottom <= kSmallSide) {
    // The tile pass's arithmetic exactly (edge_setup's adjugate of the
    // same fixed-point matrix, shade_pixel).
    int64_t r[9];
    adjugate(m, cull, r);
    const float sx2 = 2.0f / static_cast<float>(job.width);
    const float sy2 = 2.0f / static_cast<float>(job.height);
    int nf = 0;
    uint32_t offs = 0;
    uint32_t z[kInlineFrags] = {0u, 0u, 0u, 0u, 0u, 0u};
    auto scan_box = [&](const auto* edges) {
      for (int iy = b.bottom; iy < b.top; ++iy)
        for (int ix = b.left; ix < b.right; ++ix) {
          Fragment f;
          if (!shade_pixel(ix, iy, sx2, sy2, edges, v, f)) continue;
          if (nf < kInlineFrags) {
            offs |= uint32_t(((iy - b.bottom) << 2) | (ix - b.left))
                    << (4 * nf);
#pragma unroll
            for (int k = 0; k < kInlineFrags; ++k)
              if (k == nf) z[k] = __float_as_uint(f.z);
          }
          ++nf;
        }
    };
    double d[9];
    if (edges_f64(r, d)) scan_box(d); else scan_box(r);
    if (nf <= kInlineFrags) {
      e.n = nf;
      if (nf > 0) {
        e.w0 = kHeadFrag | uint32_t(nf) | (offs << 3);
        e.w1 = uint32_t(b.left) | (uint32_t(b.bottom) << 16);
        frag_z[2 * flat] = make_uint4(z[0], z[1], z[2], z[3]);
        if (nf > 4) frag_z[2 * flat + 1] = make_uint4(z[4], z[5], 0u, 0u);
      }
      return;
    }
  }
  whole_triangle<TW, TH>(job, b, e);
}

template <int TW, int TH>
__global__ __launch_bounds__(kBlock) void bin_count_kernel(
    Jobs J, int64_t total_tris, int cull, uint32_t* __restrict__ counter,
    uint2* __restrict__ head, uint4* __restrict__ frag_z) {
  const int64_t flat = int64_t(blockIdx.x) * kBlock + threadIdx.x;
  LaneEntries e;
  if (flat < total_tris) {
    const dr_raster_job& job = J.table[find_job<false>(J, flat)];
    e.t = int(flat - job.triangle_begin);
    e.tile_begin = job.tile_begin;
    e.ntx = tiles_x(job, TW);
    float4 v[3];
    if (load_triangle(job, e.t, v)) {
      if (small_coords(v))
        count_triangle<TW, TH, true>(job, v, cull, flat, e, frag_z);
      else
        count_triangle<TW, TH, false>(job, v, cull, flat, e, frag_z);
    }
    head[flat] = make_uint2(e.w0, e.w1);
  }
  // Wave-uniform loop over the k-th entry of every lane; one atomic per
  // distinct tile, all groups' leaders in one instruction.
  for (int k = 0; __ballot(k < e.n); ++k) {
    const bool has = k < e.n;
    int pix;
    const uint32_t tile = has ? entry_tile<TW, TH>(e, k, pix) : 0xffffffffu;
    int leader, rank, size;
    group_by_tile(has, tile, leader, rank, size);
    if (has && rank == 0) atomicAdd(&counter[tile], uint32_t(size));
  }
}

template <int TW, int TH>
__global__ __launch_bounds__(kBlock) void bin_fill_kernel(
    Jobs J, int64_t total_tris, uint32_t* __restrict__ counter,
    const uint32_t* __restrict__ offsets, uint64_t* __restrict__ bin_key,
    uint16_t* __restrict__ bin_pix, uint64_t capacity,
    const uint2* __restrict__ head, const uint4* __restrict__ frag_z) {
  const int64_t flat = int64_t(blockIdx.x) * kBlock + threadIdx.x;
  LaneEntries e;
  uint32_t z[kInlineFrags] = {0u, 0u, 0u, 0u, 0u, 0u};
  if (flat < total_tris) {
    const uint2 h = head[flat];
    e.w0 = h.x;
    e.w1 = h.y;
    if (e.w0 != 0) {
      const dr_raster_job& job = J.table[find_job<false>(J, flat)];
      e.t = int(flat - job.triangle_begin);
      e.tile_begin = job.tile_begin;
      e.ntx = tiles_x(job, TW);
      if (e.w0 & kHeadFrag) {
        e.n = int(e.w0 & 7u);
        const uint4 a = frag_z[2 * flat];
        z[0] = a.x; z[1] = a.y; z[2] = a.z; z[3] = a.w;
        if (e.n > 4) {
          const uint4 c = frag_z[2 * flat + 1];
          z[4] = c.x; z[5] = c.y;
        }
      } else {
        whole_from_head<TW, TH>(J, flat, e);
      }
    }
  }
  for (int k = 0; __ballot(k < e.n); ++k) {
    const bool has = k < e.n;
    int pix = 0;
    uint32_t tile = 0xffffffffu;
    if (has) tile = entry_tile<TW, TH>(e, k, pix);
    int leader, rank, size;
    group_by_tile(has, tile, leader, rank, size);
    // The tile's bin offset is read before the cursor atomic returns.
    const uint32_t tile_off = has ? ldg(offsets + tile) : 0u;
    uint32_t base = 0;
    if (has && rank == 0) base = atomicAdd(&counter[tile], uint32_t(size));
    base = __shfl(base, leader, 64);
    if (has) {
      const uint64_t pos = uint64_t(tile_off) + base + uint32_t(rank);
      if (pos < capacity) {
        uint32_t hi = kTriMarker;
        if (e.w0 & kHeadFrag) {
          uint32_t zb = z[0];
#pragma unroll
          for (int q = 1; q < kInlineFrags; ++q)
            if (k == q) zb = z[q];
          hi = z_order(__uint_as_float(zb));
          bin_pix[pos] = uint16_t(pix);
        }
        bin_key[pos] = (uint64_t(hi) << 32) | uint32_t(e.t);
      }
    }
  }
}

// Exclusive scan of counts[0..n) -> offsets, in two kernels: per-block scans
// with block totals, then each block adds the sum of the totals before it.
// The local scan also zeroes the counts it has read: the fill pass reuses
// them as per-tile cursors (no separate clearing launch).
constexpr int kScanItems = 4;  // per thread
constexpr int kScanTile = kBlock * kScanItems;

__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t x,
                                                         uint32_t* total) {
  __shared__ uint32_t wave_sums[kBlock / 64];
  const int lane = lane_id(), wave = threadIdx.x / 64;
  uint32_t incl = x;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const uint32_t y = __shfl_up(incl, d, 64);
    if (lane >= d) incl += y;
  }
  if (lane == 63) wave_sums[wave] = incl;
  __syncthreads();
  uint32_t before = 0, all = 0;
#pragma unroll
  for (int w = 0; w < kBlock / 64; ++w) {
    const uint32_t s = wave_sums[w];
    if (w < wave) before += s;
    all += s;
  }
  __syncthreads();
  *total = all;
  return before + incl - x;
}

__global__ __launch_bounds__(kBlock) void scan_local_kernel(
    uint32_t* __restrict__ counts, uint32_t* __restrict__ offsets,
    uint32_t* __restrict__ block_sums, int64_t n) {
  const int64_t base = int64_t(blockIdx.x) * kScanTile + threadIdx.x * kScanItems;
  uint32_t v[kScanItems], s = 0;
#pragma unroll
  for (int i = 0; i < kScanItems; ++i) {
    v[i] = (base + i < n) ? counts[base + i] : 0u;
    s += v[i];
  }
  uint32_t total;
  uint32_t run = block_exclusive_scan(s, &total);
#pragma unroll
  for (int i = 0; i < kScanItems; ++i) {
    if (base + i < n) {
      offsets[base + i] = run;
      counts[base + i] = 0u;
    }
    run += v[i];
  }
  if (threadIdx.x == 0) block_sums[blockIdx.x] = total;
}

// status[0] = total entries (low 32 bits), status[1] = 1 if > capacity,
// status[2] = high 32 bits of total.
__global__ __launch_bounds__(kBlock) void scan_add_kernel(
    uint32_t* __restrict__ offsets, const uint32_t* __restrict__ block_sums,
    int nblocks, int64_t n, uint32_t* __restrict__ status, uint64_t capacity) {
  __shared__ unsigned long long red[kBlock];
  unsigned long long before = 0, all = 0;
  for (int i = threadIdx.x; i < nblocks; i += kBlock) {
    const unsigned long long s = block_sums[i];
    if (i < int(blockIdx.x)) before += s;
    all += s;
  }
  red[threadIdx.x] = before;
  __syncthreads();
  for (int d = kBlock / 2; d > 0; d >>= 1) {
    if (threadIdx.x < d) red[threadIdx.x] += red[threadIdx.x + d];
    __syncthreads();
  }
  before = red[0];
  __syncthreads();
  red[threadIdx.x] = all;
  __syncthreads();
  for (int d = kBlock / 2; d > 0; d >>= 1) {
    if (threadIdx.x < d) red[threadIdx.x] += red[threadIdx.x + d];
    __syncthreads();
  }
  all = red[0];
  const int64_t base = int64_t(blockIdx.x) * kScanTile;
  for (int i = threadIdx.x; i < kScanTile; i += kBlock)
    if (base + i < n) offsets[base + i] += uint32_t(before);
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    status[0] = uint32_t(all);
    status[2] = uint32_t(all >> 32);
    status[1] = all > capacity ? 1u : 0u;
  }
}

// Inserts key into the L-deep sorted list of LDS slots at `p` (slot l at
// keys[l * TP + p]).  Each slot keeps the minimum it has seen and passes the
// larger key down, so after all insertions slot l holds the (l+1)-th smallest
// (z, id) key regardless of arrival order.
//
// That equals the reference's id-ordered insertion sort with strict '<'
// (rasterize_triangles_impl.cc:95-114) except at exact z ties: a displaced
// entry cascading into a run of equal-z entries moves past them, so which of
// the tied fragments survive, and in what order, depends on the insertion
// history.  Any comparison of two keys with equal z therefore flags the pixel
// for an exact sequential replay (replay_pixel).  A tie that can change the
// result always produces such a comparison: if keys A < B have equal z and
// adjacent ranks r, r+1 <= L, both reach slot r, and whichever arrives second
// meets the other there (slot r only ever holds keys >= A, and after B has
// arrived, keys <= B).
__device__ __forceinline__ void insert_key(uint64_t* keys, uint32_t* ties,
                                           int tp, int p, int L,
                                           uint64_t key) {
  for (int l = 0; l < L; ++l) {
    const uint64_t old = atomicMin(
        reinterpret_cast<unsigned long long*>(&keys[l * tp + p]),
        static_cast<unsigned long long>(key));
    if (old == kEmpty) return;
    if ((old >> 32) == (key >> 32)) atomicOr(&ties[p >> 5], 1u << (p & 31));
    key = old > key ? old : key;
  }
}

// Re-resolves one pixel by replaying the reference's insertion sort
// (rasterize_triangles_impl.cc:95-114) over its fragments in ascending
// triangle id.  Only fragments with z at most the largest kept z take part
// when the list is full: a fragment with larger z is never kept, and until it
// is pushed out it acts on the smaller ones exactly like an empty slot (both
// compare greater), so it cannot change which of them survive or their order.
//
// The whole block gathers the pixel's candidate fragments from the tile's
// bin in parallel into LDS; one thread orders them by id and replays.  If a
// pixel has more candidates than the LDS list holds (many fragments with
// exactly equal z), that thread rescans the bin once per candidate instead.
constexpr int kMaxCandidates = 64;

__device__ __forceinline__ bool fragment_at(const dr_raster_job& job, int t,
                                            int ix, int iy, float sx2,
                                            float sy2, int cull, Fragment& f) {
  float4 v[3];
  Box b;
  if (!load_triangle(job, t, v) ||
      !triangle_bbox(v, job.width, job.height, b))
    return false;
  if (ix < b.left || ix >= b.right || iy < b.bottom || iy >= b.top)
    return false;
  int64_t r[9];
  edge_setup(v, cull, r);
  bool hit = false;
  with_edges(r, [&](const auto* e) {
    hit = shade_pixel(ix, iy, sx2, sy2, e, v, f);
  });
  return hit;
}

// Reference insertion of (z, id) into the L-deep list lz/lid.
__device__ __forceinline__ void reference_insert(float* lz, int* lid, int L,
                                                 float z, int id) {
  for (int l = 0; l < L; ++l) {
    if (z < lz[l]) {
      const float tz = lz[l]; lz[l] = z; z = tz;
      const int ti = lid[l]; lid[l] = id; id = ti;
    }
    if (z == 1.0f) break;
  }
}

// z of an order key (z_order's inverse; -0 comes back as +0, which compares
// the same).
__device__ __forceinline__ float z_from_order(uint32_t o) {
  return __uint_as_float((o & 0x80000000u) ? (o & 0x7fffffffu) : ~o);
}

// Where a tile's items come from: its bin entries, or -- when the bins
// overflowed -- every triangle of the job through its count-pass record.
struct TileItems {
  bool overflow;
  uint64_t off;                  // first bin entry
  const uint64_t* bin_key;
  const uint16_t* bin_pix;
  const uint2* head;
  const uint32_t* frag_z;        // [T][8] words
};

// Fragment of item i at pixel (ix, iy), p its index in the tile: triangle t
// and its z, or false when item i has none there.
__device__ __forceinline__ bool item_fragment(const dr_raster_job& job,
                                              const TileItems& it, int64_t i,
                                              int ix, int iy, int p, float sx2,
                                              float sy2, int cull, int& t,
                                              float& z) {
  if (!it.overflow) {
    const uint64_t key = it.bin_key[it.off + i];
    t = int(uint32_t(key));
    const uint32_t hi = uint32_t(key >> 32);
    if (hi != kTriMarker) {
      if (int(it.bin_pix[it.off + i]) != p) return false;
      z = z_from_order(hi);
      return true;
    }
  } else {
    t = int(i);
    const int64_t flat = job.triangle_begin + i;
    const uint2 h = it.head[flat];
    if (h.x == 0) return false;
    if (h.x & kHeadFrag) {
      // A fragment record lists all of the triangle's fragments.
      const int dx = ix - int(h.y & 0xffffu), dy = iy - int(h.y >> 16);
      if (dx < 0 || dx >= kSmallSide || dy < 0 || dy >= kSmallSide)
        return false;
      const uint32_t want = uint32_t((dy << 2) | dx);
      const int n = int(h.x & 7u);
      for (int k = 0; k < n; ++k)
        if (((h.x >> (3 + 4 * k)) & 15u) == want) {
          z = __uint_as_float(it.frag_z[8 * flat + k]);
          return true;
        }
      return false;
    }
  }
  Fragment f;
  if (!fragment_at(job, t, ix, iy, sx2, sy2, cull, f)) return false;
  z = f.z;
  return true;
}

// Block-cooperative; every thread of the block must call it.  Inlined: as a
// called function its register demand became the whole tile kernel's (three
// resident waves per SIMD instead of five).
__device__ __forceinline__ void replay_pixel(
    const dr_raster_job& job, const TileItems& it, int64_t items, int ix,
    int iy, int p, float sx2, float sy2, int cull, int L, uint64_t* slot,
    int tp, uint64_t* cand, int* cand_n, float* replay_z, int* replay_id) {
  const uint64_t last = slot[(L - 1) * tp];
  const uint32_t zmax = last == kEmpty ? 0xffffffffu : uint32_t(last >> 32);
  if (threadIdx.x == 0) *cand_n = 0;
  __syncthreads();
  for (int64_t i = threadIdx.x; i < items; i += blockDim.x) {
    int t;
    float z;
    if (!item_fragment(job, it, i, ix, iy, p, sx2, sy2, cull, t, z)) continue;
    if (z_order(z) > zmax) continue;
    const int k = atomicAdd(cand_n, 1);
    if (k < kMaxCandidates)
      cand[k] = (uint64_t(uint32_t(t)) << 32) | __float_as_uint(z);
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    const int n = *cand_n;
    float* lz = replay_z;
    int* lid = replay_id;
    for (int l = 0; l < L; ++l) {
      lz[l] = 1.0f;
      lid[l] = -1;
    }
    if (n <= kMaxCandidates) {
      // Insertion sort by id (upper 32 bits; ids are distinct).
      for (int a = 1; a < n; ++a) {
        const uint64_t x = cand[a];
        int b = a - 1;
        while (b >= 0 && cand[b] > x) {
          cand[b + 1] = cand[b];
          --b;
        }
        cand[b + 1] = x;
      }
      for (int a = 0; a < n; ++a)
        reference_insert(lz, lid, L, __uint_as_float(uint32_t(cand[a])),
                         int(cand[a] >> 32));
    } else {
      int prev = -1;
      for (;;) {
        int next = INT32_MAX;
        float next_z = 1.0f;
        for (int64_t i = 0; i < items; ++i) {
          int t;
          float z;
          if (!item_fragment(job, it, i, ix, iy, p, sx2, sy2, cull, t, z))
            continue;
          if (t <= prev || t >= next) continue;
          if (z_order(z) > zmax) continue;
          next = t;
          next_z = z;
        }
        if (next == INT32_MAX) break;
        reference_insert(lz, lid, L, next_z, next);
        prev = next;
      }
    }
    for (int l = 0; l < L; ++l)
      slot[l * tp] = lid[l] < 0 ? kEmpty
                                : (uint64_t(z_order(lz[l])) << 32) |
                                      uint32_t(lid[l]);
  }
  __syncthreads();
}

struct TriRect {
  int tri;
  short x0, x1, y0, y1;  // clipped to the tile, image coordinates
};

// Rasterizes all bin entries of one tile per block, then resolves the tile.
// A fragment entry is inserted into the tile's depth lists as it is (its
// set-up and edge test ran in the count pass); a whole-triangle entry is
// set up here and its pixels walked: by its own lane when its clipped rect
// is small, else queued and walked by a whole wave, one pixel per lane.
// Five waves per SIMD (at most 96 VGPRs).
template <int TW, int TH, int kThreads>
__global__ __launch_bounds__(kThreads)
__attribute__((amdgpu_waves_per_eu(5))) void raster_tile_kernel(
    Jobs J, int L, int cull, const uint32_t* __restrict__ counts,
    const uint32_t* __restrict__ offsets,
    const uint64_t* __restrict__ bin_key, const uint16_t* __restrict__ bin_pix,
    const uint2* __restrict__ head, const uint32_t* __restrict__ frag_z,
    uint64_t capacity, bool sparse, bool prefix) {
  constexpr int TP = TW * TH;
  extern __shared__ uint64_t keys[];  // [L][TP]
  __shared__ TriRect queue[kThreads];
  __shared__ int queue_n;
  __shared__ uint32_t ties[(TP + 31) / 32];
  __shared__ uint64_t cand[kMaxCandidates];
  __shared__ int cand_n;
  // The replaying thread's L-deep list.
  __shared__ float replay_z[kMaxLayers];
  __shared__ int replay_id[kMaxLayers];

  const int64_t flat_tile = xcd_block<4>();
  // One job per block: its fields are held in scalar registers (loaded
  // once, made wave-uniform), not re-gathered per lane at every triangle.
  const int j = __builtin_amdgcn_readfirstlane(find_job<true>(J, flat_tile));
  const dr_raster_job job = load_uniform(J.table + j);
  const int W = job.width, H = job.height;
  const int ntx = tiles_x(job, TW);
  const int local = int(flat_tile - job.tile_begin);
  const int x0 = (local % ntx) * TW, y0 = (local / ntx) * TH;
  const int x1 = min(x0 + TW, W), y1 = min(y0 + TH, H);
  const float sx2 = 2.0f / static_cast<float>(W);
  const float sy2 = 2.0f / static_cast<float>(H);

  const uint32_t cnt = counts[flat_tile];
  const uint64_t off = offsets[flat_tile];
  TileItems it;
  it.overflow = off + cnt > capacity;
  it.off = off;
  it.bin_key = bin_key;
  it.bin_pix = bin_pix;
  it.head = head;
  it.frag_z = frag_z;
  const int64_t items =
      it.overflow ? int64_t(job.num_triangles) : int64_t(cnt);
  const int64_t P = int64_t(W) * H;
  if (cnt == 0) {
    // Empty tile (most of a frame): background everywhere, no LDS.
    for (int p = threadIdx.x; p < TP; p += kThreads) {
      const int ix = x0 + p % TW, iy = y0 + p / TW;
      if (ix >= x1 || iy >= y1) continue;
      const int64_t pix = int64_t(iy) * W + ix;
      for (int l = 0; l < L; ++l) {
        const int64_t o = pix + l * P;
        gst(job.triangle_ids + o, -1);
        if (!sparse) {
          gst(job.z_buffer + o, 1.0f);
          if (job.barycentrics) {
            gst(job.barycentrics + 3 * o + 0, 0.0f);
            gst(job.barycentrics + 3 * o + 1, 0.0f);
            gst(job.barycentrics + 3 * o + 2, 0.0f);
          }
        }
        if (prefix) break;
      }
    }
    return;
  }
  for (int i = threadIdx.x; i < L * TP; i += kThreads) keys[i] = kEmpty;
  for (int i = threadIdx.x; i < (TP + 31) / 32; i += kThreads) ties[i] = 0u;

  for (int64_t base = 0; base < items; base += kThreads) {
    if (threadIdx.x == 0) queue_n = 0;
    __syncthreads();
    const int64_t i = base + threadIdx.x;
    bool wave_job = false;
    TriRect rect;
    if (i < items) {
      int t = -1;  // a whole triangle to walk
      if (!it.overflow) {
        const uint64_t key = bin_key[off + i];
        if (uint32_t(key >> 32) != kTriMarker)
          insert_key(keys, ties, TP, int(bin_pix[off + i]), L, key);
        else
          t = int(uint32_t(key));
      } else {
        // Overflowed bins: every triangle of the job, through its record
        // (fragments by pixel; whole triangles by their tile range).
        const int64_t flat = job.triangle_begin + i;
        const uint2 h = head[flat];
        if (h.x & kHeadFrag) {
          const int n = int(h.x & 7u);
          for (int k = 0; k < n; ++k) {
            const uint32_t o = (h.x >> (3 + 4 * k)) & 15u;
            const int px = int(h.y & 0xffffu) + int(o & 3u);
            const int py = int(h.y >> 16) + int(o >> 2);
            if (px < x0 || px >= x1 || py < y0 || py >= y1) continue;
            const float z = __uint_as_float(frag_z[8 * flat + k]);
            insert_key(keys, ties, TP, (py - y0) * TW + (px - x0), L,
                       (uint64_t(z_order(z)) << 32) | uint32_t(i));
          }
        } else if (h.x & kHeadRecompute) {
          t = int(i);
        } else if (h.x & kHeadTri) {
          const int64_t d = flat_tile - int64_t(h.y);
          const int64_t tpr = int64_t((h.x >> 16) & 8191u);
          if (d >= 0 && d / tpr < int64_t((h.x >> 8) & 255u) &&
              d % tpr < int64_t(h.x & 255u))
            t = int(i);
        }
      }
      float4 v[3];
      Box b;
      if (t >= 0 && load_triangle(job, t, v) && triangle_bbox(v, W, H, b)) {
        rect.tri = t;
        rect.x0 = short(max(b.left, x0));
        rect.x1 = short(min(b.right, x1));
        rect.y0 = short(max(b.bottom, y0));
        rect.y1 = short(min(b.top, y1));
        const int area = max(0, rect.x1 - rect.x0) * max(0, rect.y1 - rect.y0);
        if (area > kWaveRasterArea) {
          wave_job = true;
        } else if (area > 0) {
          int64_t r[9];
          edge_setup(v, cull, r);
          with_edges(r, [&](const auto* e) {
            for (int iy = rect.y0; iy < rect.y1; ++iy)
              for (int ix = rect.x0; ix < rect.x1; ++ix) {
                Fragment f;
                if (shade_pixel(ix, iy, sx2, sy2, e, v, f))
                  insert_key(keys, ties, TP, (iy - y0) * TW + (ix - x0), L,
                             (uint64_t(z_order(f.z)) << 32) | uint32_t(t));
              }
          });
        }
      }
    }
    // Queue large triangles: one LDS atomic per wave, ranks from the ballot.
    const uint64_t mask = __ballot(wave_job);
    if (mask) {
      const int leader = __ffsll(static_cast<unsigned long long>(mask)) - 1;
      int qbase = 0;
      if (lane_id() == leader) qbase = atomicAdd(&queue_n, __popcll(mask));
      qbase = __builtin_amdgcn_readlane(qbase, leader);
      if (wave_job) queue[qbase + __popcll(mask & lanemask_lt())] = rect;
    }
    __syncthreads();
    // Each wave walks whole queued triangles, one pixel per lane.
    const int nq = queue_n;
    for (int q = threadIdx.x / 64; q < nq; q += kThreads / 64) {
      const TriRect rq = queue[q];
      float4 v[3];
      load_triangle(job, rq.tri, v);
      int64_t r[9];
      edge_setup(v, cull, r);
      const int rw = rq.x1 - rq.x0;
      const int area = rw * (rq.y1 - rq.y0);
      with_edges(r, [&](const auto* e) {
        for (int k = lane_id(); k < area; k += 64) {
          const int ix = rq.x0 + k % rw, iy = rq.y0 + k / rw;
          Fragment f;
          if (shade_pixel(ix, iy, sx2, sy2, e, v, f))
            insert_key(keys, ties, TP, (iy - y0) * TW + (ix - x0), L,
                       (uint64_t(z_order(f.z)) << 32) | uint32_t(rq.tri));
        }
      });
    }
    __syncthreads();
  }
  __syncthreads();

  // Exact replay of the reference's insertion order at pixels with z ties
  // (block-uniform loop: the flags are read from LDS by every thread).
  for (int w = 0; w < (TP + 31) / 32; ++w) {
    uint32_t bits = ties[w];
    while (bits) {
      const int p = w * 32 + __builtin_ctz(bits);
      bits &= bits - 1;
      const int ix = x0 + p % TW, iy = y0 + p / TW;
      if (ix >= x1 || iy >= y1) continue;
      replay_pixel(job, it, items, ix, iy, p, sx2, sy2, cull, L, keys + p, TP,
                   cand, &cand_n, replay_z, replay_id);
    }
  }

  // Resolve: recompute the winning fragments exactly and write the outputs.
  for (int p = threadIdx.x; p < TP; p += kThreads) {
    const int ix = x0 + p % TW, iy = y0 + p / TW;
    if (ix >= x1 || iy >= y1) continue;
    const int64_t pix = int64_t(iy) * W + ix;
    for (int l = 0; l < L; ++l) {
      const uint64_t key = keys[l * TP + p];
      const int64_t o = pix + l * P;
      int id = -1;
      Fragment f{0.0f, 0.0f, 0.0f, 1.0f};
      if (key != kEmpty) {
        id = int(uint32_t(key));
        float4 v[3];
        load_triangle(job, id, v);
        int64_t r[9];
        edge_setup(v, cull, r);
        with_edges(r, [&](const auto* e) {
          shade_pixel(ix, iy, sx2, sy2, e, v, f);
        });
      }
      gst(job.triangle_ids + o, id);
      if (!sparse || id >= 0) {
        gst(job.z_buffer + o, f.z);
        if (job.barycentrics) {
          gst(job.barycentrics + 3 * o + 0, f.b0);
          gst(job.barycentrics + 3 * o + 1, f.b1);
          gst(job.barycentrics + 3 * o + 2, f.b2);
        }
      }
      // The key lists fill front to back: no fragment behind an empty layer.
      if (prefix && id < 0) break;
    }
  }
}

// ---------------------------------------------------------------------------
// Host side.

struct Workspace {
  uint32_t* counts;
  uint32_t* offsets;
  uint32_t* block_sums;
  uint32_t* status;
  uint64_t* bin_key;   // [capacity] fragment keys / whole-triangle entries
  uint16_t* bin_pix;   // [capacity] a fragment's pixel in its tile
  uint2* head;         // [T] count-pass record heads
  uint4* frag_z;       // [T][2] the records' fragment z
  dr_raster_job* job;  // the single-job entry point's one-entry table
  size_t bytes;
};

inline size_t align256(size_t x) { return (x + 255) & ~size_t(255); }

inline bool valid_tile(int tw, int th) {
  return (tw == 8 && th == 8) || (tw == 16 && th == 16) ||
         (tw == 32 && th == 8) || (tw == 32 && th == 16) ||
         (tw == 32 && th == 32) || (tw == 64 && th == 8);
}

inline int64_t scan_blocks(int64_t total_tiles) {
  return (total_tiles + kScanTile - 1) / kScanTile;
}

Workspace carve(void* base, int64_t total_tiles, uint64_t capacity,
               int64_t total_tris) {
  Workspace w;
  char* p = static_cast<char*>(base);
  size_t o = 0;
  w.status = reinterpret_cast<uint32_t*>(p + o);
  o += align256(4 * sizeof(uint32_t));
  w.counts = reinterpret_cast<uint32_t*>(p + o);
  o += align256(sizeof(uint32_t) * total_tiles);
  w.offsets = reinterpret_cast<uint32_t*>(p + o);
  o += align256(sizeof(uint32_t) * total_tiles);
  w.block_sums = reinterpret_cast<uint32_t*>(p + o);
  o += align256(sizeof(uint32_t) * scan_blocks(total_tiles));
  const size_t nt = size_t(total_tris > 0 ? total_tris : 1);
  w.bin_key = reinterpret_cast<uint64_t*>(p + o);
  o += align256(sizeof(uint64_t) * capacity);
  w.bin_pix = reinterpret_cast<uint16_t*>(p + o);
  o += align256(sizeof(uint16_t) * capacity);
  w.head = reinterpret_cast<uint2*>(p + o);
  o += align256(sizeof(uint2) * nt);
  w.frag_z = reinterpret_cast<uint4*>(p + o);
  o += align256(2 * sizeof(uint4) * nt);
  w.job = reinterpret_cast<dr_raster_job*>(p + o);
  o += align256(sizeof(dr_raster_job));
  w.bytes = o;
  return w;
}

uint64_t capacity_of(const dr_raster_config& c, int64_t total_tris,
                     int64_t total_tiles) {
  if (c.bin_capacity > 0) return uint64_t(c.bin_capacity);
  return uint64_t(3 * total_tris + 16 * total_tiles + 4096);
}

template <int TW, int TH>
int launch_all(const Jobs& J, int64_t total_tris, int64_t total_tiles, int L,
               int cull, const Workspace& w, uint64_t capacity, int flags,
               int tile_threads, hipStream_t s) {
  // The two halves can run as separate calls on the same workspace (each
  // call re-reads the job table): binning, then the tile pass.
  const bool run_bins = (flags & DR_RASTER_TILES_ONLY) == 0;
  const bool run_tiles = (flags & DR_RASTER_BINS_ONLY) == 0;
  if (run_bins) {
    const int64_t clear_n = total_tiles;
    clear_kernel<<<dim3(unsigned((clear_n + kBlock - 1) / kBlock)), kBlock, 0,
                   s>>>(w.counts, clear_n);
    DR_CHECK_LAUNCH("raster clear");
    const unsigned tri_blocks = unsigned((total_tris + kBlock - 1) / kBlock);
    if (tri_blocks > 0) {
      bin_count_kernel<TW, TH><<<tri_blocks, kBlock, 0, s>>>(
          J, total_tris, cull, w.counts, w.head, w.frag_z);
      DR_CHECK_LAUNCH("raster bin count");
    }
    const int64_t sb = scan_blocks(total_tiles);
    scan_local_kernel<<<unsigned(sb), kBlock, 0, s>>>(w.counts, w.offsets,
                                                     w.block_sums, total_tiles);
    DR_CHECK_LAUNCH("raster scan");
    scan_add_kernel<<<unsigned(sb), kBlock, 0, s>>>(
        w.offsets, w.block_sums, int(sb), total_tiles, w.status, capacity);
    DR_CHECK_LAUNCH("raster scan add");
    // The fill pass reuses counts (zeroed by the local scan) as cursors: the
    // tile kernel needs the final counts, which equal the cursors after it.
    if (tri_blocks > 0) {
      bin_fill_kernel<TW, TH><<<tri_blocks, kBlock, 0, s>>>(
          J, total_tris, w.counts, w.offsets, w.bin_key, w.bin_pix, capacity,
          w.head, w.frag_z);
      DR_CHECK_LAUNCH("raster bin fill");
    }
  }
  if (!run_tiles) return kOk;
  const size_t lds = sizeof(uint64_t) * L * TW * TH;
  auto launch = [&](auto kernel, int threads) -> int {
    if (lds > 32768) {
      // Deep key lists: dynamic LDS plus the kernel's static arrays may pass
      // the default 64 KiB launch limit.
      DR_CHECK_HIP(hipFuncSetAttribute(
          reinterpret_cast<const void*>(kernel),
          hipFuncAttributeMaxDynamicSharedMemorySize, int(lds)));
    }
    kernel<<<unsigned(total_tiles), threads, lds, s>>>(
        J, L, cull, w.counts, w.offsets, w.bin_key, w.bin_pix, w.head,
        reinterpret_cast<const uint32_t*>(w.frag_z), capacity,
        (flags & DR_RASTER_SPARSE) != 0, (flags & DR_RASTER_PREFIX) != 0);
    DR_CHECK_LAUNCH("raster tiles");
    return kOk;
  };
  return tile_threads == 128 ? launch(&raster_tile_kernel<TW, TH, 128>, 128)
                             : launch(&raster_tile_kernel<TW, TH, 256>, 256);
}

// `table` is a device table of `count` jobs, or null with `single` the one
// job of dr_rasterize_triangles (stored into the workspace here).
int dispatch(const dr_raster_job* table, const dr_raster_job* single,
             int count, int64_t total_tris, int64_t total_tiles, int L,
             int cull, const dr_raster_config& c, void* ws, size_t ws_bytes,
             hipStream_t s) {
  DR_CHECK_ARG(L >= 1 && L <= kMaxLayers, "num_layers must be in [1, 64]");
  DR_CHECK_ARG(cull >= 0 && cull <= 2, "face_culling_mode must be 0, 1 or 2");
  DR_CHECK_ARG((c.flags & DR_RASTER_BINS_ONLY) == 0 ||
                   (c.flags & DR_RASTER_TILES_ONLY) == 0,
               "DR_RASTER_BINS_ONLY and DR_RASTER_TILES_ONLY exclude each "
               "other");
  DR_CHECK_ARG(valid_tile(c.tile_w, c.tile_h), "unsupported tile size");
  DR_CHECK_ARG(c.tile_threads == 0 || c.tile_threads == 128 ||
                   c.tile_threads == 256,
               "tile_threads must be 0, 128 or 256");
  DR_CHECK_ARG(total_tiles > 0 && total_tiles < (int64_t(1) << 31),
               "tile count out of range");
  DR_CHECK_ARG(total_tris >= 0 && total_tris < (int64_t(1) << 40),
               "triangle count out of range");
  DR_CHECK_ARG(size_t(L) * c.tile_w * c.tile_h * sizeof(uint64_t) <= 65536,
               "num_layers x tile pixels exceeds the 64 KiB LDS key budget");
  const uint64_t capacity = capacity_of(c, total_tris, total_tiles);
  DR_CHECK_ARG(capacity < (uint64_t(1) << 32), "bin capacity must be < 2^32");
  DR_CHECK_ARG(ws != nullptr, "workspace is null");
  DR_CHECK_ARG((reinterpret_cast<uintptr_t>(ws) & 255) == 0,
               "workspace must be 256-byte aligned");
  const Workspace w = carve(ws, total_tiles, capacity, total_tris);
  if (ws_bytes < w.bytes)
    return fail(kWorkspaceTooSmall,
                "workspace too small: need " + std::to_string(w.bytes) +
                    " bytes, got " + std::to_string(ws_bytes));
  if (table == nullptr) {
    store_job_kernel<<<1, 1, 0, s>>>(*single, w.job);
    DR_CHECK_LAUNCH("raster store job");
    table = w.job;
  }
  const Jobs J{table, count};
#define DR_TILE_CASE(A, B)                                              \
  if (c.tile_w == A && c.tile_h == B)                                   \
    return launch_all<A, B>(J, total_tris, total_tiles, L, cull, w,     \
                            capacity, c.flags, c.tile_threads, s);
  DR_TILE_CASE(8, 8)
  DR_TILE_CASE(16, 16)
  DR_TILE_CASE(32, 8)
  DR_TILE_CASE(32, 16)
  DR_TILE_CASE(32, 32)
  DR_TILE_CASE(64, 8)
#undef DR_TILE_CASE
  return fail(kUnsupported, "unsupported tile size");
}

// Workspace of dr_rasterize_triangles_noalloc: a file-scope device array,
// one copy per GPU, held until the process ends.  In host code `scratch`
// names a host-side stand-in; the device address of each GPU's copy comes
// from hipGetSymbolAddress, looked up once per device with that device
// current, and only that address is handed to kernels.
constexpr size_t kStaticScratchBytes = DR_STATIC_SCRATCH_BYTES;
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
void* g_scratch_ptr[kMaxDevices] = {};

}  // namespace
}  // namespace dr

static __device__ __align__(256) unsigned char scratch[DR_STATIC_SCRATCH_BYTES];

namespace dr {
namespace {

// Device address of the current device's `scratch`.
int static_scratch(void** out) {
  int dev = -1;
  DR_CHECK_HIP(hipGetDevice(&dev));
  DR_CHECK_ARG(dev >= 0 && dev < kMaxDevices, "device index out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (g_scratch_ptr[dev] == nullptr) {
    void* p = nullptr;
    DR_CHECK_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    DR_CHECK_ARG(p != nullptr && (reinterpret_cast<uintptr_t>(p) & 255) == 0,
                 "static scratch has no aligned device address");
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return kOk;
}

dr_raster_config resolve_config(const dr_raster_config* c) {
  dr_raster_config r{32, 8, 0, 0, 0};
  if (c) r = *c;
  return r;
}

}  // namespace
}  // namespace dr

extern "C" {

int dr_raster_jobs_prepare(dr_raster_job* jobs, int32_t n,
                           const dr_raster_config* config,
                           int64_t* total_triangles, int64_t* total_tiles) {
  const dr_raster_config c = dr::resolve_config(config);
  DR_CHECK_ARG(jobs != nullptr && n > 0, "need at least one job");
  DR_CHECK_ARG(dr::valid_tile(c.tile_w, c.tile_h), "unsupported tile size");
  DR_CHECK_ARG(c.tile_threads == 0 || c.tile_threads == 128 ||
                   c.tile_threads == 256,
               "tile_threads must be 0, 128 or 256");
  int64_t tri = 0, tiles = 0;
  for (int i = 0; i < n; ++i) {
    dr_raster_job& j = jobs[i];
    DR_CHECK_ARG(j.width > 0 && j.height > 0, "image size must be > 0");
    DR_CHECK_ARG(j.width <= 32767 && j.height <= 32767,
                 "image size must be <= 32767");
    DR_CHECK_ARG(j.num_triangles >= 0 && j.num_vertices >= 0,
                 "negative mesh size");
    DR_CHECK_ARG(j.num_triangles == 0 || (j.vertices && j.triangles),
                 "null mesh pointer");
    DR_CHECK_ARG(j.triangle_ids && j.z_buffer, "null output pointer");
    DR_CHECK_ARG(j.camera != nullptr ||
                     (reinterpret_cast<uintptr_t>(j.vertices) & 15) == 0,
                 "clip-space vertices must be 16-byte aligned");
    j.triangle_begin = tri;
    j.tile_begin = tiles;
    tri += j.num_triangles;
    tiles += int64_t((j.width + c.tile_w - 1) / c.tile_w) *
             ((j.height + c.tile_h - 1) / c.tile_h);
  }
  if (total_triangles) *total_triangles = tri;
  if (total_tiles) *total_tiles = tiles;
  return DR_OK;
}

size_t dr_raster_workspace_size(int64_t total_triangles, int64_t total_tiles,
                                const dr_raster_config* config) {
  const dr_raster_config c = dr::resolve_config(config);
  return dr::carve(nullptr, total_tiles,
                   dr::capacity_of(c, total_triangles, total_tiles),
                   total_triangles)
      .bytes;
}

int dr_rasterize_triangles(const float* vertices, int32_t num_vertices,
                           const int32_t* triangles, int32_t num_triangles,
                           int32_t width, int32_t height, int32_t num_layers,
                           int32_t face_culling_mode, int32_t* triangle_ids,
                           float* z_buffer, float* barycentrics,
                           const dr_raster_config* config, void* workspace,
                           size_t workspace_bytes, dr_stream_t stream) {
  const dr_raster_config c = dr::resolve_config(config);
  dr_raster_job job{};
  job.vertices = vertices;
  job.triangles = triangles;
  job.triangle_ids = triangle_ids;
  job.z_buffer = z_buffer;
  job.barycentrics = barycentrics;
  job.num_vertices = num_vertices;
  job.num_triangles = num_triangles;
  job.width = width;
  job.height = height;
  int64_t tt = 0, tiles = 0;
  const int rc = dr_raster_jobs_prepare(&job, 1, &c, &tt, &tiles);
  if (rc != DR_OK) return rc;
  return dr::dispatch(nullptr, &job, 1, tt, tiles, num_layers,
                      face_culling_mode, c, workspace, workspace_bytes,
                      static_cast<hipStream_t>(stream));
}

int dr_rasterize_triangles_noalloc(const float* vertices, int32_t num_vertices,
                                   const int32_t* triangles,
                                   int32_t num_triangles, int32_t width,
                                   int32_t height, int32_t num_layers,
                                   int32_t face_culling_mode,
                                   int32_t* triangle_ids, float* z_buffer,
                                   float* barycentrics,
                                   const dr_raster_config* config,
                                   dr_stream_t stream) {
  const dr_raster_config c = dr::resolve_config(config);
  dr_raster_job job{};
  job.vertices = vertices;
  job.triangles = triangles;
  job.triangle_ids = triangle_ids;
  job.z_buffer = z_buffer;
  job.barycentrics = barycentrics;
  job.num_vertices = num_vertices;
  job.num_triangles = num_triangles;
  job.width = width;
  job.height = height;
  int64_t tt = 0, tiles = 0;
  int rc = dr_raster_jobs_prepare(&job, 1, &c, &tt, &tiles);
  if (rc != DR_OK) return rc;
  const size_t need = dr_raster_workspace_size(tt, tiles, &c);
  if (need > dr::kStaticScratchBytes)
    return dr::fail(dr::kWorkspaceTooSmall,
                    "dr_rasterize_triangles_noalloc: this call needs " +
                        std::to_string(need) +
                        " bytes of scratch, above the static scratch of " +
                        std::to_string(dr::kStaticScratchBytes) +
                        " bytes; use dr_rasterize_triangles with a workspace");
  void* ws = nullptr;
  rc = dr::static_scratch(&ws);
  if (rc != DR_OK) return rc;
  return dr::dispatch(nullptr, &job, 1, tt, tiles, num_layers,
                      face_culling_mode, c, ws, dr::kStaticScratchBytes,
                      static_cast<hipStream_t>(stream));
}

int dr_rasterize_triangles_batched(const dr_raster_job* jobs_device,
                                   int32_t num_jobs, int64_t total_triangles,
                                   int64_t total_tiles, int32_t num_layers,
                                   int32_t face_culling_mode,
                                   const dr_raster_config* config,
                                   void* workspace, size_t workspace_bytes,
                                   dr_stream_t stream) {
  const dr_raster_config c = dr::resolve_config(config);
  DR_CHECK_ARG(jobs_device != nullptr && num_jobs > 0, "need jobs");
  return dr::dispatch(jobs_device, nullptr, num_jobs, total_triangles,
                      total_tiles, num_layers, face_culling_mode, c,
                      workspace, workspace_bytes,
                      static_cast<hipStream_t>(stream));
}

int dr_raster_bin_stats(const void* workspace, int64_t* entries_needed,
                        int32_t* overflowed, dr_stream_t stream) {
  DR_CHECK_ARG(workspace != nullptr, "workspace is null");
  uint32_t st[4] = {0, 0, 0, 0};
  DR_CHECK_HIP(hipMemcpyAsync(st, workspace, sizeof(st), hipMemcpyDeviceToHost,
                              static_cast<hipStream_t>(stream)));
  DR_CHECK_HIP(hipStreamSynchronize(static_cast<hipStream_t>(stream)));
  if (entries_needed) *entries_needed = int64_t(st[0]) | (int64_t(st[2]) << 32);
  if (overflowed) *overflowed = int32_t(st[1]);
  return DR_OK;
}

}  // extern "C"
