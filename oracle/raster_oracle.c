/*
 * TEST INFRASTRUCTURE ONLY -- the CPU oracle for the layered z-buffer
 * rasterizer.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load this; the product path never calls it.
 *
 * A plain-C restatement of diffren's CPU rasterizer:
 *   diffren/common/kernels/rasterize_triangles_impl.cc:27-118   (main loop)
 *   diffren/common/kernels/rasterize_utils.h:51-248             (math helpers)
 *   diffren/jax/internal/kernels/rasterize_triangles_cpu.cc:55-60 (output init)
 *
 * Semantics that matter for bit-exactness and how they are pinned here:
 *   - FP32 arithmetic with no contraction: this file must be compiled with
 *     -ffp-contract=off (oracle/Makefile does).
 *   - Float-to-integer conversions reproduce what the reference does on
 *     x86-64 (cvttss2si): truncation toward zero, and the "integer
 *     indefinite" value (INT_MIN / INT64_MIN) for NaN or out-of-range input.
 *     The C standard leaves those cases undefined, so they are spelled out.
 *   - Edge functions are int64 with two's-complement wrap (done in uint64 to
 *     keep the C well defined).
 *
 * Parity pin: tests/test_oracle_raster.py compares this file bit-for-bit
 * with the reference source compiled from its own files (oracle/_ref) on the
 * reference's C++ test scenes, seeded random scenes with adversarial vertices
 * (near-plane crossings, tiny/negative w, huge coordinates), and against the
 * committed golden vectors in tests/golden/.
 */
#include <math.h>
#include <stdint.h>
#include <string.h>

#define FRACTIONAL_BITS 16

/* x86-64 cvttss2si semantics, 32-bit destination. */
static int f32_to_i32_x86(float f) {
  if (f >= -2147483648.0f && f < 2147483648.0f) return (int)f;
  return INT32_MIN;
}

/* x86-64 cvttss2si semantics, 64-bit destination. */
static int64_t f32_to_i64_x86(float f) {
  if (f >= -9223372036854775808.0f && f < 9223372036854775808.0f)
    return (int64_t)f;
  return INT64_MIN;
}

/* int64 arithmetic with wrap-around (the reference relies on the hardware
 * behaviour of imul/add). */
static int64_t wmul(int64_t a, int64_t b) {
  return (int64_t)((uint64_t)a * (uint64_t)b);
}
static int64_t wadd(int64_t a, int64_t b) {
  return (int64_t)((uint64_t)a + (uint64_t)b);
}
static int64_t wsub(int64_t a, int64_t b) {
  return (int64_t)((uint64_t)a - (uint64_t)b);
}

/* rasterize_utils.h:57-59: f * 2^16 in float, truncated to int64. */
static int64_t to_fixed(float f) { return f32_to_i64_x86(f * 65536.0f); }

static int clampi(int v, int lo, int hi) {
  return v < lo ? lo : (v > hi ? hi : v);
}

/* rasterize_utils.h:63-73 */
static int clamped_floor_min(float a, float b, int lo, int hi) {
  return clampi(f32_to_i32_x86(floorf(fminf(a, b))), lo, hi);
}
static int clamped_ceil_max(float a, float b, int lo, int hi) {
  return clampi(f32_to_i32_x86(ceilf(fmaxf(a, b))), lo, hi);
}

typedef struct {
  int left, right, bottom, top;
} bbox_t;

/* rasterize_utils.h:75-84: project to pixel units and grow the box. */
static void grow(float x, float y, float w, int W, int H, bbox_t* b) {
  const float sx = 0.5f * (x / w + 1.0f) * (float)W;
  const float sy = 0.5f * (y / w + 1.0f) * (float)H;
  b->left = clamped_floor_min((float)b->left, sx, 0, W);
  b->right = clamped_ceil_max((float)b->right, sx, 0, W);
  b->bottom = clamped_floor_min((float)b->bottom, sy, 0, H);
  b->top = clamped_ceil_max((float)b->top, sy, 0, H);
}

/* rasterize_utils.h:86-99: grow by the point where edge p->q meets the near
 * plane z/w == -1. */
static void grow_near(const float* p, const float* q, int W, int H,
                      bbox_t* b) {
  const float den = p[2] - q[2] + p[3] - q[3];
  if (den != 0.0f) {
    const float t = (p[2] + p[3]) / den;
    const float x = p[0] + t * (q[0] - p[0]);
    const float y = p[1] + t * (q[1] - p[1]);
    const float w = p[3] + t * (q[3] - p[3]);
    grow(x, y, w, W, H, b);
  }
}

/* rasterize_utils.h:103-105 */
static int in_front(const float* v) { return v[3] > 0.0f && v[2] >= -v[3]; }

/* rasterize_utils.h:111-161.  Returns non-zero when the box is non-empty. */
static int triangle_bbox(const float* v[3], int W, int H, bbox_t* b) {
  b->left = W;
  b->right = 0;
  b->bottom = H;
  b->top = 0;
  const int vis[3] = {in_front(v[0]), in_front(v[1]), in_front(v[2])};
  for (int i = 0; i < 3; ++i) {
    if (!vis[i]) continue;
    const int j = (i + 1) % 3, k = (i + 2) % 3;
    grow(v[i][0], v[i][1], v[i][3], W, H, b);
    if (!vis[j]) grow_near(v[i], v[j], W, H, b);
    if (!vis[k]) grow_near(v[i], v[k], W, H, b);
  }
  return b->right > b->left && b->top > b->bottom;
}

/* rasterize_utils.h:176-206: adjugate of [[x0 x1 x2][y0 y1 y2][w0 w1 w2]]
 * (fixed point), with the culling sign applied.  mode: 0 none, 1 back,
 * 2 front (rasterize_triangles_types.h:24). */
static void adjugate(const int64_t m[9], int mode, int64_t r[9]) {
  /* m = [a11 a12 a13; a21 a22 a23; a31 a32 a33] row-major. */
  r[0] = wsub(wmul(m[4], m[8]), wmul(m[7], m[5]));
  r[1] = wsub(wmul(m[2], m[7]), wmul(m[8], m[1]));
  r[2] = wsub(wmul(m[1], m[5]), wmul(m[4], m[2]));
  r[3] = wsub(wmul(m[5], m[6]), wmul(m[8], m[3]));
  r[4] = wsub(wmul(m[0], m[8]), wmul(m[6], m[2]));
  r[5] = wsub(wmul(m[2], m[3]), wmul(m[5], m[0]));
  r[6] = wsub(wmul(m[3], m[7]), wmul(m[6], m[4]));
  r[7] = wsub(wmul(m[1], m[6]), wmul(m[7], m[0]));
  r[8] = wsub(wmul(m[0], m[4]), wmul(m[3], m[1]));
  if (mode == 0 || mode == 2) {
    const int64_t det_i =
        wadd(wadd(wmul(m[0], r[0]), wmul(m[1], r[3])), wmul(m[2], r[6]));
    const float det = (float)det_i;
    const float s = (mode == 0) ? copysignf(1.0f, det) : -1.0f;
    /* The reference multiplies each int64 entry by a float, which converts
     * the entry to float, multiplies, and truncates back to int64. */
    for (int i = 0; i < 9; ++i) r[i] = f32_to_i64_x86((float)r[i] * s);
  }
}

/* rasterize_utils.h:220-238 */
static void edge_values(int ix, int iy, int W, int H, const int64_t r[9],
                        int64_t e[3]) {
  const float px = (2.0f / (float)W) * ((float)ix + 0.5f) - 1.0f;
  const float py = (2.0f / (float)H) * ((float)iy + 0.5f) - 1.0f;
  const int64_t pxi = to_fixed(px), pyi = to_fixed(py);
  for (int i = 0; i < 3; ++i) {
    e[i] = wadd(wadd(wmul(r[3 * i], pxi), wmul(r[3 * i + 1], pyi)),
                (int64_t)((uint64_t)r[3 * i + 2] << FRACTIONAL_BITS));
  }
}

/* rasterize_utils.h:243-248 */
static int inside(const int64_t e[3]) {
  return (e[0] >= 0 && e[1] >= 0 && e[2] >= 0) &&
         (e[0] > 0 || e[1] > 0 || e[2] > 0);
}

/*
 * Fills ids/z/barycentrics exactly as the reference's CPU FFI handler plus
 * RasterizeTrianglesImpl do: ids = -1, z = 1, barycentrics = 0 on background
 * (rasterize_triangles_cpu.cc:55-60), then the per-triangle loop
 * (rasterize_triangles_impl.cc:37-116).  barycentrics may be NULL.
 */
void oracle_rasterize(const float* vertices, const int32_t* triangles,
                      int32_t triangle_count, int32_t W, int32_t H,
                      int32_t L, int32_t culling_mode, int32_t* ids,
                      float* zbuf, float* barys) {
  const int64_t P = (int64_t)W * H;
  for (int64_t i = 0; i < P * L; ++i) {
    ids[i] = -1;
    zbuf[i] = 1.0f;
  }
  if (barys) memset(barys, 0, sizeof(float) * 3 * P * L);

  for (int32_t t = 0; t < triangle_count; ++t) {
    const float* v[3];
    for (int k = 0; k < 3; ++k) v[k] = vertices + 4 * triangles[3 * t + k];
    bbox_t bb;
    if (!triangle_bbox(v, W, H, &bb)) continue;
    const int64_t m[9] = {to_fixed(v[0][0]), to_fixed(v[1][0]),
                          to_fixed(v[2][0]), to_fixed(v[0][1]),
                          to_fixed(v[1][1]), to_fixed(v[2][1]),
                          to_fixed(v[0][3]), to_fixed(v[1][3]),
                          to_fixed(v[2][3])};
    int64_t r[9];
    adjugate(m, culling_mode, r);
    for (int iy = bb.bottom; iy < bb.top; ++iy) {
      for (int ix = bb.left; ix < bb.right; ++ix) {
        int64_t e[3];
        edge_values(ix, iy, W, H, r, e);
        if (!inside(e)) continue;
        /* rasterize_triangles_impl.cc:79-89 */
        const float sum = (float)wadd(wadd(e[0], e[1]), e[2]);
        const float b0 = (float)e[0] / sum;
        const float b1 = (float)e[1] / sum;
        const float b2 = (float)e[2] / sum;
        const float cz = b0 * v[0][2] + b1 * v[1][2] + b2 * v[2][2];
        const float cw = b0 * v[0][3] + b1 * v[1][3] + b2 * v[2][3];
        const float z = cz / cw;
        if (z < -1.0f || z > 1.0f) continue; /* impl.cc:92 */
        /* impl.cc:95-114: L-deep insertion with strict '<'. */
        float zc = z, b0c = b0, b1c = b1, b2c = b2;
        int32_t idc = t;
        const int64_t p0 = (int64_t)iy * W + ix;
        for (int l = 0; l < L; ++l) {
          const int64_t p = p0 + P * l;
          if (zc < zbuf[p]) {
            float tz = zbuf[p];
            zbuf[p] = zc;
            zc = tz;
            int32_t ti = ids[p];
            ids[p] = idc;
            idc = ti;
            if (barys) {
              float tb;
              tb = barys[3 * p + 0]; barys[3 * p + 0] = b0c; b0c = tb;
              tb = barys[3 * p + 1]; barys[3 * p + 1] = b1c; b1c = tb;
              tb = barys[3 * p + 2]; barys[3 * p + 2] = b2c; b2c = tb;
            }
          }
          if (zc == 1.0f) break;
        }
      }
    }
  }
}
