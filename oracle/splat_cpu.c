/* TEST INFRASTRUCTURE ONLY: a float32 C restatement, on host threads, of the
 * rasterize-then-splat render for one view given its rasterization, its L2
 * loss against a target image, and the hand-derived adjoint of both.  It is
 * the CPU baseline's splat leg and a fast float32 checker for full-size
 * frames; nothing under diffren_amd/ loads it.
 *
 * Forward, in the reference's float32 expression order:
 *   clip      = M [v, 1]                       transforms.py:40-78
 *   interpolate clip (background 0,0,1,1) and colour (background 0)
 *             = m (b0 A0 + b1 A1 + b2 A2) + (1 - m) bg   interpolate.py:24-63
 *   viewport  = (xyz / w + 1) [W, H, 1] 0.5    splat.py:218-221
 *   splat     splat.py:26-175: for layer l, tap (u, v): offset
 *             floor(xy) + 0.5 - xy + (u, v), weight m coef exp(-2 |d|^2),
 *             surface index from the target's top z and the source's top z
 *             (argmin, first index on ties), fg / surface / bg accumulators,
 *             crop, 1 / (max(w - 1, 0) + 1) normalisation
 *   over      composite.py:20-37 of the three normalised buffers.
 * The reference pushes every source into padded accumulators in the loop
 * order (l, u, v).  Here each target pulls its sources in that same order,
 * so every accumulator receives the same float32 additions in the same
 * order (the pull parallelises over rows without atomics).
 *
 * Backward: loss = loss_scale sum (img - target)^2; the adjoint of over,
 * the normalisation, the accumulators, the tap weights (dw/dd = -4 d w), the
 * pixel-centre offset (d offset / d xy = -1), the viewport and perspective
 * divide, the barycentric interpolation (rasterization held fixed,
 * rasterize_triangles_xla.py:51-52) and the camera transform.  Per-fragment
 * records are formed in parallel, then scattered to vertices serially in
 * pixel order.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

#define MAXC 32
#define MAXL 16

static float splat_coef(void) {
  /* splat.py:63-78: (1 + 0.05) / sum of the nine taps exp(-2 (u^2 + v^2)). */
  double total = 0.0;
  for (int u = -1; u <= 1; ++u)
    for (int v = -1; v <= 1; ++v) total += exp(-2.0 * (u * u + v * v));
  return (float)(1.05 / total);
}

typedef struct {
  int L, H, W, C;
  const int* ids;
  const float* barys;
  const int* tris;
  int T;
  const float* clip;    /* [V][4] */
  const float* colors;  /* [V][C] */
  float* vx;            /* [L][P] viewport x */
  float* vy;            /* [L][P] viewport y */
  float* vz;            /* [L][P] viewport z (1 on background) */
  float* col;           /* [L][P][C] interpolated colour */
  float* cw;            /* [L][P] interpolated clip w */
  float* cx;            /* [L][P] interpolated clip x */
  float* cy;            /* [L][P] interpolated clip y */
} Frame;

static int corner_vertex(const Frame* f, int id, int k) {
  /* vertex_ids = triangles[triangle_id], id -1 wrapping to the last
   * triangle (rasterize.py:111-113; its barycentrics are 0). */
  const int t = id < 0 ? id + f->T : id;
  return f->tris[3 * t + k];
}

/* Interpolation and viewport positions of every pixel-layer. */
static void prepare(Frame* f) {
  const int64_t P = (int64_t)f->H * f->W;
  const int C = f->C;
#pragma omp parallel for schedule(static)
  for (int64_t q = 0; q < (int64_t)f->L * P; ++q) {
    const int id = f->ids[q];
    const float* b = f->barys + 3 * q;
    const float m = id >= 0 ? 1.0f : 0.0f;
    const int v0 = corner_vertex(f, id, 0), v1 = corner_vertex(f, id, 1),
              v2 = corner_vertex(f, id, 2);
    static const float bgc[4] = {0.0f, 0.0f, 1.0f, 1.0f};
    float c4[4];
    for (int j = 0; j < 4; ++j) {
      const float s = (f->clip[4 * v0 + j] * b[0] + f->clip[4 * v1 + j] * b[1]) +
                      f->clip[4 * v2 + j] * b[2];
      c4[j] = m * s + (1.0f - m) * bgc[j];
    }
    f->cx[q] = c4[0];
    f->cy[q] = c4[1];
    f->cw[q] = c4[3];
    const float nx = c4[0] / c4[3], ny = c4[1] / c4[3], nz = c4[2] / c4[3];
    f->vx[q] = ((nx + 1.0f) * (float)f->W) * 0.5f;
    f->vy[q] = ((ny + 1.0f) * (float)f->H) * 0.5f;
    f->vz[q] = ((nz + 1.0f) * 1.0f) * 0.5f;
    for (int c = 0; c < C; ++c) {
      const float s = (f->colors[(int64_t)v0 * C + c] * b[0] +
                       f->colors[(int64_t)v1 * C + c] * b[1]) +
                      f->colors[(int64_t)v2 * C + c] * b[2];
      f->col[q * C + c] = m * s + (1.0f - m) * 0.0f;
    }
  }
}

/* Surface index of source pixel s seen from target pixel t
 * (splat.py:131-145).  Out-of-image pixels have z 1 on every layer. */
static int surface_index(const Frame* f, int64_t t, int64_t s) {
  const int64_t P = (int64_t)f->H * f->W;
  const float zt0 = f->vz[t], zs0 = f->vz[s];
  int c_idx = 0, t_idx = 0;
  float c_min = INFINITY, t_min = INFINITY;
  for (int l = 0; l < f->L; ++l) {
    const float dc = fabsf(f->vz[l * P + s] - zt0);
    if (dc < c_min) { c_min = dc; c_idx = l; }
    const float dt = fabsf(f->vz[l * P + t] - zs0);
    if (dt < t_min) { t_min = dt; t_idx = l; }
  }
  return t_min < c_min ? -t_idx : c_idx;
}

static int layer_class(int surface, int l) {
  return surface > l ? 0 : (surface == l ? 1 : 2);
}

static float tap_weight(const Frame* f, int64_t q, int u, int v, float coef,
                        float* dx, float* dy) {
  const float ox = (floorf(f->vx[q]) + 0.5f) - f->vx[q];
  const float oy = (floorf(f->vy[q]) + 0.5f) - f->vy[q];
  *dx = ox + (float)u;
  *dy = oy + (float)v;
  const float d2 = *dx * *dx + *dy * *dy;
  const float m = f->ids[q] >= 0 ? 1.0f : 0.0f;
  return (m * coef) * expf(-2.0f * d2);
}

/* Forward + loss + backward for one view.
 *   positions [V][3], colors [V][C], camera [4][4] row-major,
 *   triangles [T][3], ids [L][H][W], barys [L][H][W][3],
 *   target [H][W][C] (or NULL: no loss, no backward),
 *   image out [H][W][C]; d_positions [V][3], d_colors [V][C] (or NULL).
 * Returns the loss (0 without a target), or -1 on bad sizes. */
double splat_cpu_render(const float* positions, const float* colors,
                        const float* camera, int V, int C,
                        const int* triangles, int T, const int* ids,
                        const float* barys, int L, int H, int W,
                        const float* target, float loss_scale, float* image,
                        float* d_positions, float* d_colors, int threads) {
  if (C < 1 || C > MAXC || L < 1 || L > MAXL || H < 1 || W < 1 || V < 1 ||
      T < 1)
    return -1.0;
#ifdef _OPENMP
  if (threads > 0) omp_set_num_threads(threads);
#else
  (void)threads;
#endif
  const int64_t P = (int64_t)H * W, LP = (int64_t)L * P;
  const float coef = splat_coef();
  float* clip = malloc(sizeof(float) * 4 * (size_t)V);
  Frame f = {L, H, W, C, ids, barys, triangles, T, clip, colors};
  f.vx = malloc(sizeof(float) * LP);
  f.vy = malloc(sizeof(float) * LP);
  f.vz = malloc(sizeof(float) * LP);
  f.cx = malloc(sizeof(float) * LP);
  f.cy = malloc(sizeof(float) * LP);
  f.cw = malloc(sizeof(float) * LP);
  f.col = malloc(sizeof(float) * LP * C);
  /* [P][3 classes][C + 1]: gradients of the accumulated colour and weight. */
  const int K = C + 1;
  float* grad = target ? malloc(sizeof(float) * P * 3 * K) : NULL;

  /* clip = ((m0 x + m1 y) + m2 z) + m3 w per row. */
#pragma omp parallel for schedule(static)
  for (int i = 0; i < V; ++i)
    for (int r = 0; r < 4; ++r)
      clip[4 * i + r] = ((camera[4 * r] * positions[3 * i] +
                          camera[4 * r + 1] * positions[3 * i + 1]) +
                         camera[4 * r + 2] * positions[3 * i + 2]) +
                        camera[4 * r + 3] * 1.0f;
  prepare(&f);

  double loss = 0.0;
#pragma omp parallel for schedule(dynamic, 4) reduction(+ : loss)
  for (int y = 0; y < H; ++y) {
    for (int x = 0; x < W; ++x) {
      const int64_t t = (int64_t)y * W + x;
      float a[3][MAXC + 1];
      memset(a, 0, sizeof(a));
      for (int l = 0; l < L; ++l)
        for (int u = -1; u <= 1; ++u)
          for (int v = -1; v <= 1; ++v) {
            const int sx = x - u, sy = y - v;
            if (sx < 0 || sx >= W || sy < 0 || sy >= H) continue;
            const int64_t s = (int64_t)sy * W + sx, q = l * P + s;
            float dx, dy;
            const float w = tap_weight(&f, q, u, v, coef, &dx, &dy);
            const int k = layer_class(surface_index(&f, t, s), l);
            for (int c = 0; c < C; ++c) a[k][c] += w * f.col[q * C + c];
            a[k][C] += w;
          }
      float n[3][MAXC], sc[3];
      for (int k = 0; k < 3; ++k) {
        const float d = a[k][C] - 1.0f;
        sc[k] = 1.0f / ((d > 0.0f ? d : 0.0f) + 1.0f);
        for (int c = 0; c < C; ++c) n[k][c] = a[k][c] * sc[k];
      }
      /* over([fg, surface, bg]): o = n2; o = n1 + (1 - a1) o; o = n0 + ... */
      float o1[MAXC], out[MAXC];
      for (int c = 0; c < C; ++c)
        o1[c] = n[1][c] + (1.0f - n[1][C - 1]) * n[2][c];
      for (int c = 0; c < C; ++c) {
        out[c] = n[0][c] + (1.0f - n[0][C - 1]) * o1[c];
        image[t * C + c] = out[c];
      }
      if (!target) continue;
      float dout[MAXC], dn[3][MAXC], do1[MAXC];
      for (int c = 0; c < C; ++c) {
        const float e = out[c] - target[t * C + c];
        loss += (double)(e * e);
        dout[c] = 2.0f * e * loss_scale;
      }
      /* over backward */
      float da0 = 0.0f, da1 = 0.0f;
      for (int c = 0; c < C; ++c) {
        dn[0][c] = dout[c];
        da0 -= dout[c] * o1[c];
        do1[c] = (1.0f - n[0][C - 1]) * dout[c];
      }
      dn[0][C - 1] += da0;
      for (int c = 0; c < C; ++c) {
        dn[1][c] = do1[c];
        da1 -= do1[c] * n[2][c];
        dn[2][c] = (1.0f - n[1][C - 1]) * do1[c];
      }
      dn[1][C - 1] += da1;
      /* normalisation backward: n = a s, s = 1 / (max(w - 1, 0) + 1). */
      for (int k = 0; k < 3; ++k) {
        float ds = 0.0f;
        float* g = grad + (t * 3 + k) * K;
        for (int c = 0; c < C; ++c) {
          g[c] = dn[k][c] * sc[k];
          ds += dn[k][c] * a[k][c];
        }
        g[C] = a[k][C] - 1.0f >= 0.0f ? -ds * sc[k] * sc[k] : 0.0f;
      }
    }
  }
  if (target) loss *= (double)loss_scale;

  if (target && (d_positions || d_colors)) {
    /* Per pixel-layer record: d colour [C], d clip x, y, w. */
    const int RS = C + 3;
    float* rec = calloc((size_t)LP * RS, sizeof(float));
#pragma omp parallel for schedule(dynamic, 4)
    for (int y = 0; y < H; ++y)
      for (int x = 0; x < W; ++x)
        for (int l = 0; l < L; ++l) {
          const int64_t s = (int64_t)y * W + x, q = l * P + s;
          if (ids[q] < 0) continue;
          float* r = rec + q * RS;
          float gx = 0.0f, gy = 0.0f;
          for (int u = -1; u <= 1; ++u)
            for (int v = -1; v <= 1; ++v) {
              const int tx = x + u, ty = y + v;
              if (tx < 0 || tx >= W || ty < 0 || ty >= H) continue;
              const int64_t t = (int64_t)ty * W + tx;
              float dx, dy;
              const float w = tap_weight(&f, q, u, v, coef, &dx, &dy);
              const int k = layer_class(surface_index(&f, t, s), l);
              const float* g = grad + (t * 3 + k) * K;
              float dw = g[C];
              for (int c = 0; c < C; ++c) {
                r[c] += w * g[c];
                dw += g[c] * f.col[q * C + c];
              }
              /* w = m coef exp(-2 |d|^2): dw/dd = -4 d w; d offset/d xy =
               * -1, so d xy += 4 w dw d. */
              gx += 4.0f * w * dw * dx;
              gy += 4.0f * w * dw * dy;
            }
          /* x = ((cx / cw + 1) W) 0.5 */
          const float hx = gx * 0.5f * (float)W, hy = gy * 0.5f * (float)H;
          const float iw = 1.0f / f.cw[q];
          r[C] = hx * iw;
          r[C + 1] = hy * iw;
          r[C + 2] = -(hx * f.cx[q] + hy * f.cy[q]) * iw * iw;
        }
    float* dclip = calloc((size_t)V * 4, sizeof(float));
    if (d_colors) memset(d_colors, 0, sizeof(float) * (size_t)V * C);
    for (int64_t q = 0; q < LP; ++q) {
      if (ids[q] < 0) continue;
      const float* r = rec + q * RS;
      for (int k = 0; k < 3; ++k) {
        const int vk = corner_vertex(&f, ids[q], k);
        const float b = barys[3 * q + k];
        if (d_colors)
          for (int c = 0; c < C; ++c) d_colors[(int64_t)vk * C + c] += b * r[c];
        dclip[4 * vk] += b * r[C];
        dclip[4 * vk + 1] += b * r[C + 1];
        dclip[4 * vk + 3] += b * r[C + 2];
      }
    }
    if (d_positions) {
#pragma omp parallel for schedule(static)
      for (int i = 0; i < V; ++i)
        for (int j = 0; j < 3; ++j)
          d_positions[3 * i + j] =
              (camera[j] * dclip[4 * i] + camera[4 + j] * dclip[4 * i + 1]) +
              camera[12 + j] * dclip[4 * i + 3];
    }
    free(dclip);
    free(rec);
  }
  free(grad);
  free(f.col);
  free(f.cw);
  free(f.cy);
  free(f.cx);
  free(f.vz);
  free(f.vy);
  free(f.vx);
  free(clip);
  return loss;
}
