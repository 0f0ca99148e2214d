/*
 * diffren_amd C ABI: MI355X (gfx950) kernels for diffren's
 * rasterize -> interpolate -> splat path.
 *
 * Conventions for every entry point:
 *   - All array arguments are DEVICE pointers, except where a comment says
 *     "host".  Nothing is allocated or freed inside an entry point; scratch
 *     memory comes in through `workspace` and its size is given by the
 *     matching *_workspace_size() query.
 *   - Work is enqueued on `stream` (a hipStream_t; NULL is the default stream)
 *     and the call returns without synchronising, so calls can be captured
 *     into a hipGraph.
 *   - The return value is DR_OK (0) or an error code; dr_last_error() gives
 *     the message for the calling thread.
 *   - Layouts are the reference's: layer-major [L][H][W] buffers, channels
 *     last, row-major C order, float32 / int32.
 */
#ifndef DIFFREN_AMD_H_
#define DIFFREN_AMD_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef void* dr_stream_t; /* hipStream_t */

enum {
  DR_OK = 0,
  DR_INVALID_ARGUMENT = 1,
  DR_WORKSPACE_TOO_SMALL = 2,
  DR_LAUNCH_FAILED = 3,
  DR_UNSUPPORTED = 4
};

/* Face culling, values of diffren's FaceCullingMode
 * (diffren/common/kernels/rasterize_triangles_types.h:24,
 *  diffren/jax/constants.py:20-24). */
enum { DR_CULL_NONE = 0, DR_CULL_BACK = 1, DR_CULL_FRONT = 2 };

/* Message for the last failed call on this thread ("" if none). */
const char* dr_last_error(void);

/* Library version string. */
const char* dr_version(void);

/* ------------------------------------------------------------------------ */
/* Rasterizer                                                               */
/* ------------------------------------------------------------------------ */

/* One rasterization job (one mesh seen through one camera).  `vertices` are
 * clip-space xyzw.  Outputs are written completely: background pixels get
 * triangle id -1, z 1.0 and barycentrics 0, as the reference CPU handler
 * initialises them (rasterize_triangles_cpu.cc:55-60).  `barycentrics` may be
 * NULL.  triangle_begin / tile_begin are filled by dr_raster_jobs_prepare. */
typedef struct dr_raster_job {
  const float* vertices;   /* [num_vertices, 4] clip space, 16-byte aligned;
                              or world [num_vertices, 3] when camera is set */
  const int32_t* triangles; /* [num_triangles, 3] */
  int32_t* triangle_ids;   /* [L, height, width] */
  float* z_buffer;         /* [L, height, width] */
  float* barycentrics;     /* [L, height, width, 3] or NULL */
  const float* camera;     /* [4, 4] row-major or NULL: with a camera each
                              vertex load transforms the world position
                              exactly as dr_transform_homogeneous does, so
                              no clip-space copy is written (the camera
                              matmul of rasterize.py:90-96 fused into the
                              rasterizer) */
  int32_t num_vertices;
  int32_t num_triangles;
  int32_t width;
  int32_t height;
  int64_t triangle_begin; /* exclusive prefix of num_triangles over jobs */
  int64_t tile_begin;     /* exclusive prefix of screen tiles over jobs */
} dr_raster_job;

/* Screen-tile binning parameters (chosen by the autotuner).
 * tile_w x tile_h must be one of 8x8, 16x16, 32x8, 32x16, 32x32, 64x8.
 * bin_capacity is the number of (tile, triangle) entries the workspace holds;
 * if a frame needs more, the tiles past the capacity fall back to testing
 * every triangle of their job (correct, slow) and dr_raster_overflowed()
 * reports it.  0 selects 2 * triangles + 4096. */
typedef struct dr_raster_config {
  int32_t tile_w;
  int32_t tile_h;
  int64_t bin_capacity;
  int32_t flags;     /* DR_RASTER_* bits */
  int32_t tile_threads; /* threads per tile-pass block: 0 (= 256) or 128.
                           256 suits dense frames (hundreds of bin entries
                           per tile); 128 lets the tile pass interleave with
                           kernels running on another stream (the batched
                           renderer's job groups) on sparse ones.  Results
                           are identical. */
} dr_raster_config;

/* z and barycentrics are written only where a triangle was stored
 * (triangle_ids >= 0); background entries keep their previous contents and
 * mean z 1, barycentrics 0.  For consumers that read ids first (the fused
 * render), it saves the background writes of every layer. */
enum { DR_RASTER_SPARSE = 1 };

/* A pixel's layers form a prefix: it holds a layer-l fragment only if it
 * holds one at every layer before l.  With DR_RASTER_PREFIX the entries past
 * a pixel's first empty layer are not written at all (that first empty
 * layer is, with id -1): a consumer that stops at a pixel's first id < 0,
 * as the fused render passes do, reads only written entries, and the deeper
 * background layers' writes are saved.  Other readers must first set the
 * ids past the first -1 to -1. */
enum { DR_RASTER_PREFIX = 8 };

/* A call normally bins the triangles into screen tiles and then rasterizes
 * the tiles.  DR_RASTER_BINS_ONLY runs the binning half only; a later call
 * with DR_RASTER_TILES_ONLY and the same jobs, workspace and config runs the
 * tile half over those bins (the vertices must not change in between).  For
 * timing or overlapping the halves separately. */
enum { DR_RASTER_BINS_ONLY = 2, DR_RASTER_TILES_ONLY = 4 };

/* Host: fills triangle_begin / tile_begin of jobs_host[0..num_jobs) and
 * returns the totals.  Returns DR_OK or DR_INVALID_ARGUMENT. */
int dr_raster_jobs_prepare(dr_raster_job* jobs_host, int32_t num_jobs,
                           const dr_raster_config* config,
                           int64_t* total_triangles, int64_t* total_tiles);

/* Workspace bytes for a call with these totals. */
size_t dr_raster_workspace_size(int64_t total_triangles, int64_t total_tiles,
                                const dr_raster_config* config);

/* Rasterizes one clip-space mesh into L depth layers.
 * Replaces the reference's XLA FFI target "rasterize_triangles"
 * (diffren/jax/internal/kernels/rasterize_triangles_gpu.cc:44-77, CPU twin
 * rasterize_triangles_cpu.cc:43-79) and its CUDA implementation
 * diffren::cuda::RasterizeTrianglesImpl (rasterize_triangles_impl_cuda.cu.h:35-42).
 * Triangle ids, z and barycentrics are bit-identical to the reference CPU
 * kernel diffren::RasterizeTrianglesImpl (rasterize_triangles_impl.cc:27-118).
 * Triangles with a vertex index outside [0, num_vertices) are skipped.
 * num_layers must be in [1, 64], with num_layers x tile_w x tile_h x 8
 * bytes of depth keys at most 64 KiB (8x8 tiles take up to 64 layers). */
int dr_rasterize_triangles(const float* vertices, int32_t num_vertices,
                           const int32_t* triangles, int32_t num_triangles,
                           int32_t width, int32_t height, int32_t num_layers,
                           int32_t face_culling_mode, int32_t* triangle_ids,
                           float* z_buffer, float* barycentrics,
                           const dr_raster_config* config, void* workspace,
                           size_t workspace_bytes, dr_stream_t stream);

/* Bytes of the static scratch that dr_rasterize_triangles_noalloc uses. */
#define DR_STATIC_SCRATCH_BYTES (64u << 20)

/* dr_rasterize_triangles without a workspace argument: the scratch is a
 * file-scope device array of DR_STATIC_SCRATCH_BYTES (64 MiB) per GPU, held
 * until the process ends, so the call allocates nothing.  No counterpart in
 * the reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, serving loops, handlers that forbid allocation).
 * A call whose workspace (dr_raster_workspace_size) exceeds the scratch
 * returns DR_WORKSPACE_TOO_SMALL.  The stream must belong to the current
 * device.  Every call on a GPU shares that GPU's scratch: calls that overlap
 * on one GPU (different streams, or host threads) race and must be
 * serialised by the caller.  Results are those of dr_rasterize_triangles. */
int dr_rasterize_triangles_noalloc(const float* vertices, int32_t num_vertices,
                                   const int32_t* triangles,
                                   int32_t num_triangles, int32_t width,
                                   int32_t height, int32_t num_layers,
                                   int32_t face_culling_mode,
                                   int32_t* triangle_ids, float* z_buffer,
                                   float* barycentrics,
                                   const dr_raster_config* config,
                                   dr_stream_t stream);

/* Rasterizes many (mesh, camera, resolution) jobs in one set of launches.
 * jobs_device: device copy of a table prepared by dr_raster_jobs_prepare.
 * No reference counterpart: the reference batches with
 * vmap_method='sequential_unrolled', one launch per image
 * (rasterize_triangles_xla.py:68-75). */
int dr_rasterize_triangles_batched(const dr_raster_job* jobs_device,
                                   int32_t num_jobs, int64_t total_triangles,
                                   int64_t total_tiles, int32_t num_layers,
                                   int32_t face_culling_mode,
                                   const dr_raster_config* config,
                                   void* workspace, size_t workspace_bytes,
                                   dr_stream_t stream);

/* Host: reads back (synchronously, on `stream`) the binning statistics of the
 * last call that used `workspace`: entries needed and whether the capacity
 * overflowed.  For the autotuner and tests; not for hot loops. */
int dr_raster_bin_stats(const void* workspace, int64_t* entries_needed,
                        int32_t* overflowed, dr_stream_t stream);


/* ------------------------------------------------------------------------ */
/* Vertex transform                                                         */
/* ------------------------------------------------------------------------ */

/* out[v] = M @ [x, y, z, (w or 1)] for a row-major [4,4] device `matrix`,
 * summed in the fixed order ((m0 x + m1 y) + m2 z) + m3 w.
 * Replaces transforms.transform_homogeneous (diffren/jax/utils/transforms.py:
 * 40-78) as used by rasterize.py:90-96 and splat.py:209-212.
 * vertices [num_vertices, dims] (dims 3 or 4), out [num_vertices, 4]. */
int dr_transform_homogeneous(const float* matrix, const float* vertices,
                             int32_t dims, int64_t num_vertices, float* out,
                             dr_stream_t stream);

/* One mesh through `views` matrices [views, 4, 4]: out [views, V, 4], the
 * same arithmetic as dr_transform_homogeneous, one launch. */
int dr_transform_homogeneous_multi(const float* matrices, int32_t views,
                                   const float* vertices, int32_t dims,
                                   int64_t num_vertices, float* out,
                                   dr_stream_t stream);

/* d_vertices[v, j] = sum_i M[i, j] d_clip[v, i], j < dims (overwrites). */
int dr_transform_homogeneous_bwd(const float* matrix, const float* d_clip,
                                 int32_t dims, int64_t num_vertices,
                                 float* d_vertices, dr_stream_t stream);

/* ------------------------------------------------------------------------ */
/* Differentiable barycentrics                                              */
/* ------------------------------------------------------------------------ */

/* barycentrics [L, H, W, 3] recomputed from the clip-space vertices
 * [num_vertices, 4] (16-byte aligned) for the triangle id of every
 * pixel-layer (triangle_ids [L, H, W] from the rasterizer): normalised
 * 2D-homogeneous edge functions at pixel centres px = (2x+1)/W - 1,
 * py = (2y+1)/H - 1, nan_to_num of r / sum(r), 0 on background.  Replaces
 * differentiable_barycentrics, normalized_pixel_coordinates,
 * compute_triangle_matrix and compute_barycentric_coordinates
 * (diffren/jax/internal/rasterize.py:127-256), called from
 * rasterize_triangles(compute_diff_barys=True) (:120-121). */
int dr_diff_barycentrics_fwd(const float* clip_vertices, int64_t num_vertices,
                             const int32_t* triangles, int32_t num_triangles,
                             const int32_t* triangle_ids, int32_t width,
                             int32_t height, int32_t num_layers,
                             float* barycentrics, dr_stream_t stream);

/* d_clip_vertices [num_vertices, 4] += the VJP of dr_diff_barycentrics_fwd
 * for d_barycentrics [L, H, W, 3] (x, y, w columns; z gets none).  Float
 * atomics: the summation order is not fixed; the caller clears the output.
 * No reference counterpart (the reference uses XLA autodiff of
 * rasterize.py:206-256). */
int dr_diff_barycentrics_bwd(const float* clip_vertices, int64_t num_vertices,
                             const int32_t* triangles, int32_t num_triangles,
                             const int32_t* triangle_ids, int32_t width,
                             int32_t height, int32_t num_layers,
                             const float* d_barycentrics,
                             float* d_clip_vertices, dr_stream_t stream);

/* ------------------------------------------------------------------------ */
/* Interpolation                                                            */
/* ------------------------------------------------------------------------ */

/* out[p, c] = (b0 A[v0,c] + b1 A[v1,c]) + b2 A[v2,c] where triangle_ids[p]
 * is a triangle (v0,v1,v2) = triangles[id], else background[c] (0 if
 * background is NULL).  num_pixels = L*H*W pixel-layers; barycentrics
 * [num_pixels, 3]; attributes [num_vertices, channels]; out [num_pixels,
 * channels].  Replaces interpolate._interpolate_vertex_attribute_layer and
 * interpolate_vertex_attribute (diffren/jax/internal/interpolate.py:24-63,
 * 118-145), including the vertex_ids gather of rasterize.py:111-113. */
int dr_interpolate_fwd(const int32_t* triangle_ids, const float* barycentrics,
                       const int32_t* triangles, int32_t num_triangles,
                       const float* attributes, int32_t num_vertices,
                       int32_t channels, const float* background,
                       int64_t num_pixels, float* out, dr_stream_t stream);

/* d_attributes[v_k, c] += b_k d_out[p, c] over foreground p (accumulates:
 * the caller clears d_attributes).  Float atomics: the summation order is
 * not fixed.  No reference counterpart (the reference uses XLA autodiff of
 * interpolate.py:49-50). */
int dr_interpolate_bwd(const int32_t* triangle_ids, const float* barycentrics,
                       const int32_t* triangles, int32_t num_triangles,
                       const float* d_out, int32_t num_vertices,
                       int32_t channels, int64_t num_pixels,
                       float* d_attributes, dr_stream_t stream);

/* d_barycentrics[p, k] = sum_c d_out[p, c] attributes[v_k, c] for
 * foreground p, 0 on background (written, not accumulated): the
 * barycentric half of the interpolation VJP, used when the barycentrics
 * come from dr_diff_barycentrics_fwd.  No reference counterpart (XLA
 * autodiff of interpolate.py:49-50 through the differentiable barycentrics
 * of rasterize.py:120-121). */
int dr_interpolate_bwd_barycentrics(const int32_t* triangle_ids,
                                    const int32_t* triangles,
                                    int32_t num_triangles,
                                    const float* attributes,
                                    int32_t num_vertices, int32_t channels,
                                    const float* d_out, int64_t num_pixels,
                                    float* d_barycentrics,
                                    dr_stream_t stream);

/* ------------------------------------------------------------------------ */
/* Splatting                                                                */
/* ------------------------------------------------------------------------ */

/* The splat coefficient 1.05 / sum_{u,v in -1..1} exp(-2 (u^2+v^2))
 * (splat.py:63-78), as float. */
float dr_splat_coefficient(void);

/* 3x3 Gaussian splat of every pixel-layer sample at its pixel centre,
 * accumulated into foreground / surface / background layers, normalised and
 * composited with `over`.  Replaces splat.splat_at_pixel_centers
 * (diffren/jax/internal/splat.py:26-175).
 *   xyz [L,H,W,3] viewport positions, color [L,H,W,C], mask [L,H,W]
 *   out [H,W,C]; optional (NULL to skip): accum [3,H,W,C], norm [3,H,W,C],
 *   accum_weights [3,H,W].  C one of 1-16, 20, 24, 28, 32 (callers pad
 *   other counts with zero channels before the alpha channel, or split
 *   them), any L >= 1 (up to 8 layers a pixel's layer depths stay in
 *   registers; deeper frames read them from memory where needed). */
int dr_splat_fwd(const float* xyz, const float* color, const float* mask,
                 int32_t num_layers, int32_t height, int32_t width,
                 int32_t channels, float* out, float* accum, float* norm,
                 float* accum_weights, dr_stream_t stream);

/* Workspace bytes for dr_splat_bwd: H*W*3*(C+1) floats. */
size_t dr_splat_bwd_workspace_size(int32_t height, int32_t width,
                                   int32_t channels);

/* Vector-Jacobian product of dr_splat_fwd's `out`: writes d_color [L,H,W,C]
 * and d_xyz [L,H,W,3] (z component 0: z only selects layers).  No reference
 * counterpart (the reference uses XLA autodiff of splat.py:26-175). */
int dr_splat_bwd(const float* xyz, const float* color, const float* mask,
                 int32_t num_layers, int32_t height, int32_t width,
                 int32_t channels, const float* d_out, float* d_color,
                 float* d_xyz, void* workspace, size_t workspace_bytes,
                 dr_stream_t stream);


/* ------------------------------------------------------------------------ */
/* Fused, batched rasterize-then-splat render (forward + backward)          */
/* ------------------------------------------------------------------------ */

/* One render job: a rasterized view of a mesh and the colours to splat --
 * either per-vertex `colors` under an identity shader (render.py:123-132
 * with shading_function returning the interpolated colours), or the shaded
 * pixels of any shading function (`shaded`, splat.py:178-229's
 * shaded_pixels).  All pointers are device pointers; C = channels, the last
 * channel is alpha.  When C is a multiple of 4 the [.., C] frames (image,
 * target, d_image, accum, norm, shaded, d_shaded) must be 16-byte aligned:
 * their rows move as 16-byte loads and stores. */
typedef struct dr_splat_job {
  const float* clip_vertices;   /* [V, 4] clip space, 16-byte aligned; or
                                   NULL to transform `positions` by `camera`
                                   on load (as dr_transform_homogeneous) */
  const float* positions;       /* [V, 3] world positions when
                                   clip_vertices is NULL */
  const int32_t* triangles;     /* [T, 3] */
  const float* colors;          /* [V, C] */
  const float* camera;          /* [4, 4] row-major, or NULL (see d_vertices) */
  const int32_t* triangle_ids;  /* [L, H, W] from the rasterizer */
  const float* z_buffer;        /* [L, H, W] from the rasterizer */
  const float* barycentrics;    /* [L, H, W, 3] from the rasterizer */
  float* image;                 /* [H, W, C] output, or NULL */
  const float* target;          /* [H, W, C] L2 target (grad_mode 1) */
  const float* d_image;         /* [H, W, C] upstream gradient (grad_mode 2) */
  float* grad_scratch;          /* [H, W, dr_render_scratch_floats(C)]
                                   per-job workspace, 16-byte aligned */
  float* d_records;             /* [V, R] accumulated gradient records, or
                                   NULL; R = dr_render_record_size(C).
                                   cols 0-2: d vertex = M^T d clip when
                                   camera is set, else cols 0, 1, 3: d clip
                                   x, y, w; cols 4 .. 4+C-1: d colours.
                                   Jobs that share a mesh share it. */
  float* accum;                 /* [3, H, W, C] accumulated colours of the
                                   fg / surface / bg layers, or NULL */
  float* norm;                  /* [3, H, W, C] normalised, or NULL */
  float* d_camera;              /* [4, 4] accumulated gradient of the loss
                                   with respect to `camera` through the
                                   splat's transform (splat.py:209-212;
                                   rows x, y, w; the rasterization is held
                                   fixed, rasterize_triangles_xla.py:51-52),
                                   or NULL; needs camera and positions.
                                   Caller clears. */
  float* accum_weights;         /* [3, H, W] accumulated weights, or NULL
                                   (accum and norm are the reference's
                                   return_accum_buffers outputs,
                                   splat.py:166-175) */
  const float* shaded;          /* [L, H, W, C] shaded pixels, or NULL for
                                   per-vertex `colors`.  When set, `colors`
                                   is not read, d_records carry only the
                                   position gradient (R = 4) and the colour
                                   gradient goes to d_shaded. */
  float* d_shaded;              /* [L, H, W, C] gradient of the loss with
                                   respect to `shaded`, written completely
                                   by the backward (0 on background), or
                                   NULL */
  int32_t num_vertices;
  int32_t num_triangles;
  int32_t width;
  int32_t height;
  int64_t tile_begin;           /* filled by dr_splat_jobs_prepare */
  int32_t fixed_shift;          /* 0: d_records and d_camera accumulate
                                   float atomics (summation order not
                                   fixed).  1-62: deterministic mode: both
                                   hold int64 (8-byte aligned) and each
                                   contribution adds round(v * 2^shift)
                                   exactly, so the result does not depend
                                   on the order; dr_fixed_to_float
                                   converts.  Range |sum| < 2^(63-shift),
                                   resolution 2^-shift. */
  int32_t reserved;             /* 0 */
} dr_splat_job;

/* Host: fills tile_begin of jobs[0..n) and returns the total tile count. */
int dr_splat_jobs_prepare(dr_splat_job* jobs_host, int32_t num_jobs,
                          int32_t channels, int64_t* total_tiles);

/* Floats per pixel of a job's grad_scratch (0 if the channel count is not
 * one of 1-16, 20, 24, 28). */
int32_t dr_render_scratch_floats(int32_t channels);

/* Floats per vertex gradient record for `channels` colour channels (8, 16
 * or 32; 0 if unsupported).  Jobs with `shaded` pixels use 4-float records
 * (the position gradient only). */
int32_t dr_render_record_size(int32_t channels);

enum { DR_RENDER_ACCUM_BUFFERS = 4, DR_RENDER_SHADED = 8 };

/* Workspace bytes of dr_render_splat_forward (per-tile loss partials). */
size_t dr_render_splat_workspace_size(int64_t total_tiles);

/* Target pass over all jobs in one launch: interpolate colours and viewport
 * positions (interpolate.py:24-63, splat.py:209-221), splat
 * (splat.py:26-175) and composite (composite.py:20-37) into `image`.
 * grad_mode 0: image only.  1: also the L2 loss sum((image - target)^2) *
 * loss_scale added to *loss (device float, caller clears; summed per tile,
 * then in a fixed order, so it is deterministic) and its gradient
 * 2 (image - target) loss_scale pushed into grad_scratch.  2: the given
 * d_image pushed into grad_scratch.  grad_mode | DR_RENDER_ACCUM_BUFFERS
 * also writes the jobs' accum / norm / accum_weights outputs (ignored
 * without it).  channels 1-16, 20, 24 or 28.  The workspace
 * is needed for grad_mode 1 only.  A job whose target (mode 1) or d_image
 * (mode 2) or grad_scratch is NULL gets its image only (no loss, no
 * gradient).
 * No reference counterpart as one call: it fuses render.py:123-132 for an
 * identity shader, batched where the reference runs one launch per image. */
int dr_render_splat_forward(const dr_splat_job* jobs_device, int32_t num_jobs,
                            int64_t total_tiles, int32_t num_layers,
                            int32_t channels, int32_t grad_mode,
                            float loss_scale, float* loss, void* workspace,
                            size_t workspace_bytes, dr_stream_t stream);

/* out[i] = fixed[i] * 2^-fixed_shift: the deterministic mode's integer
 * gradient sums as float32.  No reference counterpart. */
int dr_fixed_to_float(const int64_t* fixed, int64_t n, int32_t fixed_shift,
                      float* out, dr_stream_t stream);

/* Source pass: from grad_scratch (written by a forward with grad_mode 1 or
 * 2), accumulate the d_records (and d_camera) of all jobs (float atomics, or
 * fixed point) and write the d_shaded of jobs with shaded pixels; jobs with
 * no grad_scratch, or none of d_records, d_camera and d_shaded, are
 * skipped.  The hand-written VJP of splat.py:178-229 + interpolate.py:24-63
 * + transforms.py:40-78; the reference uses XLA autodiff.  flags 0
 * differentiates the jobs with per-vertex colours, DR_RENDER_SHADED the jobs
 * with shaded pixels (the others are skipped). */
int dr_render_splat_backward(const dr_splat_job* jobs_device,
                             int32_t num_jobs, int64_t total_tiles,
                             int32_t num_layers, int32_t channels,
                             int32_t flags, dr_stream_t stream);


/* ------------------------------------------------------------------------ */
/* Texture sampling                                                         */
/* ------------------------------------------------------------------------ */

enum { DR_PAD_CONSTANT = 0, DR_PAD_EDGE = 1 };

/* out [num_samples, C] = bilinear samples of image [H, W, C] (HWC float32)
 * at locations [num_samples, 2] (x, y in image units, 8-byte aligned; texel
 * (i, j) is centred at (j + 0.5, i + 0.5)), outside the image padded with 0
 * (DR_PAD_CONSTANT) or the edge texel (DR_PAD_EDGE).  Replaces
 * image.bilinear_resample (diffren/jax/utils/image.py:21-118) for
 * image_order = locations_order = 'HWC'; shaders.texture_map
 * (diffren/jax/utils/shaders.py:47-72) is DR_PAD_EDGE at uv * (W, H). */
int dr_bilinear_resample_fwd(const float* image, int32_t height,
                             int32_t width, int32_t channels,
                             const float* locations, int64_t num_samples,
                             int32_t pad_mode, float* out,
                             dr_stream_t stream);

/* VJP of dr_bilinear_resample_fwd for d_out [num_samples, C]:
 * d_image [H, W, C] += (float atomics; NULL to skip) and d_locations
 * [num_samples, 2] = (written; NULL to skip).  The weights' clip to [0, 1]
 * is differentiated as JAX differentiates minimum(maximum(t, 0), 1): 1/2 at
 * a bound.  No reference counterpart (XLA autodiff of image.py:52-92). */
int dr_bilinear_resample_bwd(const float* image, int32_t height,
                             int32_t width, int32_t channels,
                             const float* locations, int64_t num_samples,
                             int32_t pad_mode, const float* d_out,
                             float* d_image, float* d_locations,
                             dr_stream_t stream);

/* ------------------------------------------------------------------------ */
/* Deferred shading: G-buffer, many-light Phong, learned MLP shader         */
/* ------------------------------------------------------------------------ */

/* G-buffer [num_pixels, 12] = world position, unit normal, unit view
 * direction (eye - position), albedo; zeros on background.  Attributes are
 * interpolated as dr_interpolate_fwd does; normals and view directions are
 * normalised as transforms.l2_normalize (transforms.py:26-37) does.
 * positions/normals/albedo [num_vertices, 3]; eye [3] (device). */
int dr_gbuffer(const int32_t* triangle_ids, const float* barycentrics,
               const int32_t* triangles, int32_t num_triangles,
               const float* positions, const float* normals,
               const float* albedo, int32_t num_vertices, const float* eye,
               int64_t num_pixels, float* gbuffer, dr_stream_t stream);

typedef struct dr_phong_params {
  float ka, kd, ks, ns;
  float attenuation_constant, attenuation_linear, attenuation_quadratic;
} dr_phong_params;

/* out [num_pixels, 3] = albedo * (ka + sum_i I_i decay_i (kd nol_i +
 * ks noh_i^ns)), lights [num_lights, 4] = (x, y, z, intensity).  The
 * many-light form of shaders.point_light (shaders.py:100-173) with the
 * ambient term counted once. */
int dr_phong_shade(const float* gbuffer, int64_t num_pixels,
                   const float* lights, int32_t num_lights,
                   const dr_phong_params* params, float* out,
                   dr_stream_t stream);

/* MLP 12 -> 64 -> 64 -> 64 -> 64 -> 3 (ReLU), bfloat16 MFMA operands, float32
 * accumulation.  No reference counterpart (a learned replacement for the
 * closed-form shader above).  Parameters are a flat float32 vector of
 * dr_mlp_num_params() entries: W1[12,64] b1 W2[64,64] b2 W3 b3 W4 b4
 * W5[64,3] b5, W[k_in][unit_out]. */
int32_t dr_mlp_num_params(void);

/* Bytes of the packed bfloat16 operand image written by dr_mlp_pack. */
size_t dr_mlp_packed_bytes(void);

/* Packs float32 parameters into the MFMA operand image (forward and
 * backward orders); call after every parameter update. */
int dr_mlp_pack(const float* params, void* packed, dr_stream_t stream);

/* Input normalisation and output options of the MLP entry points:
 * input_norm (device, may be NULL = identity) holds [shift 12 | scale 12]
 * and the network sees (x - shift) * scale; with DR_MLP_ALBEDO_MODULATE the
 * image is albedo * network output, albedo being x[:, 9:12] (raw).  With
 * DR_MLP_GRAD_OVERWRITE dr_mlp_train_step writes grad and loss instead of
 * adding to them. */
enum { DR_MLP_ALBEDO_MODULATE = 1, DR_MLP_GRAD_OVERWRITE = 2 };

/* y [num_pixels, 3] = MLP(x [num_pixels, 12]).  num_blocks <= 0: default. */
int dr_mlp_forward(const void* packed, const float* params, const float* x,
                   int64_t num_pixels, float* y, const float* input_norm,
                   int32_t flags, int32_t num_blocks, dr_stream_t stream);

/* Workspace bytes of dr_mlp_train_step for num_blocks (<= 0: default). */
size_t dr_mlp_train_workspace_size(int32_t num_blocks);

/* One training pass: loss += loss_scale * sum_p mask[p] |y_p - target_p|^2
 * and grad += d loss / d params (float32, accumulated: the caller clears
 * grad and loss, or passes DR_MLP_GRAD_OVERWRITE; per-block partials are
 * summed in a fixed order, so the result is deterministic).  mask
 * [num_pixels] may be NULL (all ones); x must be 16-byte aligned and
 * num_pixels at most 2^35. */
int dr_mlp_train_step(const void* packed, const float* params, const float* x,
                      const float* target, const float* mask,
                      int64_t num_pixels, float loss_scale,
                      const float* input_norm, int32_t flags, float* grad,
                      float* loss, int32_t num_blocks, void* workspace,
                      size_t workspace_bytes, dr_stream_t stream);

/* Adam update of params with moments m, v (all float32, num_params);
 * step counts from 1. */
int dr_mlp_adam(float* params, const float* grad, float* m, float* v,
                int32_t step, float lr, float beta1, float beta2, float eps,
                dr_stream_t stream);

#ifdef __cplusplus
}
#endif

#endif /* DIFFREN_AMD_H_ */
