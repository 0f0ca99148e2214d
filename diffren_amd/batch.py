"""Batched, fused rasterize-then-splat rendering with a hand-written backward.

`BatchedSplatRenderer` renders many (mesh, camera, resolution) jobs -- of
uneven triangle counts and image sizes -- through one set of launches per
stage:
  transform (one launch per mesh, all its views) -> batched tile rasterizer
  (one launch per stage for all jobs of a group) -> fused target pass
  (interpolate + splat + over + L2 gradient) -> fused source pass (splat VJP
  + interpolate VJP + transform VJP, scattering into d colours and
  d vertices).
Jobs can be split into `streams` groups that run concurrently on their own
streams (the tail of one group's kernels overlaps the other's work).  All
buffers are allocated once at construction, so a step allocates nothing and
is captured in a HIP graph (`capture()` / `replay()`).

This is the batched counterpart of calling render.render_triangles(...,
shading_function=lambda a: a["colors"], compositing_mode=SPLAT_OVER) once per
view (render.py:28-140 under jax.vmap, which issues one rasterizer launch per
image: rasterize_triangles_xla.py:68-75) followed by jax.grad of an L2 loss.
With `shading_function` set, any shader runs between the rasterizer and the
fused splat: each job's mesh attributes are interpolated (interp kernels),
shaded by the function (torch, differentiable), splatted from the shaded
pixels by the fused kernels, and the shaded pixels' gradient is pushed back
through the shader and the interpolation by autograd.
"""
import ctypes
import json

import torch

from diffren_amd import _native
from diffren_amd import framebuffer
from diffren_amd import interpolate
from diffren_amd import rasterize


class _Group:
  """Jobs of one stream: their raster and splat job tables and workspaces."""

  def __init__(self, indices):
    self.indices = indices


class BatchedSplatRenderer:
  """Preplanned fused forward+backward over a fixed batch of render jobs.

  meshes:  list of dicts with device tensors "vertices" [V,3] f32,
           "triangles" [T,3] i32, "colors" [V,C] f32; jobs reference a mesh
           by index, so several views of one mesh share its gradient buffers.
  jobs:    list of (mesh_index, camera [4,4] device f32, width, height).
           Cameras are copied at construction.
  tile:    raster tile (w, h), None for the default or "auto" for the
           autotuner (diffren_amd.autotune).
  streams: number of job groups run concurrently (each on its own stream),
           or "auto": the autotuner times the captured step with 1, 2 and 4
           groups on this batch and keeps the fastest (streams_timings),
           then times that step with the tile pass in 256- and 128-thread
           blocks (tile_threads_timings).
  tile_threads: threads per tile-pass block (dr_raster_config.tile_threads):
           128 or 256, "auto" to time both on the captured step, or None:
           "auto" with streams="auto", else 256.
  fused_transform: when True no clip-space buffers are written: the
           rasterizer and the splat passes read the world-space vertices
           and transform them by the job's camera as they load them (same
           float expression as the transform launch, so results are
           identical).

  return_accum_buffers: also keep, per job, the splat's accumulated and
           normalised fg / surface / bg layers accum[i], norm[i]
           [3, H, W, C] (render_triangles(return_accum_buffers=True),
           render.py:130-132) and their weights accum_weights[i] [3, H, W].

  camera_grads: also accumulate d loss / d camera of every job into
           d_cameras[i] [4, 4] (through the splat's transform, splat.py:
           209-212; rasterization held fixed as rasterize_triangles_xla.py:
           51-52 holds it).

  deterministic: accumulate the vertex (and camera) gradients as exact
           64-bit fixed-point sums with resolution 2^-fixed_shift (range
           |g| < 2^(63 - fixed_shift)), converted to float32 after the
           backward: the gradients then do not depend on the order in which
           the GPU's atomics arrive, so repeated steps give identical bits.
           The loss is always summed in a fixed order.

  shading_function: None for the identity shader on the meshes' "colors".
           Otherwise a function of a dict of interpolated attributes
           ([L, H, W, Ca] each, the meshes' "attributes" dicts of [V, Ca]
           device tensors, background 0) returning [L, H, W, channels]
           shaded pixels, alpha last -- render_triangles' shading_function
           (render.py:123-129).  The step then interpolates, shades, splats
           the shaded pixels with the fused kernels, and back-propagates the
           shaded pixels' gradient through the shader and the interpolation
           (torch.autograd.grad): d_attributes[m][name] holds the gradient
           with respect to mesh m's attribute `name` (for the attributes
           that require grad), d_vertices the gradient through the splat
           positions, d_colors is None.  `channels` is then required.

  After step(): images[i], loss (sum over jobs of loss_scale * |img - tgt|^2),
  d_vertices[m] [V,3] and d_colors[m] [V,C] (views of records[m]).  ids[i]
  holds triangle ids; z[i] / barys[i] are defined under triangles only.
  """

  def __init__(self, meshes, jobs, num_layers, tile=None, bin_capacity=0,
               loss_scale=None, autotune_cache=None, streams=1,
               fused_transform=False, return_accum_buffers=False,
               camera_grads=False, deterministic=False, fixed_shift=40,
               shading_function=None, channels=None, tile_threads=None):
    lib = _native.lib()
    self.meshes = meshes
    self.jobs = jobs
    self.L = int(num_layers)
    dev = meshes[0]["vertices"].device
    self.device = dev
    self.shading_function = shading_function
    if shading_function is not None:
      if channels is None:
        raise ValueError("shading_function needs `channels`, the channel "
                         "count of its output")
      C = int(channels)
    else:
      C = meshes[0]["colors"].shape[1]
    for m in meshes:
      v, t = m["vertices"], m["triangles"]
      _native.require_device(v, t)
      if (v.dtype != torch.float32 or v.dim() != 2 or v.shape[1] != 3 or
          not v.is_contiguous()):
        raise ValueError("mesh vertices must be contiguous float32 [V, 3]")
      if (t.dtype != torch.int32 or t.dim() != 2 or t.shape[1] != 3 or
          not t.is_contiguous()):
        raise ValueError("mesh triangles must be contiguous int32 [T, 3]")
      if shading_function is not None:
        attrs = m.get("attributes")
        if not isinstance(attrs, dict) or not attrs:
          raise ValueError("with a shading_function every mesh needs an "
                           "\"attributes\" dict of [V, Ca] tensors")
        for a in attrs.values():
          _native.require_device(a)
          if a.dim() != 2 or a.shape[0] != v.shape[0]:
            raise ValueError("mesh attributes must be [V, Ca]")
        continue
      c = m["colors"]
      _native.require_device(c)
      if (c.dtype != torch.float32 or c.dim() != 2 or
          c.shape[0] != v.shape[0] or not c.is_contiguous()):
        raise ValueError("mesh colors must be contiguous float32 [V, C]")
      if c.shape[1] != C:
        raise ValueError("all meshes need the same channel count")
    for mi, cam, w, h in jobs:
      if not 0 <= mi < len(meshes):
        raise ValueError(f"job references mesh {mi} of {len(meshes)}")
      if tuple(cam.shape) != (4, 4) or w <= 0 or h <= 0:
        raise ValueError("jobs are (mesh index, camera [4, 4], width, "
                         "height) with positive sizes")
    self.C = C
    self.fused_transform = bool(fused_transform)
    # Per-mesh gradient records [V, R]: d position in cols 0..2, d colour in
    # cols 4..4+C-1 (one contiguous record per vertex for grouped atomics).
    R = lib.dr_render_record_size(C)
    if R > 0 and shading_function is not None:
      R = 4  # shaded pixels: the records carry the position gradient only
    if R <= 0:
      raise ValueError(
          f"the fused path takes 1-16, 20, 24 or 28 colour channels, not {C}:"
          " pad the colours with zero channels before the alpha channel, or "
          "render through render.render_triangles, which takes any count")
    self.records = [torch.zeros((m["vertices"].shape[0], R),
                                dtype=torch.float32, device=dev)
                    for m in meshes]
    self.d_vertices = [r[:, 0:3] for r in self.records]
    self.d_colors = (None if shading_function is not None else
                     [r[:, 4:4 + C] for r in self.records])
    self.fixed_shift = int(fixed_shift) if deterministic else 0
    if deterministic and not 1 <= self.fixed_shift <= 62:
      raise ValueError("fixed_shift must be in [1, 62]")
    self.records_fixed = ([torch.zeros(r.shape, dtype=torch.int64, device=dev)
                           for r in self.records] if deterministic else [])
    self.loss = torch.zeros(1, dtype=torch.float32, device=dev)
    self._ids, self.z, self.barys = [], [], []
    self.images, self.targets, self.scratch = [], [], []
    # Clip-space vertices: one [views, V, 4] block per mesh, its views laid
    # out group by group (_layout_views), so each job group transforms its
    # own views with one launch per mesh, on its own stream.
    self.mesh_views = {}
    for i, (mi, _, _, _) in enumerate(jobs):
      self.mesh_views.setdefault(mi, []).append(i)
    self.clip_blocks, self.camera_blocks = {}, {}
    for mi, views in ([] if self.fused_transform else
                      self.mesh_views.items()):
      V = meshes[mi]["vertices"].shape[0]
      self.clip_blocks[mi] = torch.empty((len(views), V, 4),
                                         dtype=torch.float32, device=dev)
      self.camera_blocks[mi] = torch.stack(
          [jobs[i][1].to(torch.float32).reshape(4, 4) for i in views])
    self.clip = [None if self.fused_transform else
                 self.clip_blocks[mi][self.mesh_views[mi].index(i)]
                 for i, (mi, _, _, _) in enumerate(jobs)]
    sf = lib.dr_render_scratch_floats(C)
    self.d_cameras = (torch.zeros((len(jobs), 4, 4), dtype=torch.float32,
                                  device=dev) if camera_grads else None)
    self.d_cameras_fixed = (torch.zeros((len(jobs), 4, 4), dtype=torch.int64,
                                        device=dev)
                            if camera_grads and deterministic else None)
    self.accum, self.norm, self.accum_weights = [], [], []
    if return_accum_buffers:
      for _, _, w, h in jobs:
        self.accum.append(torch.zeros((3, h, w, C), dtype=torch.float32,
                                      device=dev))
        self.norm.append(torch.zeros((3, h, w, C), dtype=torch.float32,
                                     device=dev))
        self.accum_weights.append(torch.zeros((3, h, w), dtype=torch.float32,
                                              device=dev))
    npix = 0
    for mi, cam, w, h in jobs:
      self._ids.append(torch.empty((self.L, h, w), dtype=torch.int32,
                                   device=dev))
      self.z.append(torch.empty((self.L, h, w), dtype=torch.float32,
                                device=dev))
      self.barys.append(torch.empty((self.L, h, w, 3), dtype=torch.float32,
                                    device=dev))
      self.images.append(torch.empty((h, w, C), dtype=torch.float32,
                                      device=dev))
      self.targets.append(torch.zeros((h, w, C), dtype=torch.float32,
                                      device=dev))
      self.scratch.append(torch.empty((h, w, sf), dtype=torch.float32,
                                      device=dev))
      npix += w * h * C
    self.shaded, self.d_shaded = [], []
    self.d_attributes = None
    if shading_function is not None:
      self.d_attributes = [
          {k: torch.zeros(a.shape, dtype=torch.float32, device=dev)
           for k, a in m["attributes"].items() if a.requires_grad}
          for m in meshes]
      for _, _, w, h in jobs:
        self.shaded.append(torch.zeros((self.L, h, w, C), dtype=torch.float32,
                                       device=dev))
        self.d_shaded.append(torch.zeros((self.L, h, w, C),
                                         dtype=torch.float32, device=dev))
    self.loss_scale = 1.0 / npix if loss_scale is None else float(loss_scale)
    self.cameras = [j[1].to(torch.float32).contiguous() for j in jobs]
    self.autotune_cache = autotune_cache
    self.autotune_timings = None
    self.streams_timings = None
    self.tile_threads_timings = None
    auto_groups = streams == "auto"
    self._set_groups(1 if auto_groups else streams)
    # The fused passes read ids first, so background z / barycentrics are
    # never written (DR_RASTER_SPARSE), and stop at a pixel's first empty
    # layer, so the layers behind it are not written either
    # (DR_RASTER_PREFIX; the ids property completes them for other readers).
    # A shading function interpolates every layer: it gets complete ids.
    self._raster_flags = _native.RASTER_SPARSE | (
        _native.RASTER_PREFIX if shading_function is None else 0)
    if tile == "auto":
      tile, bin_capacity = self._autotune()
    self.raster_config = rasterize._config(tile, bin_capacity,
                                           self._raster_flags)
    self.graph = None
    self._graph_forward_only = False
    if auto_groups:
      self._set_groups(self._autotune_groups())
    for g in self.groups:
      self._plan_raster(g)
      self._plan_splat(g)
    if tile_threads is None:
      tile_threads = "auto" if auto_groups else 256
    if tile_threads == "auto":
      self._autotune_tile_threads()
    else:
      if int(tile_threads) not in (128, 256):
        raise ValueError("tile_threads must be 128, 256 or 'auto'")
      cfg = self.raster_config
      self.raster_config = _native.RasterConfig(
          cfg.tile_w, cfg.tile_h, cfg.bin_capacity, cfg.flags,
          int(tile_threads))
    # Bin-overflow watch.  Each group's rasterizer status words (entries
    # needed, overflow flag) are copied into pinned host memory at the end
    # of its raster stage, inside a captured graph too.  The next step() or
    # replay() reads them once the previous step has completed; when a
    # group's bins overflowed (its tiles then fell back to scanning every
    # triangle record of their job: correct, slow), the bin capacity grows to
    # twice the entries needed, the workspaces are re-planned and a captured
    # graph is captured again.  `regrows` counts these.
    self._status_event = None
    self.regrows = 0

  # -- planning --------------------------------------------------------------

  def _set_groups(self, n):
    """Splits the jobs into n groups (each with its own stream and loss
    slot); the groups still need _plan_raster / _plan_splat."""
    dev = self.device
    self.groups = self._split(max(1, int(n)))
    # One loss slot per group, summed in a fixed order after the groups join.
    self.loss_parts = torch.zeros(len(self.groups), dtype=torch.float32,
                                  device=dev)
    for k, g in enumerate(self.groups):
      g.loss_slot = self.loss_parts[k:k + 1]
    self.streams = [torch.cuda.Stream(device=dev) for _ in self.groups] \
        if len(self.groups) > 1 else []
    # Each group's raster status words (entries needed, overflow flag),
    # copied to pinned host memory inside every step (see _check_overflow).
    self._status_host = torch.zeros((len(self.groups), 16), dtype=torch.uint8,
                                    pin_memory=True)
    self._layout_views()

  def _layout_views(self):
    """Orders each mesh's views in its clip block group by group and gives
    every group its (mesh, first view, view count) transform ranges."""
    for g in self.groups:
      g.transforms = []
    if self.fused_transform:
      return
    for mi in self.mesh_views:
      order = [i for g in self.groups for i in g.indices
               if self.jobs[i][0] == mi]
      self.mesh_views[mi] = order
      self.camera_blocks[mi] = torch.stack(
          [self.jobs[i][1].to(torch.float32).reshape(4, 4) for i in order])
      for k, i in enumerate(order):
        self.clip[i] = self.clip_blocks[mi][k]
      start = 0
      for g in self.groups:
        n = sum(1 for i in g.indices if self.jobs[i][0] == mi)
        if n:
          g.transforms.append((mi, start, n))
        start += n

  def _autotune_groups(self):
    """Number of job groups from diffren_amd.autotune.tune_groups: the
    captured step is timed for each candidate count on this batch."""
    from diffren_amd import autotune

    def run_with(n):
      self._set_groups(n)
      for g in self.groups:
        self._plan_raster(g)
        self._plan_splat(g)
      return self.capture().replay

    n, self.streams_timings = autotune.tune_groups(
        self._tune_key(), run_with, len(self.jobs),
        cache_path=self.autotune_cache)
    self.graph = None
    return n

  def _tune_key(self):
    """Autotuner key of this batch's step (the raster autotuner's signature
    plus what else fixes the step)."""
    from diffren_amd import autotune
    sizes = [(w, h, self.meshes[mi]["triangles"].shape[0])
             for mi, _, w, h in self.jobs]
    return autotune.signature(sizes, self.L, 1) + json.dumps(
        [self.raster_config.tile_w, self.raster_config.tile_h, self.C,
         bool(self.fused_transform), bool(self.fixed_shift),
         self.shading_function is not None])

  def _autotune_tile_threads(self):
    """Tile-pass block size from diffren_amd.autotune.tune_tile_threads: the
    captured step is timed with each candidate on this batch and group
    count."""
    from diffren_amd import autotune
    cfg = self.raster_config

    def run_with(t):
      self.raster_config = _native.RasterConfig(
          cfg.tile_w, cfg.tile_h, cfg.bin_capacity, cfg.flags, t)
      return self.capture().replay

    t, self.tile_threads_timings = autotune.tune_tile_threads(
        self._tune_key() + f"/{len(self.groups)}", run_with,
        cache_path=self.autotune_cache)
    self.raster_config = _native.RasterConfig(
        cfg.tile_w, cfg.tile_h, cfg.bin_capacity, cfg.flags, t)
    self.graph = None

  def _split(self, n):
    """Splits jobs into n groups of balanced pixel counts (greedy)."""
    n = min(n, len(self.jobs))
    order = sorted(range(len(self.jobs)),
                   key=lambda i: -self.jobs[i][2] * self.jobs[i][3])
    load = [0] * n
    parts = [[] for _ in range(n)]
    for i in order:
      k = load.index(min(load))
      parts[k].append(i)
      load[k] += self.jobs[i][2] * self.jobs[i][3]
    return [_Group(sorted(p)) for p in parts if p]

  def _raster_table(self, indices):
    table = (_native.RasterJob * len(indices))()
    for k, i in enumerate(indices):
      mi, _, w, h = self.jobs[i]
      m = self.meshes[mi]
      if self.fused_transform:
        table[k].vertices = m["vertices"].data_ptr()
        table[k].camera = self.cameras[i].data_ptr()
      else:
        table[k].vertices = self.clip[i].data_ptr()
      table[k].triangles = m["triangles"].data_ptr()
      table[k].triangle_ids = self._ids[i].data_ptr()
      table[k].z_buffer = self.z[i].data_ptr()
      table[k].barycentrics = self.barys[i].data_ptr()
      table[k].num_vertices = m["vertices"].shape[0]
      table[k].num_triangles = m["triangles"].shape[0]
      table[k].width = w
      table[k].height = h
    return table

  def _autotune(self):
    """Tile shape and bin capacity from diffren_amd.autotune."""
    from diffren_amd import autotune
    self.transform()
    sizes = [(w, h, self.meshes[mi]["triangles"].shape[0])
             for mi, _, w, h in self.jobs]
    tw, th, cap, timings = autotune.tune_raster(
        self._raster_table(list(range(len(self.jobs)))), len(self.jobs),
        self.L, 1, self.device, sizes, cache_path=self.autotune_cache,
        flags=self._raster_flags)
    self.autotune_timings = timings
    return (tw, th), cap

  def _plan_raster(self, g):
    lib = _native.lib()
    table = self._raster_table(g.indices)
    tt, tiles = ctypes.c_int64(), ctypes.c_int64()
    _native.check(lib.dr_raster_jobs_prepare(table, len(g.indices),
                                             ctypes.byref(self.raster_config),
                                             ctypes.byref(tt),
                                             ctypes.byref(tiles)),
                  "dr_raster_jobs_prepare")
    g.raster_table = _native.table_to_device(table, self.device)
    g.total_triangles, g.raster_tiles = tt.value, tiles.value
    g.raster_ws = _native.workspace(
        lib.dr_raster_workspace_size(tt.value, tiles.value,
                                     ctypes.byref(self.raster_config)),
        self.device)

  def _plan_splat(self, g):
    lib = _native.lib()
    table = (_native.SplatJob * len(g.indices))()
    for k, i in enumerate(g.indices):
      mi, _, w, h = self.jobs[i]
      m = self.meshes[mi]
      j = table[k]
      j.positions = m["vertices"].data_ptr()
      if not self.fused_transform:
        j.clip_vertices = self.clip[i].data_ptr()
      if self.d_cameras is not None:
        j.d_camera = (self.d_cameras_fixed if self.d_cameras_fixed is not None
                      else self.d_cameras)[i].data_ptr()
      j.triangles = m["triangles"].data_ptr()
      if self.shading_function is not None:
        j.shaded = self.shaded[i].data_ptr()
        j.d_shaded = self.d_shaded[i].data_ptr()
      else:
        j.colors = m["colors"].data_ptr()
      j.camera = self.cameras[i].data_ptr()
      j.triangle_ids = self._ids[i].data_ptr()
      j.z_buffer = self.z[i].data_ptr()
      j.barycentrics = self.barys[i].data_ptr()
      j.image = self.images[i].data_ptr()
      j.target = self.targets[i].data_ptr()
      j.d_image = None
      j.grad_scratch = self.scratch[i].data_ptr()
      j.d_records = (self.records_fixed[mi] if self.records_fixed else
                     self.records[mi]).data_ptr()
      j.fixed_shift = self.fixed_shift
      if self.accum:
        j.accum = self.accum[i].data_ptr()
        j.norm = self.norm[i].data_ptr()
        j.accum_weights = self.accum_weights[i].data_ptr()
      j.num_vertices = m["vertices"].shape[0]
      j.num_triangles = m["triangles"].shape[0]
      j.width = w
      j.height = h
    tiles = ctypes.c_int64()
    _native.check(lib.dr_splat_jobs_prepare(table, len(g.indices), self.C,
                                            ctypes.byref(tiles)),
                  "dr_splat_jobs_prepare")
    g.splat_table = _native.table_to_device(table, self.device)
    g.splat_tiles = tiles.value
    g.splat_ws = _native.workspace(
        lib.dr_render_splat_workspace_size(tiles.value), self.device)

  # -- stages ----------------------------------------------------------------

  def transform(self):
    """Clip-space vertices of every job (one launch per mesh)."""
    if self.fused_transform:
      return
    for mi, views in self.mesh_views.items():
      self._transform_views(mi, 0, len(views))

  def _transform_views(self, mi, start, count):
    """Views start .. start + count - 1 of mesh mi's clip block."""
    v = self.meshes[mi]["vertices"]
    _native.check(
        _native.lib().dr_transform_homogeneous_multi(
            _native.ptr(self.camera_blocks[mi][start]), count, _native.ptr(v),
            v.shape[1], v.shape[0], _native.ptr(self.clip_blocks[mi][start]),
            _native.stream_handle()),
        "dr_transform_homogeneous_multi")

  def _rasterize(self, g, half=0):
    """half: 0 for the whole rasterizer, or _native.RASTER_BINS_ONLY /
    RASTER_TILES_ONLY for one half of it."""
    cfg = self.raster_config
    if half:
      cfg = _native.RasterConfig(cfg.tile_w, cfg.tile_h, cfg.bin_capacity,
                                 cfg.flags | half, cfg.tile_threads)
    _native.check(
        _native.lib().dr_rasterize_triangles_batched(
            _native.ptr(g.raster_table), len(g.indices), g.total_triangles,
            g.raster_tiles, self.L, 1, ctypes.byref(cfg),
            _native.ptr(g.raster_ws), g.raster_ws.numel(),
            _native.stream_handle()), "dr_rasterize_triangles_batched")

  def _splat_forward(self, g, grad_mode):
    _native.check(
        _native.lib().dr_render_splat_forward(
            _native.ptr(g.splat_table), len(g.indices), g.splat_tiles, self.L,
            self.C,
            grad_mode | (_native.RENDER_ACCUM_BUFFERS if self.accum else 0),
            self.loss_scale, _native.ptr(g.loss_slot),
            _native.ptr(g.splat_ws), g.splat_ws.numel(),
            _native.stream_handle()), "dr_render_splat_forward")

  def _splat_backward(self, g):
    _native.check(
        _native.lib().dr_render_splat_backward(
            _native.ptr(g.splat_table), len(g.indices), g.splat_tiles, self.L,
            self.C,
            _native.RENDER_SHADED if self.shading_function is not None else 0,
            _native.stream_handle()), "dr_render_splat_backward")

  def rasterize(self):
    for g in self.groups:
      self._rasterize(g)

  def splat_forward(self, grad_mode=1):
    for g in self.groups:
      self._splat_forward(g, grad_mode)
    self._sum_loss()

  def splat_backward(self):
    for g in self.groups:
      self._splat_backward(g)
    self._convert_fixed()

  def _sum_loss(self):
    # Fixed order: elementwise adds, no reduction kernel.
    self.loss.copy_(self.loss_parts[0:1])
    for k in range(1, len(self.groups)):
      self.loss.add_(self.loss_parts[k:k + 1])

  def _convert_fixed(self):
    """Deterministic mode: integer gradient sums -> float32 records."""
    if not self.fixed_shift:
      return
    lib = _native.lib()
    pairs = list(zip(self.records_fixed, self.records))
    if self.d_cameras_fixed is not None:
      pairs.append((self.d_cameras_fixed, self.d_cameras))
    for src, dst in pairs:
      _native.check(lib.dr_fixed_to_float(
          _native.ptr(src), src.numel(), self.fixed_shift, _native.ptr(dst),
          _native.stream_handle()), "dr_fixed_to_float")

  @property
  def ids(self):
    """Triangle ids [L, h, w] of each job (-1 on background).  The fused
    raster writes a pixel's layers only up to its first empty one
    (DR_RASTER_PREFIX); reading them here first sets the layers behind it
    to -1, in place, on the current stream."""
    if self._raster_flags & _native.RASTER_PREFIX:
      for t in self._ids:
        for l in range(1, self.L):
          t[l].masked_fill_(t[l - 1] < 0, -1)
    return self._ids

  def zero_grads(self):
    # Fill kernels, not memsets, so the step can be graph-captured.
    for t in self.records + self.records_fixed:
      t.zero_()
    for t in (self.d_cameras, self.d_cameras_fixed):
      if t is not None:
        t.zero_()
    self.loss_parts.zero_()
    self.loss.zero_()

  def _group_raster(self, g):
    for mi, start, count in g.transforms:
      self._transform_views(mi, start, count)
    self._rasterize(g)
    self._status_host[self.groups.index(g)].copy_(g.raster_ws[:16],
                                                  non_blocking=True)

  def _group_splat(self, g, grad_mode):
    self._splat_forward(g, grad_mode)
    if grad_mode:
      self._splat_backward(g)

  def _group_step(self, g, grad_mode):
    self._group_raster(g)
    self._group_splat(g, grad_mode)

  def _on_streams(self, fn):
    """fn(group) for every group, each on its own stream (forked from and
    joined back into the current stream)."""
    if not self.streams:
      fn(self.groups[0])
      return
    cur = torch.cuda.current_stream()
    for s in self.streams:
      s.wait_stream(cur)
    for g, s in zip(self.groups, self.streams):
      with torch.cuda.stream(s):
        fn(g)
    for s in self.streams:
      cur.wait_stream(s)

  def _shade(self, grad):
    """Interpolates every job's attributes, runs the shading function and
    stores its output in the job's shaded buffer; returns the outputs that
    carry gradients (for the backward)."""
    outs = []
    with torch.set_grad_enabled(grad):
      for i, (mi, _, w, h) in enumerate(self.jobs):
        m = self.meshes[mi]
        fb = framebuffer.Framebuffer(
            barycentrics=self.barys[i], triangle_id=self._ids[i].unsqueeze(-1),
            triangles=m["triangles"], z_buffer=self.z[i])
        # Background 0 (render.py:124), written by the kernel itself (no
        # background tensor: nothing is copied from the host, so the step
        # can be graph-captured).
        attrs = interpolate.interpolate_vertex_attributes(m["attributes"], fb)
        out = self.shading_function(attrs)
        if tuple(out.shape) != (self.L, h, w, self.C):
          raise ValueError(
              f"shading_function returned {tuple(out.shape)} for job {i}, "
              f"expected {(self.L, h, w, self.C)}")
        self.shaded[i].copy_(out.detach())
        if grad and out.requires_grad:
          outs.append((out, self.d_shaded[i]))
    return outs

  def _attribute_grads(self, outs):
    """d_attributes from the shaded pixels' gradient (written by the source
    pass into d_shaded) through the shader and the interpolation."""
    want = [(mi, k, a) for mi, m in enumerate(self.meshes)
            for k, a in m["attributes"].items() if a.requires_grad]
    if not want:
      return
    grads = (torch.autograd.grad([o for o, _ in outs], [a for _, _, a in want],
                                 [d for _, d in outs], allow_unused=True)
             if outs else [None] * len(want))
    for (mi, k, _), g in zip(want, grads):
      if g is None:
        self.d_attributes[mi][k].zero_()
      else:
        self.d_attributes[mi][k].copy_(g)

  def _run_groups(self, grad_mode):
    if self.shading_function is None:
      self._on_streams(lambda g: self._group_step(g, grad_mode))
    else:
      self._on_streams(self._group_raster)
      outs = self._shade(grad_mode != 0)
      self._on_streams(lambda g: self._group_splat(g, grad_mode))
      if grad_mode:
        self._attribute_grads(outs)
    if grad_mode == 1:
      self._sum_loss()
    if grad_mode:
      self._convert_fixed()

  def _check_overflow(self):
    """Regrows the bins if the last completed step overflowed them."""
    ev = self._status_event
    if ev is None or not ev.query():
      return
    self._status_event = None
    st = self._status_host.view(torch.int32).reshape(len(self.groups), 4)
    st = st.tolist()
    if not any(row[1] for row in st):
      return
    need = max((row[0] & 0xffffffff) | (row[2] << 32) for row in st)
    cap = max(int(2 * need) + 4096, int(self.raster_config.bin_capacity))
    self.raster_config = _native.RasterConfig(
        self.raster_config.tile_w, self.raster_config.tile_h, cap,
        self.raster_config.flags, self.raster_config.tile_threads)
    for g in self.groups:
      self._plan_raster(g)
    self.regrows += 1
    if self.graph is not None:
      self.graph = None
      self.capture(forward_only=self._graph_forward_only)

  def _mark_step(self):
    ev = torch.cuda.Event()
    ev.record()
    self._status_event = ev

  def step(self):
    """One forward + L2 loss + backward over all jobs (enqueue only)."""
    self._check_overflow()
    self._enqueue_step()
    self._mark_step()

  def _enqueue_step(self):
    self.zero_grads()
    self._run_groups(1)

  def render(self):
    """Forward only: fills self.images."""
    self._check_overflow()
    self._enqueue_render()
    self._mark_step()

  def _enqueue_render(self):
    self._run_groups(0)

  # -- graphs ----------------------------------------------------------------

  def capture(self, warmup=2, forward_only=False):
    """Captures step() (or render() when forward_only) in a HIP graph;
    replay() runs it.  Returns the graph."""
    fn = self._enqueue_render if forward_only else self._enqueue_step
    self._graph_forward_only = forward_only
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
      for _ in range(warmup):
        fn()
    torch.cuda.current_stream().wait_stream(s)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
      fn()
    torch.cuda.synchronize()
    self.graph = g
    return g

  def replay(self):
    self._check_overflow()
    if self.graph is None:
      self.capture()
    self.graph.replay()
    self._mark_step()

  def bin_stats(self):
    """(entries needed, overflowed) of the last raster launch of each group,
    summed / or-ed."""
    need, over = 0, False
    for g in self.groups:
      n, o = ctypes.c_int64(), ctypes.c_int32()
      _native.check(_native.lib().dr_raster_bin_stats(
          _native.ptr(g.raster_ws), ctypes.byref(n), ctypes.byref(o),
          _native.stream_handle()), "dr_raster_bin_stats")
      need += n.value
      over = over or bool(o.value)
    return need, over
