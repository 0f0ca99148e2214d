"""Batched, fused rasterize-then-splat rendering with a hand-written backward.

`BatchedSplatRenderer` renders many (mesh, camera, resolution) jobs -- of
uneven triangle counts and image sizes -- through one set of launches per
stage:
  transform (per job) -> batched tile rasterizer (one launch per stage for all
  jobs) -> fused target pass (interpolate + splat + over + L2 gradient) ->
  fused source pass (splat VJP + interpolate VJP + transform VJP, scattering
  into d colours and d vertices).
All buffers are allocated once at construction, so a step allocates nothing
and can be captured in a HIP graph (`capture()` / `replay()`).

This is the batched counterpart of calling render.render_triangles(...,
shading_function=lambda a: a["colors"], compositing_mode=SPLAT_OVER) once per
view (render.py:28-140 under jax.vmap, which issues one rasterizer launch per
image: rasterize_triangles_xla.py:68-75) followed by jax.grad of an L2 loss.
"""
import ctypes

import torch

from diffren_amd import _native
from diffren_amd import rasterize


class BatchedSplatRenderer:
  """Preplanned fused forward+backward over a fixed batch of render jobs.

  meshes:  list of dicts with device tensors "vertices" [V,3] f32,
           "triangles" [T,3] i32, "colors" [V,C] f32; jobs reference a mesh
           by index, so several views of one mesh share its gradient buffers.
  jobs:    list of (mesh_index, camera [4,4] device f32, width, height).
  """

  def __init__(self, meshes, jobs, num_layers, tile=None, bin_capacity=0,
               loss_scale=None, autotune_cache=None):
    self.meshes = meshes
    self.jobs = jobs
    self.L = int(num_layers)
    dev = meshes[0]["vertices"].device
    self.device = dev
    C = meshes[0]["colors"].shape[1]
    for m in meshes:
      if m["colors"].shape[1] != C:
        raise ValueError("all meshes need the same channel count")
    self.C = C
    # Per-mesh gradient records [V, R]: d position in cols 0..2, d colour in
    # cols 4..4+C-1 (one contiguous record per vertex for grouped atomics).
    R = _native.lib().dr_render_record_size(C)
    if R <= 0:
      raise ValueError(f"unsupported channel count {C}")
    self.records = [torch.zeros((m["vertices"].shape[0], R),
                                dtype=torch.float32, device=dev)
                    for m in meshes]
    self.d_vertices = [r[:, 0:3] for r in self.records]
    self.d_colors = [r[:, 4:4 + C] for r in self.records]
    self.loss = torch.zeros(1, dtype=torch.float32, device=dev)
    self.clip, self.ids, self.z, self.barys = [], [], [], []
    self.images, self.targets, self.scratch = [], [], []
    # Clip-space vertices: one [views, V, 4] block per mesh, one transform
    # launch per mesh.
    self.mesh_views = {}
    for i, (mi, _, _, _) in enumerate(jobs):
      self.mesh_views.setdefault(mi, []).append(i)
    self.clip_blocks, self.camera_blocks = {}, {}
    for mi, views in self.mesh_views.items():
      V = meshes[mi]["vertices"].shape[0]
      self.clip_blocks[mi] = torch.empty((len(views), V, 4),
                                         dtype=torch.float32, device=dev)
      self.camera_blocks[mi] = torch.stack(
          [jobs[i][1].to(torch.float32).reshape(4, 4) for i in views])
    slot = {}
    for i, (mi, _, _, _) in enumerate(jobs):
      slot[i] = self.mesh_views[mi].index(i)
    npix = 0
    for i, (mi, cam, w, h) in enumerate(jobs):
      self.clip.append(self.clip_blocks[mi][slot[i]])
      self.ids.append(torch.empty((self.L, h, w), dtype=torch.int32,
                                  device=dev))
      self.z.append(torch.empty((self.L, h, w), dtype=torch.float32,
                                device=dev))
      self.barys.append(torch.empty((self.L, h, w, 3), dtype=torch.float32,
                                    device=dev))
      self.images.append(torch.empty((h, w, C), dtype=torch.float32,
                                      device=dev))
      self.targets.append(torch.zeros((h, w, C), dtype=torch.float32,
                                      device=dev))
      self.scratch.append(torch.empty(
          (h, w, _native.lib().dr_render_scratch_floats(C)),
          dtype=torch.float32, device=dev))
      npix += w * h * C
    self.loss_scale = 1.0 / npix if loss_scale is None else float(loss_scale)
    self.cameras = [j[1].to(torch.float32).contiguous() for j in jobs]
    self.autotune_cache = autotune_cache
    self.autotune_timings = None
    self._plan_raster(tile, bin_capacity)
    self._plan_splat()
    self.graph = None

  # -- planning --------------------------------------------------------------

  def _plan_raster(self, tile, bin_capacity):
    lib = _native.lib()
    if tile == "auto":
      tile, bin_capacity = self._autotune()
    # The fused passes read ids first, so background z / barycentrics are
    # never written (DR_RASTER_SPARSE).
    self.raster_config = rasterize._config(tile, bin_capacity,
                                           _native.RASTER_SPARSE)
    n = len(self.jobs)
    table = self._raster_table()
    tt, tiles = ctypes.c_int64(), ctypes.c_int64()
    _native.check(lib.dr_raster_jobs_prepare(table, n,
                                             ctypes.byref(self.raster_config),
                                             ctypes.byref(tt),
                                             ctypes.byref(tiles)),
                  "dr_raster_jobs_prepare")
    self.raster_table = _native.table_to_device(table, self.device)
    self.total_triangles, self.raster_tiles = tt.value, tiles.value
    nbytes = lib.dr_raster_workspace_size(tt.value, tiles.value,
                                          ctypes.byref(self.raster_config))
    self.raster_ws = _native.workspace(nbytes, self.device)

  def _raster_table(self):
    n = len(self.jobs)
    table = (_native.RasterJob * n)()
    for i, (mi, _, w, h) in enumerate(self.jobs):
      m = self.meshes[mi]
      table[i].vertices = self.clip[i].data_ptr()
      table[i].triangles = m["triangles"].data_ptr()
      table[i].triangle_ids = self.ids[i].data_ptr()
      table[i].z_buffer = self.z[i].data_ptr()
      table[i].barycentrics = self.barys[i].data_ptr()
      table[i].num_vertices = m["vertices"].shape[0]
      table[i].num_triangles = m["triangles"].shape[0]
      table[i].width = w
      table[i].height = h
    return table

  def _autotune(self):
    """Tile shape and bin capacity from diffren_amd.autotune."""
    from diffren_amd import autotune
    self.transform()
    sizes = [(w, h, self.meshes[mi]["triangles"].shape[0])
             for mi, _, w, h in self.jobs]
    tw, th, cap, timings = autotune.tune_raster(
        self._raster_table(), len(self.jobs), self.L, 1, self.device, sizes,
        cache_path=self.autotune_cache, flags=_native.RASTER_SPARSE)
    self.autotune_timings = timings
    return (tw, th), cap

  def _plan_splat(self):
    lib = _native.lib()
    n = len(self.jobs)
    table = (_native.SplatJob * n)()
    for i, (mi, _, w, h) in enumerate(self.jobs):
      m = self.meshes[mi]
      j = table[i]
      j.clip_vertices = self.clip[i].data_ptr()
      j.triangles = m["triangles"].data_ptr()
      j.colors = m["colors"].data_ptr()
      j.camera = self.cameras[i].data_ptr()
      j.triangle_ids = self.ids[i].data_ptr()
      j.z_buffer = self.z[i].data_ptr()
      j.barycentrics = self.barys[i].data_ptr()
      j.image = self.images[i].data_ptr()
      j.target = self.targets[i].data_ptr()
      j.d_image = None
      j.grad_scratch = self.scratch[i].data_ptr()
      j.d_records = self.records[mi].data_ptr()
      j.num_vertices = m["vertices"].shape[0]
      j.num_triangles = m["triangles"].shape[0]
      j.width = w
      j.height = h
    tiles = ctypes.c_int64()
    _native.check(lib.dr_splat_jobs_prepare(table, n, self.C,
                                            ctypes.byref(tiles)),
                  "dr_splat_jobs_prepare")
    self.splat_table = _native.table_to_device(table, self.device)
    self.splat_tiles = tiles.value
    self.splat_ws = _native.workspace(
        lib.dr_render_splat_workspace_size(tiles.value), self.device)

  # -- stages ----------------------------------------------------------------

  def transform(self):
    lib = _native.lib()
    s = _native.stream_handle()
    for mi, views in self.mesh_views.items():
      v = self.meshes[mi]["vertices"]
      _native.check(
          lib.dr_transform_homogeneous_multi(
              _native.ptr(self.camera_blocks[mi]), len(views), _native.ptr(v),
              v.shape[1], v.shape[0], _native.ptr(self.clip_blocks[mi]), s),
          "dr_transform_homogeneous_multi")

  def rasterize(self):
    lib = _native.lib()
    _native.check(
        lib.dr_rasterize_triangles_batched(
            _native.ptr(self.raster_table), len(self.jobs),
            self.total_triangles, self.raster_tiles, self.L, 1,
            ctypes.byref(self.raster_config), _native.ptr(self.raster_ws),
            self.raster_ws.numel(), _native.stream_handle()),
        "dr_rasterize_triangles_batched")

  def splat_forward(self, grad_mode=1):
    _native.check(
        _native.lib().dr_render_splat_forward(
            _native.ptr(self.splat_table), len(self.jobs), self.splat_tiles,
            self.L, self.C, grad_mode, self.loss_scale,
            _native.ptr(self.loss), _native.ptr(self.splat_ws),
            self.splat_ws.numel(), _native.stream_handle()),
        "dr_render_splat_forward")

  def splat_backward(self):
    _native.check(
        _native.lib().dr_render_splat_backward(
            _native.ptr(self.splat_table), len(self.jobs), self.splat_tiles,
            self.L, self.C, _native.stream_handle()),
        "dr_render_splat_backward")

  def zero_grads(self):
    # Fill kernels, not memsets, so the step can be graph-captured.
    for t in self.records:
      t.zero_()
    self.loss.zero_()

  def step(self):
    """One forward + L2 loss + backward over all jobs (enqueue only)."""
    self.zero_grads()
    self.transform()
    self.rasterize()
    self.splat_forward(grad_mode=1)
    self.splat_backward()

  def render(self):
    """Forward only: fills self.images."""
    self.transform()
    self.rasterize()
    self.splat_forward(grad_mode=0)

  # -- graphs ----------------------------------------------------------------

  def capture(self, warmup=2):
    """Captures step() in a HIP graph; returns it."""
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
      for _ in range(warmup):
        self.step()
    torch.cuda.current_stream().wait_stream(s)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
      self.step()
    torch.cuda.synchronize()
    self.graph = g
    return g

  def replay(self):
    if self.graph is None:
      self.capture()
    self.graph.replay()

  def bin_stats(self):
    need, over = ctypes.c_int64(), ctypes.c_int32()
    _native.check(_native.lib().dr_raster_bin_stats(
        _native.ptr(self.raster_ws), ctypes.byref(need), ctypes.byref(over),
        _native.stream_handle()), "dr_raster_bin_stats")
    return need.value, bool(over.value)
