"""Rasterization: the layered z-buffer primitive and its framebuffer wrapper.

`rasterize_triangles_primitive` mirrors the reference's JAX primitive
`rasterize_triangles_xla.rasterize_triangles(vertices, triangles, image_width,
image_height, num_layers, face_culling_mode)`
(diffren/jax/internal/kernels/rasterize_triangles_xla.py:32-78): same arguments,
same outputs (triangle_id [L,H,W] i32, z [L,H,W] f32, barycentrics [L,H,W,3]
f32), background id -1 / z 1 / barycentrics 0 as the reference CPU handler
writes them (rasterize_triangles_cpu.cc:55-60).  It runs the gfx950 tile
rasterizer through the C ABI; results are bit-identical to the reference CPU
kernel.  Leading batch dimensions are accepted: `batched=False` launches one
image at a time (the reference's vmap_method='sequential_unrolled', :68-75),
`batched=True` rasterizes all images in one set of launches.

`rasterize_triangles` mirrors diffren.jax.internal.rasterize.rasterize_triangles
(diffren/jax/internal/rasterize.py:26-124), including compute_diff_barys=True:
`differentiable_barycentrics` (:127-162) recomputes the barycentrics from the
clip-space vertices on the GPU with a hand-written backward, so gradients
reach the vertices through every attribute interpolated with them.
`normalized_pixel_coordinates`, `compute_triangle_matrix` and
`compute_barycentric_coordinates` (:165-256) are the same functions on torch
tensors, for callers that use them directly.
"""
import ctypes
import dataclasses

import torch

from diffren_amd import _native
from diffren_amd import constants
from diffren_amd import framebuffer as fb
from diffren_amd import transforms

DEFAULT_TILE = (32, 8)
# The rasterizer keeps num_layers x tile pixels 64-bit depth keys per tile in
# 64 KiB of LDS: up to 32 layers on the default 32x8 tiles, 64 on 8x8 tiles.
MAX_LAYERS = 64


def _config(tile=None, bin_capacity=0, flags=0, num_layers=1):
  tw, th = tile or DEFAULT_TILE
  if tile is None and num_layers * tw * th * 8 > 65536:
    tw, th = 8, 8
  return _native.RasterConfig(int(tw), int(th), int(bin_capacity),
                              int(flags), 0)


def _check_inputs(vertices, triangles, image_width, image_height, num_layers,
                  face_culling_mode):
  if vertices.shape[-1] != 4:
    raise ValueError("vertices must be [..., V, 4] clip-space xyzw")
  if triangles.shape[-1] != 3:
    raise ValueError("triangles must be [..., T, 3]")
  if not image_width > 0:
    raise ValueError(f"image_width must be > 0 but is {image_width}.")
  if not image_height > 0:
    raise ValueError(f"image_height must be > 0 but is {image_height}.")
  if not 0 < num_layers <= MAX_LAYERS:
    raise ValueError(f"num_layers must be in [1, {MAX_LAYERS}] but is "
                     f"{num_layers}.")
  if int(face_culling_mode) not in (0, 1, 2):
    raise ValueError(f"invalid face_culling_mode {face_culling_mode}")


def _rasterize_one(vertices, triangles, image_width, image_height, num_layers,
                   face_culling_mode, config, out=None, want_barys=True):
  dev = vertices.device
  L, H, W = num_layers, image_height, image_width
  if out is None:
    ids = torch.empty((L, H, W), dtype=torch.int32, device=dev)
    z = torch.empty((L, H, W), dtype=torch.float32, device=dev)
    b = (torch.empty((L, H, W, 3), dtype=torch.float32, device=dev)
         if want_barys else None)
  else:
    ids, z, b = out
  lib = _native.lib()
  nbytes = lib.dr_raster_workspace_size(
      triangles.shape[0],
      -(-W // config.tile_w) * -(-H // config.tile_h), ctypes.byref(config))
  ws = _native.workspace(nbytes, dev)
  rc = lib.dr_rasterize_triangles(
      _native.ptr(vertices), vertices.shape[0], _native.ptr(triangles),
      triangles.shape[0], W, H, L, int(face_culling_mode), _native.ptr(ids),
      _native.ptr(z), _native.ptr(b), ctypes.byref(config),
      _native.ptr(ws), ws.numel(), _native.stream_handle())
  _native.check(rc, "dr_rasterize_triangles")
  return ids, z, b


def rasterize_batched(vertices_list, triangles_list, sizes, num_layers,
                      face_culling_mode, tile=None, bin_capacity=0,
                      want_barys=True):
  """Rasterizes several (mesh, size) jobs in one set of launches.

  vertices_list[i] [Vi,4] f32 and triangles_list[i] [Ti,3] i32 are device
  tensors; sizes[i] = (width, height).  Returns a list of (ids, z, barys).
  """
  config = _config(tile, bin_capacity, num_layers=num_layers)
  n = len(vertices_list)
  dev = vertices_list[0].device
  jobs = (_native.RasterJob * n)()
  outs = []
  for i, (v, t, (w, h)) in enumerate(zip(vertices_list, triangles_list,
                                         sizes)):
    v = v.contiguous()
    t = t.contiguous()
    _native.require_device(v, t)
    ids = torch.empty((num_layers, h, w), dtype=torch.int32, device=dev)
    z = torch.empty((num_layers, h, w), dtype=torch.float32, device=dev)
    b = (torch.empty((num_layers, h, w, 3), dtype=torch.float32, device=dev)
         if want_barys else None)
    outs.append((ids, z, b, v, t))
    jobs[i].vertices = v.data_ptr()
    jobs[i].triangles = t.data_ptr()
    jobs[i].triangle_ids = ids.data_ptr()
    jobs[i].z_buffer = z.data_ptr()
    jobs[i].barycentrics = b.data_ptr() if b is not None else None
    jobs[i].num_vertices = v.shape[0]
    jobs[i].num_triangles = t.shape[0]
    jobs[i].width = w
    jobs[i].height = h
  lib = _native.lib()
  tt = ctypes.c_int64()
  tiles = ctypes.c_int64()
  _native.check(lib.dr_raster_jobs_prepare(jobs, n, ctypes.byref(config),
                                           ctypes.byref(tt),
                                           ctypes.byref(tiles)),
                "dr_raster_jobs_prepare")
  table_host = torch.frombuffer(bytearray(bytes(jobs)), dtype=torch.uint8)
  table = table_host.to(dev)
  nbytes = lib.dr_raster_workspace_size(tt.value, tiles.value,
                                        ctypes.byref(config))
  ws = _native.workspace(nbytes, dev)
  rc = lib.dr_rasterize_triangles_batched(
      _native.ptr(table), n, tt.value, tiles.value, num_layers,
      int(face_culling_mode), ctypes.byref(config), _native.ptr(ws),
      ws.numel(), _native.stream_handle())
  _native.check(rc, "dr_rasterize_triangles_batched")
  # Keep the job table and inputs alive until the launches have run.
  torch.cuda.current_stream().synchronize()
  return [(ids, z, b) for ids, z, b, _, _ in outs]


def rasterize_triangles_primitive(vertices, triangles, image_width,
                                  image_height, num_layers, face_culling_mode,
                                  batched=False, tile=None, bin_capacity=0):
  """See module docstring; rasterize_triangles_xla.py:32-78."""
  _check_inputs(vertices, triangles, image_width, image_height, num_layers,
                face_culling_mode)
  _native.require_device(vertices, triangles)
  vertices = vertices.detach().to(torch.float32).contiguous()
  triangles = triangles.to(torch.int32).contiguous()
  config = _config(tile, bin_capacity, num_layers=num_layers)
  if vertices.dim() == 2 and triangles.dim() == 2:
    return _rasterize_one(vertices, triangles, image_width, image_height,
                          num_layers, face_culling_mode, config)
  batch = vertices.shape[:-2] if vertices.dim() > 2 else triangles.shape[:-2]
  nb = 1
  for s in batch:
    nb *= s
  vs = vertices.reshape(-1, *vertices.shape[-2:]).expand(
      nb, *vertices.shape[-2:]) if vertices.dim() > 2 else [vertices] * nb
  ts = triangles.reshape(-1, *triangles.shape[-2:]) if triangles.dim() > 2 \
      else [triangles] * nb
  if batched:
    res = rasterize_batched([vs[i] for i in range(nb)],
                            [ts[i] for i in range(nb)],
                            [(image_width, image_height)] * nb, num_layers,
                            face_culling_mode, tile, bin_capacity)
  else:
    res = [_rasterize_one(vs[i], ts[i], image_width, image_height, num_layers,
                          face_culling_mode, config) for i in range(nb)]
  shape = tuple(batch) + (num_layers, image_height, image_width)
  ids = torch.stack([r[0] for r in res]).reshape(shape)
  z = torch.stack([r[1] for r in res]).reshape(shape)
  b = torch.stack([r[2] for r in res]).reshape(shape + (3,))
  return ids, z, b


def rasterize_triangles(vertices, triangles, camera, image_width,
                        image_height, num_layers=1,
                        face_culling_mode=constants.FaceCullingMode.NONE,
                        compute_diff_barys=False):
  """Rasterizes triangles into a Framebuffer (rasterize.py:26-124).

  vertices: [V, 3] or [V, 4] f32 (clip space if camera is None);
  triangles: [T, 3] i32; camera: [4, 4] or None.
  """
  if vertices.dim() != 2:
    raise ValueError(f"vertices must have shape [V, 3|4] but has shape "
                     f"{tuple(vertices.shape)}")
  if triangles.dim() != 2 or triangles.shape[-1] != 3:
    raise ValueError(f"triangles must have shape [T, 3] but has shape "
                     f"{tuple(triangles.shape)}")
  if camera is not None:
    if tuple(camera.shape) != (4, 4):
      raise ValueError(f"camera must have shape [4, 4] but has shape "
                       f"{tuple(camera.shape)}")
    if vertices.shape[-1] not in (3, 4):
      raise ValueError("vertices must be xyz or xyzw")
    clip = transforms.transform_homogeneous(camera, vertices)
  else:
    if vertices.shape[-1] != 4:
      raise ValueError("vertices must be [V, 4] when camera is None")
    clip = vertices
  tris = triangles.to(torch.int32)
  ids, z, b = rasterize_triangles_primitive(clip.detach(), tris, image_width,
                                            image_height, num_layers,
                                            face_culling_mode)
  out = fb.Framebuffer(barycentrics=b, triangle_id=ids.unsqueeze(-1),
                       triangles=tris, z_buffer=z)
  if compute_diff_barys:
    out = differentiable_barycentrics(out, clip, tris)
  return out


class _DiffBarycentrics(torch.autograd.Function):
  """dr_diff_barycentrics_fwd / _bwd: barycentrics [L, H, W, 3] as a
  differentiable function of the clip-space vertices [V, 4]."""

  @staticmethod
  def forward(ctx, clip, triangles, ids):
    c = clip.detach().to(torch.float32).contiguous()
    L, H, W = ids.shape
    out = torch.empty((L, H, W, 3), dtype=torch.float32, device=c.device)
    _native.check(
        _native.lib().dr_diff_barycentrics_fwd(
            _native.ptr(c), c.shape[0], _native.ptr(triangles),
            triangles.shape[0], _native.ptr(ids), W, H, L, _native.ptr(out),
            _native.stream_handle()), "dr_diff_barycentrics_fwd")
    ctx.save_for_backward(c, triangles, ids)
    return out

  @staticmethod
  def backward(ctx, d_barys):
    c, triangles, ids = ctx.saved_tensors
    L, H, W = ids.shape
    d_clip = torch.zeros_like(c)
    d_barys = d_barys.to(torch.float32).contiguous()
    _native.check(
        _native.lib().dr_diff_barycentrics_bwd(
            _native.ptr(c), c.shape[0], _native.ptr(triangles),
            triangles.shape[0], _native.ptr(ids), W, H, L,
            _native.ptr(d_barys), _native.ptr(d_clip),
            _native.stream_handle()), "dr_diff_barycentrics_bwd")
    return d_clip, None, None


def differentiable_barycentrics(framebuffer, clip_space_vertices, triangles):
  """Framebuffer with barycentrics recomputed differentiably from the
  clip-space vertices [V, 4] (rasterize.py:127-162); zero on background."""
  if clip_space_vertices.dim() != 2 or clip_space_vertices.shape[-1] != 4:
    raise ValueError("clip_space_vertices must be [V, 4]")
  _native.require_device(clip_space_vertices)
  ids = framebuffer.triangle_id[..., 0].contiguous()
  if ids.dim() != 3:
    raise ValueError("differentiable_barycentrics takes one [L, H, W] "
                     "framebuffer")
  tris = triangles.to(torch.int32).contiguous()
  b = _DiffBarycentrics.apply(clip_space_vertices, tris, ids)
  return dataclasses.replace(framebuffer, barycentrics=b)


def normalized_pixel_coordinates(image_width, image_height, device=None):
  """(x, y) pixel-centre coordinates [H, W] in [-1, 1] (rasterize.py:165-189):
  (2i + 1) / size - 1 in float32."""
  x = (2 * torch.arange(image_width, device=device) + 1).to(
      torch.float32) / float(image_width) - 1
  y = (2 * torch.arange(image_height, device=device) + 1).to(
      torch.float32) / float(image_height) - 1
  yy, xx = torch.meshgrid(y, x, indexing="ij")
  return xx, yy


def compute_triangle_matrix(clip_space_vertices, triangle):
  """[3, 3] adjugate of the triangle's (x, y, w) columns times the sign of
  its determinant (rasterize.py:192-225)."""
  xyzw = clip_space_vertices[triangle.long(), :]
  xyw = torch.cat((xyzw[:, :2], xyzw[:, 3:4]), dim=-1).T
  d11 = xyw[1, 1] * xyw[2, 2] - xyw[1, 2] * xyw[2, 1]
  d21 = xyw[1, 2] * xyw[2, 0] - xyw[1, 0] * xyw[2, 2]
  d31 = xyw[1, 0] * xyw[2, 1] - xyw[1, 1] * xyw[2, 0]
  d12 = xyw[2, 1] * xyw[0, 2] - xyw[2, 2] * xyw[0, 1]
  d22 = xyw[2, 2] * xyw[0, 0] - xyw[2, 0] * xyw[0, 2]
  d32 = xyw[2, 0] * xyw[0, 1] - xyw[2, 1] * xyw[0, 0]
  d13 = xyw[0, 1] * xyw[1, 2] - xyw[0, 2] * xyw[1, 1]
  d23 = xyw[0, 2] * xyw[1, 0] - xyw[0, 0] * xyw[1, 2]
  d33 = xyw[0, 0] * xyw[1, 1] - xyw[0, 1] * xyw[1, 0]
  matrix = torch.stack([torch.stack([d11, d12, d13]),
                        torch.stack([d21, d22, d23]),
                        torch.stack([d31, d32, d33])])
  det = xyw[0, 0] * d11 + xyw[1, 0] * d12 + xyw[2, 0] * d13
  sign = torch.sign(det) + (det == 0).to(det.dtype)
  return sign * matrix


def compute_barycentric_coordinates(triangle_matrix, px, py):
  """[3] barycentrics of a point (rasterize.py:228-256), nan_to_num'ed."""
  b = (triangle_matrix[:, 0] * px + triangle_matrix[:, 1] * py +
       triangle_matrix[:, 2])
  return torch.nan_to_num(b / torch.sum(b))
