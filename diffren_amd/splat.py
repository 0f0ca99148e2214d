"""Rasterize-then-splat on the GPU (diffren/jax/internal/splat.py:26-229).

splat_at_pixel_centers(xyz_layers, color_layers, foreground_mask)
  -> (output [H,W,C], accum [3,H,W,C], normalized [3,H,W,C])
splat_shaded_pixels(vertices, camera_matrices, rasterized_fb, shaded_pixels,
                    return_accum_buffers)
Both keep the reference's signatures.  `output` has a hand-written backward
with respect to the positions and the colours (dr_splat_bwd for
splat_at_pixel_centers; the fused render kernels' target and source passes
for splat_shaded_pixels, which also give the camera gradient); the
accumulation buffers are returned for inspection without gradients.
"""
import ctypes

import torch

from diffren_amd import _native
from diffren_amd import interpolate
from diffren_amd import transforms


class _Splat(torch.autograd.Function):

  @staticmethod
  def forward(ctx, xyz, color, mask, want_buffers):
    xyz = xyz.detach().to(torch.float32).contiguous()
    color = color.detach().to(torch.float32).contiguous()
    mask = mask.detach().to(torch.float32).contiguous()
    L, H, W, C = color.shape
    dev = color.device
    out = torch.empty((H, W, C), dtype=torch.float32, device=dev)
    accum = norm = None
    if want_buffers:
      accum = torch.empty((3, H, W, C), dtype=torch.float32, device=dev)
      norm = torch.empty((3, H, W, C), dtype=torch.float32, device=dev)
    _native.check(
        _native.lib().dr_splat_fwd(
            _native.ptr(xyz), _native.ptr(color), _native.ptr(mask), L, H, W,
            C, _native.ptr(out), _native.ptr(accum), _native.ptr(norm), None,
            _native.stream_handle()), "dr_splat_fwd")
    ctx.save_for_backward(xyz, color, mask)
    if want_buffers:
      ctx.mark_non_differentiable(accum, norm)
      return out, accum, norm
    return out, None, None

  @staticmethod
  def backward(ctx, d_out, _d_accum, _d_norm):
    xyz, color, mask = ctx.saved_tensors
    L, H, W, C = color.shape
    d_out = d_out.to(torch.float32).contiguous()
    d_color = torch.empty_like(color)
    d_xyz = torch.empty_like(xyz)
    lib = _native.lib()
    ws = _native.workspace(lib.dr_splat_bwd_workspace_size(H, W, C),
                           color.device)
    _native.check(
        lib.dr_splat_bwd(_native.ptr(xyz), _native.ptr(color),
                         _native.ptr(mask), L, H, W, C, _native.ptr(d_out),
                         _native.ptr(d_color), _native.ptr(d_xyz),
                         _native.ptr(ws), ws.numel(), _native.stream_handle()),
        "dr_splat_bwd")
    return d_xyz, d_color, None, None


# Channel counts the splat kernels are instantiated for.
KERNEL_CHANNELS = tuple(range(1, 17)) + (20, 24, 28, 32)


def _with_alpha_last(color, k):
  """color [..., c] padded to k channels with zeros before the last (alpha)
  channel: the splat treats every channel independently except the last,
  which `over` uses as alpha for all of them (composite.py:33-37), so the
  padding changes no other channel and its outputs are dropped."""
  c = color.shape[-1]
  if k == c:
    return color
  zeros = color.new_zeros(color.shape[:-1] + (k - c,))
  return torch.cat([color[..., :-1], zeros, color[..., -1:]], -1)


def _splat_kernel(xyz, color, mask, want_buffers):
  """One kernel call for any channel count up to 32 (padded)."""
  c = color.shape[-1]
  k = next(n for n in KERNEL_CHANNELS if n >= c)
  out, accum, norm = _Splat.apply(xyz, _with_alpha_last(color, k), mask,
                                  want_buffers)
  if k == c:
    return out, accum, norm
  keep = lambda t: None if t is None else torch.cat(
      [t[..., :c - 1], t[..., -1:]], -1)
  return keep(out), keep(accum), keep(norm)


def splat_at_pixel_centers(xyz_layers, color_layers, foreground_mask,
                           return_buffers=True):
  """splat.py:26-175.  xyz [L,H,W,3], color [L,H,W,C], mask [L,H,W,1].

  Any channel count: counts the kernels are not instantiated for are padded
  up to the next one; above 32 the colour channels are splatted in groups of
  31, each with the alpha channel, whose outputs come from the first group
  (so the alpha gradient, a sum over all channels, is the sum of the
  groups')."""
  if xyz_layers.dim() != 4 or xyz_layers.shape[-1] != 3:
    raise ValueError("xyz_layers must be [L, H, W, 3]")
  if color_layers.dim() != 4 or color_layers.shape[:3] != xyz_layers.shape[:3]:
    raise ValueError("color_layers must be [L, H, W, C] matching xyz_layers")
  _native.require_device(xyz_layers, color_layers, foreground_mask)
  mask = foreground_mask.reshape(xyz_layers.shape[:3])
  c = color_layers.shape[-1]
  if c <= KERNEL_CHANNELS[-1]:
    return _splat_kernel(xyz_layers, color_layers, mask, return_buffers)
  alpha = color_layers[..., -1:]
  parts = []
  for c0 in range(0, c - 1, 31):
    group = torch.cat([color_layers[..., c0:min(c0 + 31, c - 1)], alpha], -1)
    parts.append(_splat_kernel(xyz_layers, group, mask, return_buffers))

  def join(i):
    if parts[0][i] is None:
      return None
    return torch.cat([p[i][..., :-1] for p in parts] + [parts[0][i][..., -1:]],
                     -1)

  return join(0), join(1), join(2)


# Channel counts the fused render kernels are instantiated for.
FUSED_CHANNELS = tuple(range(1, 17)) + (20, 24, 28)
FUSED_MAX_LAYERS = 8


class _FusedSplatShaded(torch.autograd.Function):
  """splat_shaded_pixels on the fused render kernels (csrc/render.hip): one
  target pass interpolates the clip positions, maps them to the viewport,
  splats and composites; the backward runs the target pass again with the
  incoming d image (writing the per-pixel gradient scratch) and the source
  pass, which writes d shaded per pixel-layer and scatters the position
  gradient to the vertices (and the camera).

  geometry is world [V, 3] positions when `camera` is given (the kernels
  transform them on load, as dr_transform_homogeneous does), else clip-space
  [V, 4] vertices."""

  @staticmethod
  def forward(ctx, geometry, camera, shaded, ids, z, barys, triangles,
              want_buffers):
    lib = _native.lib()
    dev = shaded.device
    L, H, W, C = shaded.shape
    geometry = geometry.detach().to(torch.float32).contiguous()
    cam = None if camera is None else camera.detach().to(
        torch.float32).contiguous()
    shaded = shaded.detach().to(torch.float32).contiguous()
    image = torch.empty((H, W, C), dtype=torch.float32, device=dev)
    bufs = (None, None, None)
    if want_buffers:
      bufs = (torch.empty((3, H, W, C), dtype=torch.float32, device=dev),
              torch.empty((3, H, W, C), dtype=torch.float32, device=dev),
              torch.empty((3, H, W), dtype=torch.float32, device=dev))
    table = _FusedSplatShaded._job(geometry, cam, shaded, ids, z, barys,
                                   triangles, W, H)
    j = table[0]
    j.image = image.data_ptr()
    if want_buffers:
      j.accum, j.norm, j.accum_weights = (t.data_ptr() for t in bufs)
    tiles = _FusedSplatShaded._prepare(table, C)
    dev_table = _native.table_to_device(table, dev)
    mode = _native.RENDER_ACCUM_BUFFERS if want_buffers else 0
    _native.check(
        lib.dr_render_splat_forward(_native.ptr(dev_table), 1, tiles, L, C,
                                    mode, 1.0, None, None, 0,
                                    _native.stream_handle()),
        "dr_render_splat_forward")
    ctx.save_for_backward(geometry, cam, shaded, ids, z, barys, triangles)
    ctx.dims = (L, H, W, C)
    if want_buffers:
      ctx.mark_non_differentiable(bufs[0], bufs[1])
    return image, bufs[0], bufs[1]

  @staticmethod
  def _job(geometry, cam, shaded, ids, z, barys, triangles, W, H):
    table = (_native.SplatJob * 1)()
    j = table[0]
    if cam is not None:
      j.positions = geometry.data_ptr()
      j.camera = cam.data_ptr()
    else:
      j.clip_vertices = geometry.data_ptr()
    j.triangles = triangles.data_ptr()
    j.shaded = shaded.data_ptr()
    j.triangle_ids = ids.data_ptr()
    j.z_buffer = z.data_ptr()
    j.barycentrics = barys.data_ptr()
    j.num_vertices = geometry.shape[0]
    j.num_triangles = triangles.shape[0]
    j.width = W
    j.height = H
    return table

  @staticmethod
  def _prepare(table, C):
    tiles = ctypes.c_int64()
    _native.check(_native.lib().dr_splat_jobs_prepare(table, 1, C,
                                                      ctypes.byref(tiles)),
                  "dr_splat_jobs_prepare")
    return tiles.value

  @staticmethod
  def backward(ctx, d_image, _d_accum, _d_norm):
    geometry, cam, shaded, ids, z, barys, triangles = ctx.saved_tensors
    L, H, W, C = ctx.dims
    want_geom, want_cam, want_shaded = ctx.needs_input_grad[:3]
    if not (want_geom or want_cam or want_shaded):
      return (None,) * 8
    lib = _native.lib()
    dev = shaded.device
    d_image = d_image.to(torch.float32).contiguous()
    scratch = torch.empty(H * W * lib.dr_render_scratch_floats(C),
                          dtype=torch.float32, device=dev)
    records = (torch.zeros((geometry.shape[0], 4), dtype=torch.float32,
                           device=dev) if want_geom else None)
    d_cam = (torch.zeros((4, 4), dtype=torch.float32, device=dev)
             if want_cam else None)
    d_shaded = torch.empty_like(shaded) if want_shaded else None
    table = _FusedSplatShaded._job(geometry, cam, shaded, ids, z, barys,
                                   triangles, W, H)
    j = table[0]
    j.d_image = d_image.data_ptr()
    j.grad_scratch = scratch.data_ptr()
    j.d_records = None if records is None else records.data_ptr()
    j.d_camera = None if d_cam is None else d_cam.data_ptr()
    j.d_shaded = None if d_shaded is None else d_shaded.data_ptr()
    tiles = _FusedSplatShaded._prepare(table, C)
    dev_table = _native.table_to_device(table, dev)
    stream = _native.stream_handle()
    _native.check(
        lib.dr_render_splat_forward(_native.ptr(dev_table), 1, tiles, L, C,
                                    2, 1.0, None, None, 0, stream),
        "dr_render_splat_forward")
    _native.check(
        lib.dr_render_splat_backward(_native.ptr(dev_table), 1, tiles, L, C,
                                     _native.RENDER_SHADED, stream),
        "dr_render_splat_backward")
    d_geom = None
    if records is not None:
      # cols 0-2: d world position (camera given); else cols 0, 1, 3:
      # d clip x, y, w and a zero z column.
      d_geom = records[:, :3] if cam is not None else records
    return d_geom, d_cam, d_shaded, None, None, None, None, None


def _fused_splat_ok(vertices, camera_matrices, rasterized_fb, shaded):
  """Whether the fused kernels can take this splat_shaded_pixels call: one
  view, a z buffer from the rasterizer (it equals the interpolated clip z / w
  when the framebuffer was rasterized from these vertices and camera, as
  rasterize_triangles produces it), barycentrics that carry no gradient
  (differentiable barycentrics route gradients through the interpolated
  positions, which the unfused form handles), and layers and channels the
  kernels are built for."""
  if (rasterized_fb.z_buffer is None or rasterized_fb.triangles is None or
      rasterized_fb.is_batched):
    return False
  if rasterized_fb.barycentrics.requires_grad:
    return False
  if shaded.dim() != 4 or vertices.dim() != 2:
    return False
  L, _, _, c = shaded.shape
  if L > FUSED_MAX_LAYERS or c > FUSED_CHANNELS[-1]:
    return False
  if camera_matrices is not None:
    return camera_matrices.dim() == 2 and vertices.shape[-1] in (3, 4)
  return vertices.shape[-1] == 4


def _splat_shaded_fused(vertices, camera_matrices, rasterized_fb, shaded,
                        return_accum_buffers):
  _native.require_device(vertices, shaded)
  geometry, camera = vertices, camera_matrices
  if camera is not None and vertices.shape[-1] == 4:
    # Homogeneous world points: transform them (differentiably) first.
    geometry, camera = transforms.transform_homogeneous(camera, vertices), None
  c = shaded.shape[-1]
  k = next(n for n in FUSED_CHANNELS if n >= c)
  ids = rasterized_fb.triangle_id[..., 0].to(torch.int32).contiguous()
  z = rasterized_fb.z_buffer.detach().to(torch.float32).contiguous()
  b = rasterized_fb.barycentrics.detach().to(torch.float32).contiguous()
  tris = rasterized_fb.triangles.to(torch.int32).contiguous()
  out, accum, norm = _FusedSplatShaded.apply(
      geometry, camera, _with_alpha_last(shaded, k), ids, z, b, tris,
      bool(return_accum_buffers))
  if k != c:
    keep = lambda t: None if t is None else torch.cat(
        [t[..., :c - 1], t[..., -1:]], -1)
    out, accum, norm = keep(out), keep(accum), keep(norm)
  return out, accum, norm


def splat_shaded_pixels(vertices, camera_matrices, rasterized_fb,
                        shaded_pixels, return_accum_buffers=False,
                        fused=None):
  """splat.py:178-229: splat shaded pixels at their clip-space positions.

  A single view with a rasterizer z buffer and constant barycentrics (what
  rasterize_triangles returns) goes through the fused render kernels: one
  launch forward, two backward, no intermediate frames.  Other calls
  (batched framebuffers, differentiable barycentrics, more than 8 layers or
  28 channels) take the unfused composition: transform, interpolate the clip
  positions, viewport, splat_at_pixel_centers.  `fused` (not a reference
  argument): None picks as above, False forces the unfused composition,
  True raises ValueError where the fused kernels cannot take the call."""
  ok = fused is not False and _fused_splat_ok(
      vertices, camera_matrices, rasterized_fb, shaded_pixels)
  if fused and not ok:
    raise ValueError("splat_shaded_pixels(fused=True) needs one view, a z "
                     "buffer, constant barycentrics, at most 8 layers and "
                     "28 channels")
  if ok:
    out, accum, norm = _splat_shaded_fused(vertices, camera_matrices,
                                           rasterized_fb, shaded_pixels,
                                           return_accum_buffers)
  else:
    out, accum, norm = _splat_shaded_unfused(vertices, camera_matrices,
                                             rasterized_fb, shaded_pixels,
                                             return_accum_buffers)
  if return_accum_buffers:
    return out, accum, norm
  return out


def _splat_shaded_unfused(vertices, camera_matrices, rasterized_fb,
                          shaded_pixels, return_accum_buffers):
  H, W = shaded_pixels.shape[1:3]
  projected = vertices
  if camera_matrices is not None:
    projected = transforms.transform_homogeneous(camera_matrices, vertices)
  clip = interpolate.interpolate_vertex_attribute(
      projected, rasterized_fb, (0.0, 0.0, 1.0, 1.0))
  ndc = clip[..., :3] / clip[..., 3:4]
  scale = torch.tensor([W, H, 1], dtype=torch.float32, device=clip.device)
  viewport = (ndc + 1.0) * scale * 0.5
  return splat_at_pixel_centers(viewport, shaded_pixels,
                                rasterized_fb.foreground_mask,
                                return_accum_buffers)
