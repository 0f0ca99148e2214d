"""Barycentric vertex-attribute interpolation on the GPU.

Mirrors diffren.jax.internal.interpolate (interpolate.py:24-63, 118-170):
interpolate_vertex_attribute(attribute [V,C], framebuffer, background_value)
-> [L,H,W,C], with a hand-written backward kernel (scatter-add of
barycentric-weighted gradients to the three vertices of each pixel's
triangle).  Gradients flow to the attribute, and to the barycentrics when
they are differentiable (rasterize_triangles(compute_diff_barys=True),
rasterize.py:120-121); the rasterizer's own barycentrics and ids are
constants (rasterize_triangles_xla.py:51-52).
"""
import torch

from diffren_amd import _native


class _Interpolate(torch.autograd.Function):

  @staticmethod
  def forward(ctx, attribute, ids, barys, triangles, background):
    a = attribute.detach().to(torch.float32).contiguous()
    C = a.shape[1]
    n = ids.numel()
    out = torch.empty(tuple(ids.shape) + (C,), dtype=torch.float32,
                      device=a.device)
    _native.check(
        _native.lib().dr_interpolate_fwd(
            _native.ptr(ids), _native.ptr(barys), _native.ptr(triangles),
            triangles.shape[0], _native.ptr(a), a.shape[0], C,
            _native.ptr(background), n, _native.ptr(out),
            _native.stream_handle()), "dr_interpolate_fwd")
    ctx.save_for_backward(ids, barys, triangles, a)
    return out

  @staticmethod
  def backward(ctx, d_out):
    ids, barys, triangles, a = ctx.saved_tensors
    V, C = a.shape
    d_out = d_out.to(torch.float32).contiguous()
    da = db = None
    if ctx.needs_input_grad[0]:
      da = torch.zeros((V, C), dtype=torch.float32, device=d_out.device)
      _native.check(
          _native.lib().dr_interpolate_bwd(
              _native.ptr(ids), _native.ptr(barys), _native.ptr(triangles),
              triangles.shape[0], _native.ptr(d_out), V, C, ids.numel(),
              _native.ptr(da), _native.stream_handle()), "dr_interpolate_bwd")
    if ctx.needs_input_grad[2]:
      db = torch.empty(tuple(ids.shape) + (3,), dtype=torch.float32,
                       device=d_out.device)
      _native.check(
          _native.lib().dr_interpolate_bwd_barycentrics(
              _native.ptr(ids), _native.ptr(triangles), triangles.shape[0],
              _native.ptr(a), V, C, _native.ptr(d_out), ids.numel(),
              _native.ptr(db), _native.stream_handle()),
          "dr_interpolate_bwd_barycentrics")
    return da, None, db, None, None


def _background(background_value, C, device):
  if background_value is None:
    return None
  bg = torch.as_tensor(background_value, dtype=torch.float32, device=device)
  bg = bg.reshape(-1)
  if bg.numel() == 1 and C > 1:
    bg = bg.expand(C)
  if bg.numel() != C:
    raise ValueError(f"background_value has {bg.numel()} entries for {C} "
                     "channels")
  return bg.contiguous()


def interpolate_vertex_attribute(attribute, framebuffer,
                                 background_value=None):
  """[V, C] attribute -> [L, H, W, C] (interpolate.py:118-145)."""
  if attribute.dim() != 2:
    raise ValueError("attribute must be [num_vertices, num_channels]")
  _native.require_device(attribute)
  ids = framebuffer.triangle_id[..., 0].contiguous()
  barys = framebuffer.barycentrics.contiguous()
  tris = framebuffer.triangles
  bg = _background(background_value, attribute.shape[1], attribute.device)
  return _Interpolate.apply(attribute, ids, barys, tris, bg)


def interpolate_vertex_attributes(attributes, framebuffer,
                                  background_value=None):
  """Dict version (interpolate.py:148-170)."""
  return {k: interpolate_vertex_attribute(v, framebuffer, background_value)
          for k, v in attributes.items()}
