"""Rasterize-then-splat on the GPU (diffren/jax/internal/splat.py:26-229).

splat_at_pixel_centers(xyz_layers, color_layers, foreground_mask)
  -> (output [H,W,C], accum [3,H,W,C], normalized [3,H,W,C])
splat_shaded_pixels(vertices, camera_matrices, rasterized_fb, shaded_pixels,
                    return_accum_buffers)
Both keep the reference's signatures.  `output` has a hand-written backward
(dr_splat_bwd) with respect to the xy positions and the colours; the
accumulation buffers are returned for inspection without gradients.
"""
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


def splat_shaded_pixels(vertices, camera_matrices, rasterized_fb,
                        shaded_pixels, return_accum_buffers=False):
  """splat.py:178-229: splat shaded pixels at their clip-space positions."""
  H, W = shaded_pixels.shape[1:3]
  projected = vertices
  if camera_matrices is not None:
    projected = transforms.transform_homogeneous(camera_matrices, vertices)
  clip = interpolate.interpolate_vertex_attribute(
      projected, rasterized_fb, (0.0, 0.0, 1.0, 1.0))
  ndc = clip[..., :3] / clip[..., 3:4]
  scale = torch.tensor([W, H, 1], dtype=torch.float32, device=clip.device)
  viewport = (ndc + 1.0) * scale * 0.5
  out, accum, norm = splat_at_pixel_centers(viewport, shaded_pixels,
                                            rasterized_fb.foreground_mask,
                                            return_accum_buffers)
  if return_accum_buffers:
    return out, accum, norm
  return out
