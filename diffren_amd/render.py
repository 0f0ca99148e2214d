"""render_triangles: rasterize -> interpolate -> shade -> splat/over.

Mirrors diffren.jax.render.render_triangles (diffren/jax/render.py:28-140),
same arguments and return values, on the gfx950 kernels.  The shading
function receives a dict of [L, H, W, C] attribute buffers (the reference
vmaps it over layers; torch shaders broadcast over the leading dim).

With SPLAT_OVER the shaded pixels of any shading function are splatted by
the fused render kernels (splat.splat_shaded_pixels).

recompute_rasterization=True (an addition to the reference's arguments)
keeps no framebuffer between forward and backward: the forward runs without
saving intermediates and the backward rasterizes again (the rasterization is
a constant of the gradient, rasterize_triangles_xla.py:51-52, so recomputing
it gives the same ids and barycentrics bit for bit), interpolates, shades and
splats with gradients, then back-propagates.  It trades one more forward for
the L x H x W framebuffer, interpolated attributes and splat inputs that the
backward would otherwise hold.
"""
import torch

from diffren_amd import composite
from diffren_amd import constants
from diffren_amd import interpolate
from diffren_amd import rasterize
from diffren_amd import splat


def _render(vertices, attributes, triangles, camera_matrices, image_width,
            image_height, shading_function, num_layers, face_culling_mode,
            compositing_mode, return_accum_buffers, compute_diff_barys,
            fused_splat):
  rasterized = rasterize.rasterize_triangles(
      vertices, triangles, camera_matrices, image_width, image_height,
      num_layers=num_layers, face_culling_mode=face_culling_mode,
      compute_diff_barys=compute_diff_barys)
  interpolated = interpolate.interpolate_vertex_attributes(
      attributes, rasterized, (0.0,))
  shaded = shading_function(interpolated)
  if compositing_mode == constants.CompositingMode.SPLAT_OVER:
    return splat.splat_shaded_pixels(vertices, camera_matrices, rasterized,
                                     shaded, return_accum_buffers,
                                     fused=fused_splat)
  shaded = shaded * rasterized.foreground_mask
  if compositing_mode == constants.CompositingMode.OVER:
    return composite.over(shaded)
  return shaded


def render_triangles(vertices, attributes, triangles, camera_matrices,
                     image_width, image_height, shading_function,
                     num_layers=1,
                     face_culling_mode=constants.FaceCullingMode.BACK,
                     compositing_mode=constants.CompositingMode.SPLAT_OVER,
                     return_accum_buffers=False, compute_diff_barys=False,
                     recompute_rasterization=False, fused_splat=None):
  """render.py:28-140.  fused_splat is splat_shaded_pixels' `fused`."""
  if (compositing_mode == constants.CompositingMode.SPLAT_OVER and
      face_culling_mode != constants.FaceCullingMode.BACK):
    raise ValueError("Backface culling must be enabled when using splatting.")
  args = (vertices, attributes, triangles, camera_matrices, image_width,
          image_height, shading_function, num_layers, face_culling_mode,
          compositing_mode, return_accum_buffers, compute_diff_barys,
          fused_splat)
  if recompute_rasterization and torch.is_grad_enabled():
    return torch.utils.checkpoint.checkpoint(_render, *args,
                                             use_reentrant=False)
  return _render(*args)
