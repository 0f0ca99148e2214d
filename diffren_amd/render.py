"""render_triangles: rasterize -> interpolate -> shade -> splat/over.

Mirrors diffren.jax.render.render_triangles (diffren/jax/render.py:28-140),
same arguments and return values, on the gfx950 kernels.  The shading
function receives a dict of [L, H, W, C] attribute buffers (the reference
vmaps it over layers; torch shaders broadcast over the leading dim).
"""
from diffren_amd import composite
from diffren_amd import constants
from diffren_amd import interpolate
from diffren_amd import rasterize
from diffren_amd import splat


def render_triangles(vertices, attributes, triangles, camera_matrices,
                     image_width, image_height, shading_function,
                     num_layers=1,
                     face_culling_mode=constants.FaceCullingMode.BACK,
                     compositing_mode=constants.CompositingMode.SPLAT_OVER,
                     return_accum_buffers=False, compute_diff_barys=False):
  """render.py:28-140."""
  if (compositing_mode == constants.CompositingMode.SPLAT_OVER and
      face_culling_mode != constants.FaceCullingMode.BACK):
    raise ValueError("Backface culling must be enabled when using splatting.")
  rasterized = rasterize.rasterize_triangles(
      vertices, triangles, camera_matrices, image_width, image_height,
      num_layers=num_layers, face_culling_mode=face_culling_mode,
      compute_diff_barys=compute_diff_barys)
  interpolated = interpolate.interpolate_vertex_attributes(
      attributes, rasterized, (0.0,))
  shaded = shading_function(interpolated)
  if compositing_mode == constants.CompositingMode.SPLAT_OVER:
    return splat.splat_shaded_pixels(vertices, camera_matrices, rasterized,
                                     shaded, return_accum_buffers)
  shaded = shaded * rasterized.foreground_mask
  if compositing_mode == constants.CompositingMode.OVER:
    return composite.over(shaded)
  return shaded
