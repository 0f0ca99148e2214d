"""splat_shaded_pixels on the fused render kernels (shaded pixels of any
shading function as the splat's colours) against the unfused composition and
the float64 oracle (oracle/splat_oracle.py, autograd).

Tolerances: images bit-identical to the unfused GPU path and within 1e-5
relative of the oracle; gradients (vertices through the splat positions, the
shader's inputs through the shaded pixels, the camera) within 1e-4 relative
of float64.
"""
import numpy as np
import pytest
import torch

import helpers
from diffren_amd import mesh
from diffren_amd import rasterize
from diffren_amd import render
from diffren_amd import scenes
from diffren_amd import shading
from diffren_amd import splat
from oracle import splat_oracle as so

pytestmark = pytest.mark.gpu

FWD_TOL = 1e-5
GRAD_TOL = 1e-4
LIGHT = (1.5, 2.0, 4.0)


def _point_light_shader(attrs):
  """RGBA: one point light on the interpolated colour, normal and position,
  alpha from the interpolated alpha attribute (shaders.point_light,
  shaders.py:100-173)."""
  rgb = shading.point_light(attrs["color"][..., :3],
                            shading.l2_normalize(attrs["normal"]),
                            attrs["position"], LIGHT, ks=0.5, ns=8.0)
  return torch.cat([rgb, attrs["color"][..., 3:]], -1)


def _attributes(v, col, tris):
  return {"position": v, "normal": mesh.compute_vertex_normals(v, tris),
          "color": col}


def _render(s, w, h, layers, device, fused, cam_grad=True):
  v = torch.as_tensor(s["vertices"], device=device).requires_grad_(True)
  rgb = torch.as_tensor(s["colors"][:, :3], device=device)
  col = torch.cat([rgb, torch.ones_like(rgb[:, :1])], 1).requires_grad_(True)
  cam = torch.as_tensor(s["camera"], device=device).requires_grad_(cam_grad)
  tris = torch.as_tensor(s["triangles"], device=device)
  img = render.render_triangles(v, _attributes(v, col, tris), tris, cam, w, h,
                                _point_light_shader, num_layers=layers,
                                fused_splat=fused)
  return img, v, col, cam


@pytest.mark.parametrize("layers", [1, 3])
def test_point_light_render_fused_vs_unfused_and_oracle(layers, device):
  """render_triangles with a point-light shader: the fused splat gives the
  unfused path's image (to 1e-6: the shader's vertex normals are summed by
  index_add_'s float atomics, so two renders' shaded pixels can differ in
  the last bit; with equal shaded pixels the images are bit-identical, see
  test_splat_shaded_pixels_channels), and vertex, colour and camera
  gradients (through the splat positions, the shader, the normals and the
  interpolation) within 1e-4 of the float64 oracle."""
  w, h = 160, 128
  s = scenes.icosphere_scene(seed=9)
  target = torch.rand((h, w, 4), device=device,
                      generator=torch.Generator(device).manual_seed(3))
  out = {}
  for fused in (True, False):
    img, v, col, cam = _render(s, w, h, layers, device, fused)
    ((img - target) ** 2).sum().backward()
    out[fused] = (img.detach(), v.grad, col.grad, cam.grad)
  assert helpers.rel_err(out[True][0].cpu(), out[False][0].cpu()) <= 1e-6

  # Oracle: the GPU rasterization (bit-exact to the oracle rasterizer, tested
  # elsewhere) with everything after it in float64 autograd.
  fb = rasterize.rasterize_triangles(
      torch.as_tensor(s["vertices"], device=device),
      torch.as_tensor(s["triangles"], device=device),
      torch.as_tensor(s["camera"], device=device), w, h, num_layers=layers,
      face_culling_mode=1)
  ids = fb.triangle_id[..., 0].cpu().numpy()
  b = fb.barycentrics.cpu().numpy()
  v64 = torch.as_tensor(s["vertices"], dtype=torch.float64).requires_grad_()
  c64 = torch.cat([torch.as_tensor(s["colors"][:, :3], dtype=torch.float64),
                   torch.ones((v64.shape[0], 1), dtype=torch.float64)],
                  1).requires_grad_()
  m64 = torch.as_tensor(s["camera"], dtype=torch.float64).requires_grad_()
  t64 = torch.as_tensor(s["triangles"]).long()
  attrs = {k: so.interpolate(a, ids, b, s["triangles"])
           for k, a in _attributes(v64, c64, t64).items()}
  shaded = _point_light_shader(attrs)
  clip = so.transform(m64, v64)
  o64, _, _ = so.splat_shaded_pixels(clip, ids, b, s["triangles"], shaded)
  ((o64 - target.cpu().double()) ** 2).sum().backward()
  img, dv, dc, dm = out[True]
  assert helpers.rel_err(img.cpu(), o64.detach()) <= FWD_TOL
  assert helpers.rel_err(dv.cpu(), v64.grad) <= GRAD_TOL
  assert helpers.rel_err(dc.cpu(), c64.grad) <= GRAD_TOL
  assert helpers.rel_err(dm.cpu(), m64.grad) <= GRAD_TOL
  # The unfused path's float32 autodiff evaluates the tap derivatives at the
  # rounded splat offsets; at these frame sizes it agrees to 1e-4 as well.
  assert helpers.rel_err(out[False][1].cpu(), v64.grad) <= GRAD_TOL


@pytest.mark.parametrize("channels", [1, 3, 4, 7, 12, 17, 30])
@pytest.mark.parametrize("clip_space", [False, True])
def test_splat_shaded_pixels_channels(channels, clip_space, device):
  """Random shaded pixels of C channels (17 runs padded to 20, 30 takes the
  unfused composition), from world vertices + camera or from clip-space
  vertices: image, d shaded and d vertices against the float64 oracle, and
  the image against the unfused GPU path."""
  w, h, L = 96, 80, 3
  s = scenes.icosphere_scene(seed=channels)
  cam = torch.as_tensor(s["camera"], device=device)
  tris = torch.as_tensor(s["triangles"], device=device)
  verts = torch.as_tensor(s["vertices"], device=device)
  fb = rasterize.rasterize_triangles(verts, tris, cam, w, h, num_layers=L,
                                     face_culling_mode=1)
  rng = np.random.default_rng(channels)
  sh = rng.uniform(0, 1, (L, h, w, channels)).astype(np.float32)
  d_out = torch.as_tensor(rng.normal(size=(h, w, channels)).astype(np.float32),
                          device=device)
  clip32 = so.transform(s["camera"], s["vertices"], torch.float32)
  geom_np = clip32.numpy() if clip_space else s["vertices"]
  results = {}
  for fused in (None, False):
    g = torch.as_tensor(geom_np, device=device).requires_grad_(True)
    shaded = torch.as_tensor(sh, device=device).requires_grad_(True)
    out = splat.splat_shaded_pixels(g, None if clip_space else cam, fb,
                                    shaded, fused=fused)
    (out * d_out).sum().backward()
    results[fused] = (out.detach(), g.grad, shaded.grad)
  assert torch.equal(results[None][0], results[False][0])
  ids = fb.triangle_id[..., 0].cpu().numpy()
  b = fb.barycentrics.cpu().numpy()
  g64 = torch.as_tensor(geom_np, dtype=torch.float64).requires_grad_()
  s64 = torch.as_tensor(sh, dtype=torch.float64).requires_grad_()
  clip64 = g64 if clip_space else so.transform(s["camera"], g64)
  o64, _, _ = so.splat_shaded_pixels(clip64, ids, b, s["triangles"], s64)
  (o64 * d_out.cpu().double()).sum().backward()
  img, dg, ds = results[None]
  assert helpers.rel_err(img.cpu(), o64.detach()) <= FWD_TOL
  assert helpers.rel_err(ds.cpu(), s64.grad) <= GRAD_TOL
  assert helpers.rel_err(dg.cpu(), g64.grad) <= GRAD_TOL
  # d shaded is exactly 0 on background pixel-layers.
  bg = fb.triangle_id[..., 0] < 0
  assert not ds[bg].any()


def test_fused_splat_accum_buffers_and_refusals(device):
  """return_accum_buffers through the fused kernels equals the unfused
  path's buffers; fused=True refuses what the fused kernels cannot take."""
  w, h, L = 64, 48, 2
  s = scenes.icosphere_scene(seed=1)
  cam = torch.as_tensor(s["camera"], device=device)
  tris = torch.as_tensor(s["triangles"], device=device)
  verts = torch.as_tensor(s["vertices"], device=device)
  fb = rasterize.rasterize_triangles(verts, tris, cam, w, h, num_layers=L,
                                     face_culling_mode=1)
  sh = torch.rand((L, h, w, 4), device=device)
  a = splat.splat_shaded_pixels(verts, cam, fb, sh, True, fused=True)
  b = splat.splat_shaded_pixels(verts, cam, fb, sh, True, fused=False)
  for x, y in zip(a, b):
    assert helpers.rel_err(x.cpu(), y.cpu()) <= FWD_TOL
  fb_diff = rasterize.rasterize_triangles(verts, tris, cam, w, h,
                                          num_layers=L, face_culling_mode=1,
                                          compute_diff_barys=True)
  fb_diff.barycentrics.requires_grad_(True)
  with pytest.raises(ValueError):
    splat.splat_shaded_pixels(verts, cam, fb_diff, sh, fused=True)
  with pytest.raises(ValueError):
    splat.splat_shaded_pixels(verts, cam, fb,
                              torch.rand((L, h, w, 29), device=device),
                              fused=True)
