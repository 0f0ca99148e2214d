"""Differentiable barycentrics (rasterize_triangles(compute_diff_barys=True))
on the GPU vs the CPU oracle (oracle/diffbary_oracle.py).

Tolerances: forward barycentrics bit-exact to the float32 oracle (same
expressions, same order); the golden Jacobian of rasterize_test.py:216-224
to 6 decimals; barycentric and colour gradients within 1e-4 relative of the
float64 oracle (evaluated at the kernel's float32 forward values); vertex
gradients of interpolated or splatted images within 1e-3 (DIFF_GRAD_TOL).
"""
import numpy as np
import pytest
import torch

import helpers
import scenes as golden_scenes
from diffren_amd import constants
from diffren_amd import interpolate
from diffren_amd import rasterize
from diffren_amd import render
from diffren_amd import scenes
from oracle import diffbary_oracle as dbo
from oracle import raster_oracle
from oracle import splat_oracle as so
from test_oracle_diffbary import GOLDEN_JACOBIAN, GOLDEN_NDC, TRI

pytestmark = pytest.mark.gpu

GRAD_TOL = 1e-4
# Vertex gradients that pass through interpolation into the coordinates: the
# VJP of coordinates normalised to sum 1 subtracts the common part of the
# three vertex terms, so on small triangles the float32 inputs (d image, the
# interpolated attribute's vertex values) lose digits, as in any float32
# autodiff of the same function; measured 2e-5 (over) to 4e-4 (splat).
DIFF_GRAD_TOL = 1e-3


def _diff_fb(clip, tris, w, h, layers, device, cull=0):
  c = torch.as_tensor(clip, device=device)
  t = torch.as_tensor(tris, device=device)
  return rasterize.rasterize_triangles(c, t, None, w, h, layers, cull,
                                       compute_diff_barys=True)


@pytest.mark.parametrize("case", ["cube", "sphere", "random0", "random1"])
def test_forward_bit_exact_vs_oracle(case, device):
  if case == "cube":
    clip, tris, w, h, L = (golden_scenes.CUBE_VERTICES,
                           golden_scenes.CUBE_TRIANGLES, 640, 480, 2)
  elif case == "sphere":
    s = scenes.icosphere_scene(seed=3, subdivisions=4)
    clip, tris, w, h, L = (scenes.to_clip(s["vertices"], s["camera"]),
                           s["triangles"], 256, 192, 3)
  else:
    seed = int(case[-1])
    clip, tris = golden_scenes.adversarial_scene(2000 + seed)
    w, h, L = 77, 45, 3
  fb = _diff_fb(clip, tris, w, h, L, device)
  ids = fb.triangle_id[..., 0].cpu().numpy()
  want = dbo.barycentrics_f32(clip, tris, ids, w, h)
  assert helpers.same_bits(fb.barycentrics.detach().cpu().numpy(), want)


def test_golden_jacobian(device):
  tris = torch.as_tensor(TRI, device=device)
  eye = torch.eye(4, device=device)

  def f(ndc):
    fb = rasterize.rasterize_triangles(ndc, tris, eye, 64, 48,
                                       compute_diff_barys=True)
    return fb.barycentrics[0, 24, 32]

  jac = torch.autograd.functional.jacobian(
      f, torch.as_tensor(GOLDEN_NDC, device=device))
  np.testing.assert_array_almost_equal(jac.cpu().numpy(), GOLDEN_JACOBIAN)
  # Without compute_diff_barys the barycentrics are constants
  # (rasterize_test.py:208-210).  The default culling is NONE, as in the
  # reference (rasterize.py:32): this triangle is back-facing.
  ndc = torch.as_tensor(GOLDEN_NDC, device=device).requires_grad_(True)
  fb = rasterize.rasterize_triangles(ndc, tris, eye, 64, 48)
  assert not fb.barycentrics.requires_grad


@pytest.mark.parametrize("w,fixture", [
    ((1.0, 1.0, 1.0), "Simple_Triangle.png"),
    ((0.2, 0.5, 2.0), "Perspective_Corrected_Triangle.png"),
])
def test_reference_pngs(w, fixture, device):
  v = golden_scenes.SIMPLE_TRIANGLE * np.asarray(w, np.float32)[:, None]
  fb = _diff_fb(v, TRI, 640, 480, 1, device)
  ok, msg = helpers.check_image(fb.barycentrics[0].detach().cpu().numpy(),
                                fixture)
  assert ok, msg


def test_reference_cube_png(device):
  fb = _diff_fb(golden_scenes.CUBE_VERTICES, golden_scenes.CUBE_TRIANGLES,
                640, 480, 1, device)
  ok, msg = helpers.check_image(fb.barycentrics[0].detach().cpu().numpy(),
                                "Barycentrics_Cube.png",
                                add_transparency=False)
  assert ok, msg


def _b64(s, v64, ids, w, h):
  """Float64 gradient oracle of the coordinates at the kernel's float32
  forward values (bit-identical to barycentrics_f32)."""
  clip32 = so.transform(s["camera"], s["vertices"], torch.float32).numpy()
  b32 = dbo.barycentrics_f32(clip32, s["triangles"], ids, w, h)
  clip64 = so.transform(s["camera"], v64, torch.float64)
  return dbo.barycentrics_torch_at(clip64, s["triangles"], ids, w, h, b32)


def _scene64(seed, w, h, layers):
  s = scenes.icosphere_scene(seed=seed, subdivisions=3)
  clip32 = so.transform(s["camera"], s["vertices"], torch.float32).numpy()
  ids, _, _ = raster_oracle.rasterize(clip32, s["triangles"], w, h, layers, 1)
  return s, ids


@pytest.mark.parametrize("layers", [1, 3])
def test_barycentric_gradient_matches_float64(layers, device):
  w, h = 96, 80
  s, ids = _scene64(5, w, h, layers)
  rng = np.random.default_rng(layers)
  wgt = rng.normal(size=(layers, h, w, 3)).astype(np.float32)
  v = torch.as_tensor(s["vertices"], device=device).requires_grad_(True)
  cam = torch.as_tensor(s["camera"], device=device)
  tris = torch.as_tensor(s["triangles"], device=device)
  fb = rasterize.rasterize_triangles(v, tris, cam, w, h, layers,
                                     constants.FaceCullingMode.BACK,
                                     compute_diff_barys=True)
  assert (fb.triangle_id[..., 0].cpu().numpy() == ids).all()
  (fb.barycentrics * torch.as_tensor(wgt, device=device)).sum().backward()
  v64 = torch.as_tensor(s["vertices"], dtype=torch.float64).requires_grad_(
      True)
  clip64 = so.transform(s["camera"], v64, torch.float64)
  b64 = dbo.barycentrics_torch(clip64, s["triangles"], ids, w, h)
  (b64 * torch.as_tensor(wgt, dtype=torch.float64)).sum().backward()
  err = helpers.rel_err(v.grad.cpu(), v64.grad)
  assert err <= GRAD_TOL


def test_interpolation_gradient_through_barycentrics(device):
  w, h, layers = 96, 80, 2
  s, ids = _scene64(6, w, h, layers)
  v = torch.as_tensor(s["vertices"], device=device).requires_grad_(True)
  col = torch.as_tensor(s["colors"], device=device).requires_grad_(True)
  cam = torch.as_tensor(s["camera"], device=device)
  tris = torch.as_tensor(s["triangles"], device=device)
  fb = rasterize.rasterize_triangles(v, tris, cam, w, h, layers,
                                     constants.FaceCullingMode.BACK,
                                     compute_diff_barys=True)
  out = interpolate.interpolate_vertex_attribute(col, fb)
  target = torch.rand_like(out)
  ((out - target) ** 2).sum().backward()
  v64 = torch.as_tensor(s["vertices"], dtype=torch.float64).requires_grad_(
      True)
  c64 = torch.as_tensor(s["colors"], dtype=torch.float64).requires_grad_(True)
  b64 = _b64(s, v64, ids, w, h)
  o64 = so.interpolate(c64, ids, b64, s["triangles"])
  err = helpers.rel_err(out.detach().cpu(), o64.detach())
  assert err <= 1e-5
  ((o64 - target.cpu().double()) ** 2).sum().backward()
  err_c = helpers.rel_err(col.grad.cpu(), c64.grad)
  err_v = helpers.rel_err(v.grad.cpu(), v64.grad)
  assert err_c <= GRAD_TOL and err_v <= DIFF_GRAD_TOL, (err_c, err_v)


@pytest.mark.parametrize("mode", ["splat", "over"])
def test_render_with_diff_barys(mode, device):
  w = h = 96
  layers = 2
  s, ids = _scene64(7, w, h, layers)
  v = torch.as_tensor(s["vertices"], device=device).requires_grad_(True)
  col = torch.as_tensor(s["colors"], device=device).requires_grad_(True)
  cam = torch.as_tensor(s["camera"], device=device)
  tris = torch.as_tensor(s["triangles"], device=device)
  cm = (constants.CompositingMode.SPLAT_OVER if mode == "splat" else
        constants.CompositingMode.OVER)
  img = render.render_triangles(v, {"c": col}, tris, cam, w, h,
                                lambda a: a["c"], num_layers=layers,
                                compositing_mode=cm, compute_diff_barys=True)
  target = torch.rand_like(img)
  ((img - target) ** 2).mean().backward()
  v64 = torch.as_tensor(s["vertices"], dtype=torch.float64).requires_grad_(
      True)
  c64 = torch.as_tensor(s["colors"], dtype=torch.float64).requires_grad_(True)
  clip64 = so.transform(s["camera"], v64, torch.float64)
  b64 = _b64(s, v64, ids, w, h)
  shaded = so.interpolate(c64, ids, b64, s["triangles"], (0.0,))
  if mode == "splat":
    o64, _, _ = so.splat_shaded_pixels(clip64, ids, b64, s["triangles"],
                                       shaded)
  else:
    mask = torch.as_tensor(ids >= 0, dtype=torch.float64)[..., None]
    o64 = so.over(shaded * mask)
  err = helpers.rel_err(img.detach().cpu(), o64.detach())
  assert err <= 1e-5
  ((o64 - target.cpu().double()) ** 2).mean().backward()
  err_c = helpers.rel_err(col.grad.cpu(), c64.grad)
  err_v = helpers.rel_err(v.grad.cpu(), v64.grad)
  assert err_c <= GRAD_TOL and err_v <= DIFF_GRAD_TOL, (err_c, err_v)


def test_background_and_degenerate(device):
  # Background pixels are 0; a degenerate triangle is never rasterized, so
  # no coordinate is computed for it.
  fb = _diff_fb(golden_scenes.DEGENERATE, TRI, 4, 4, 1, device)
  assert (fb.barycentrics.detach() == 0).all()
  # An empty frame (all background) has zero gradient.
  clip = torch.as_tensor(golden_scenes.SIMPLE_TRIANGLE,
                         device=device).requires_grad_(True)
  t = torch.as_tensor(TRI, device=device)
  fb = rasterize.rasterize_triangles(clip + torch.tensor(
      [5.0, 0.0, 0.0, 0.0], device=device), t, None, 8, 8, 1, 0,
                                     compute_diff_barys=True)
  fb.barycentrics.sum().backward()
  assert (clip.grad == 0).all()


def test_exact_barycentric_gradient_for_user_coordinates(device):
  # Barycentrics that do not come from differentiable_barycentrics get the
  # exact (uncentred) interpolation VJP.
  w, h, layers = 64, 48, 2
  s, ids = _scene64(8, w, h, layers)
  tris = torch.as_tensor(s["triangles"], device=device)
  rng = np.random.default_rng(3)
  b = torch.as_tensor(rng.uniform(size=(layers, h, w, 3)).astype(np.float32),
                      device=device).requires_grad_(True)
  fbuf = rasterize.fb.Framebuffer(
      barycentrics=b, triangle_id=torch.as_tensor(ids, device=device)[
          ..., None], triangles=tris)
  col = torch.as_tensor(s["colors"], device=device)
  out = interpolate.interpolate_vertex_attribute(col, fbuf)
  gout = torch.as_tensor(rng.normal(size=tuple(out.shape)).astype(np.float32),
                         device=device)
  (out * gout).sum().backward()
  b64 = torch.as_tensor(b.detach().cpu().numpy(),
                        dtype=torch.float64).requires_grad_(True)
  o64 = so.interpolate(torch.as_tensor(s["colors"], dtype=torch.float64),
                       ids, b64, s["triangles"])
  (o64 * gout.cpu().double()).sum().backward()
  err = helpers.rel_err(b.grad.cpu(), b64.grad)
  assert err <= 1e-5
