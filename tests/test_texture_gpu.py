"""Bilinear texture resampling (image.bilinear_resample, shading.texture_map)
on the GPU vs the CPU oracle (oracle/texture_oracle.py).

Tolerances: forward bit-exact to the float32 oracle (same expressions, same
order); gradients within 1e-5 relative of the float64 oracle.
"""
import os

import numpy as np
import pytest
import torch

import helpers
from diffren_amd import image
from diffren_amd import render
from diffren_amd import shading
from oracle import texture_oracle as to
from oracle import raster_oracle
from oracle import splat_oracle as so
from test_oracle_texture import SQUARE, SQUARE_TEXTURE, SQUARE_TRIS, SQUARE_UV
from test_oracle_texture import boundary_inputs
from test_oracle_texture import SPOT_CASES, spot_oracle, spot_scene, spot_shade

pytestmark = pytest.mark.gpu


def _locations(rng, n, H, W):
  loc = rng.uniform(-2.0, [W + 2.0, H + 2.0], size=(n, 2)).astype(np.float32)
  # Exact texel centres, texel corners and integers: the clip's ties.
  k = n // 4
  loc[:k] = np.floor(loc[:k]) + 0.5
  loc[k:2 * k] = np.floor(loc[k:2 * k])
  return loc


@pytest.mark.parametrize("pad", ["constant", "edge"])
@pytest.mark.parametrize("C", [1, 3, 4])
def test_forward_bit_exact(pad, C, device):
  rng = np.random.default_rng(C)
  H, W = 37, 53
  img = rng.uniform(size=(H, W, C)).astype(np.float32)
  loc = _locations(rng, 4000, H, W).reshape(40, 100, 2)
  got = image.bilinear_resample(torch.as_tensor(img, device=device),
                                torch.as_tensor(loc, device=device),
                                pad_mode=pad)
  want = to.bilinear_resample(torch.as_tensor(img), torch.as_tensor(loc), pad)
  assert tuple(got.shape) == (40, 100, C)
  assert helpers.same_bits(got.cpu().numpy(), want.numpy())


@pytest.mark.parametrize("pad", ["constant", "edge"])
def test_backward_matches_float64(pad, device):
  rng = np.random.default_rng(5)
  H, W, C = 19, 23, 3
  img = rng.uniform(size=(H, W, C)).astype(np.float32)
  loc = _locations(rng, 3000, H, W).reshape(30, 100, 2)
  wgt = rng.normal(size=(30, 100, C)).astype(np.float32)
  gi = torch.as_tensor(img, device=device).requires_grad_(True)
  gl = torch.as_tensor(loc, device=device).requires_grad_(True)
  out = image.bilinear_resample(gi, gl, pad_mode=pad)
  (out * torch.as_tensor(wgt, device=device)).sum().backward()
  oi = torch.as_tensor(img, dtype=torch.float64).requires_grad_(True)
  ol = torch.as_tensor(loc, dtype=torch.float64).requires_grad_(True)
  (to.bilinear_resample(oi, ol, pad) *
   torch.as_tensor(wgt, dtype=torch.float64)).sum().backward()
  err_i = helpers.rel_err(gi.grad.cpu(), oi.grad)
  err_l = helpers.rel_err(gl.grad.cpu(), ol.grad)
  assert err_i <= 1e-5 and err_l <= 1e-5, (err_i, err_l)


def test_chw_orders(device):
  rng = np.random.default_rng(2)
  img = torch.as_tensor(rng.uniform(size=(9, 11, 4)).astype(np.float32),
                        device=device)
  loc = torch.as_tensor(_locations(rng, 60, 9, 11).reshape(6, 10, 2),
                        device=device)
  hwc = image.bilinear_resample(img, loc)
  chw = image.bilinear_resample(img.permute(2, 0, 1), loc.permute(2, 0, 1),
                                image_order="CHW", locations_order="CHW")
  assert torch.equal(chw.permute(1, 2, 0), hwc)
  with pytest.raises(ValueError):
    image.bilinear_resample(img, loc, pad_mode="reflect")
  with pytest.raises(ValueError):
    image.bilinear_resample(img, loc[..., :1])


@pytest.mark.parametrize("order", ["HWC", "CHW"])
def test_recovers_input(order, device):
  # image_test.py:25-44.
  n = 64
  img = np.arange(3 * n * n, dtype=np.float32).reshape(n, n, 3)
  x, y = np.meshgrid(np.arange(n), np.arange(n))
  loc = (np.stack((x, y), -1) + 0.5).astype(np.float32)
  if order == "CHW":
    img_in, loc_in = img.transpose(2, 0, 1), loc.transpose(2, 0, 1)
  else:
    img_in, loc_in = img, loc
  out = image.bilinear_resample(
      torch.as_tensor(np.ascontiguousarray(img_in), device=device),
      torch.as_tensor(np.ascontiguousarray(loc_in), device=device),
      image_order=order, locations_order=order).cpu().numpy()
  assert np.array_equal(out, img_in)


@pytest.mark.parametrize("order", ["HWC", "CHW"])
def test_boundary_cases(order, device):
  # image_test.py:46-108, infinities included; bit-exact to the oracle.
  img, loc, want = boundary_inputs()
  if order == "CHW":
    img_in, loc_in = img.transpose(2, 0, 1), loc.transpose(2, 0, 1)
    want = want.transpose(2, 0, 1)
  else:
    img_in, loc_in = img, loc
  out = image.bilinear_resample(
      torch.as_tensor(np.ascontiguousarray(img_in), device=device),
      torch.as_tensor(np.ascontiguousarray(loc_in), device=device),
      image_order=order, locations_order=order).cpu().numpy()
  np.testing.assert_array_almost_equal(want, out, decimal=5)
  ref = to.bilinear_resample(torch.as_tensor(img), torch.as_tensor(loc))
  ref = ref.numpy()
  if order == "CHW":
    ref = ref.transpose(2, 0, 1)
  assert helpers.same_bits(out, np.ascontiguousarray(ref))


def test_textured_square_matches_reference_png(device):
  # shaders_test.py:33-71 through the GPU render path (SPLAT_OVER).
  tex = torch.as_tensor(SQUARE_TEXTURE, device=device)
  img = render.render_triangles(
      torch.as_tensor(SQUARE, device=device),
      {"uv": torch.as_tensor(SQUARE_UV, device=device)},
      torch.as_tensor(SQUARE_TRIS, device=device), None, 640, 480,
      lambda a: shading.texture_map(a["uv"], tex))
  ok, msg = helpers.check_image(img.cpu().numpy(), "Textured_Square.png")
  assert ok, msg


def test_texture_gradient_through_render(device):
  rng = np.random.default_rng(9)
  tex = rng.uniform(size=(8, 8, 4)).astype(np.float32)
  w, h = 96, 72
  t = torch.as_tensor(tex, device=device).requires_grad_(True)
  img = render.render_triangles(
      torch.as_tensor(SQUARE, device=device),
      {"uv": torch.as_tensor(SQUARE_UV, device=device)},
      torch.as_tensor(SQUARE_TRIS, device=device), None, w, h,
      lambda a: shading.texture_map(a["uv"], t))
  target = torch.rand_like(img)
  ((img - target) ** 2).mean().backward()
  ids, _, b = raster_oracle.rasterize(SQUARE, SQUARE_TRIS, w, h, 1, 1)
  t64 = torch.as_tensor(tex, dtype=torch.float64).requires_grad_(True)
  uv = so.interpolate(torch.as_tensor(SQUARE_UV, dtype=torch.float64), ids, b,
                      SQUARE_TRIS, (0.0,))
  o64, _, _ = so.splat_shaded_pixels(
      torch.as_tensor(SQUARE, dtype=torch.float64), ids, b, SQUARE_TRIS,
      to.texture_map(uv, t64))
  err = helpers.rel_err(img.detach().cpu(), o64.detach())
  assert err <= 1e-5
  ((o64 - target.cpu().double()) ** 2).mean().backward()
  err = helpers.rel_err(t.grad.cpu(), t64.grad)
  assert err <= 1e-4


@pytest.mark.parametrize("kind,fixture", SPOT_CASES)
def test_spot_textured_matches_reference_png(kind, fixture, device):
  # shaders_test.py:75-172 through the GPU render path (OVER): texture
  # lookup on the resampling kernel, the library's lights and normals.
  from diffren_amd import constants
  from diffren_amd import mesh
  pos, uv, t, proj = spot_scene()
  p = torch.as_tensor(pos, device=device)
  attrs = {"uvs": torch.as_tensor(uv, device=device), "positions": p,
           "normals": mesh.compute_vertex_normals(
               p, torch.as_tensor(t, device=device))}
  path = os.path.join(helpers.FIXTURES, "spot_texture.png")

  def shader(a):
    textured = shading.texture_map(a["uvs"], path)
    return spot_shade(kind, textured, a["normals"], a["positions"],
                      (shading.diffuse_light, shading.point_light))

  img = render.render_triangles(
      p, attrs, torch.as_tensor(t, device=device),
      torch.as_tensor(proj, device=device), 640, 480, shader,
      compositing_mode=constants.CompositingMode.OVER)
  ok, msg = helpers.check_image(img.cpu().numpy(), fixture)
  assert ok, msg
  assert helpers.rel_err(img.cpu(), spot_oracle(kind)) <= 1e-5
