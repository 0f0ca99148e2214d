"""The texture-resampling oracle (oracle/texture_oracle.py) pinned on the
CPU: the reference's Textured_Square.png through the oracle render path,
closed forms and float64 finite differences."""
import os

import numpy as np
import pytest
import torch

import helpers
from oracle import raster_oracle
from oracle import splat_oracle as so
from oracle import texture_oracle as to

# shaders_test.py:33-71
SQUARE = np.array([[-0.5, -0.5, 0.0, 1.0], [0.5, -0.5, 0.0, 1.0],
                   [0.5, 0.5, 0.0, 1.0], [-0.5, 0.5, 0.0, 1.0]], np.float32)
SQUARE[:, 0] *= 0.75
SQUARE_UV = np.array([[0.0, 0.0], [0.0, 1.0], [1.0, 1.0], [1.0, 0.0]],
                     np.float32)
SQUARE_TRIS = np.array([[0, 1, 2], [0, 2, 3]], np.int32)
SQUARE_TEXTURE = np.array([[1.0, 0.0, 0.0], [0.0, 0.0, 0.0], [0.0, 0.0, 0.0],
                           [0.0, 0.0, 1.0]], np.float32).reshape(2, 2, 3)


def test_oracle_textured_square_matches_reference_png():
  ids, _, b = raster_oracle.rasterize(SQUARE, SQUARE_TRIS, 640, 480, 1, 1)
  uv = so.interpolate(torch.as_tensor(SQUARE_UV), ids, b, SQUARE_TRIS,
                      (0.0,))
  shaded = to.texture_map(uv, torch.as_tensor(SQUARE_TEXTURE))
  img, _, _ = so.splat_shaded_pixels(torch.as_tensor(SQUARE), ids, b,
                                     SQUARE_TRIS, shaded)
  ok, msg = helpers.check_image(img.numpy(), "Textured_Square.png")
  assert ok, msg


def test_oracle_closed_forms():
  rng = np.random.default_rng(0)
  img = torch.as_tensor(rng.uniform(size=(5, 7, 3)))
  # A texel centre returns the texel; a midpoint the mean of two texels.
  out = to.bilinear_resample(img, torch.tensor([[3.5, 2.5], [3.0, 2.5]],
                                               dtype=torch.float64))
  assert torch.allclose(out[0], img[2, 3])
  assert torch.allclose(out[1], 0.5 * (img[2, 2] + img[2, 3]))
  # Edge padding clamps; constant padding fades to 0 half a texel out.
  far = torch.tensor([[-3.0, 1.5], [100.0, 4.5]], dtype=torch.float64)
  e = to.bilinear_resample(img, far, "edge")
  assert torch.allclose(e[0], img[1, 0]) and torch.allclose(e[1], img[4, 6])
  c = to.bilinear_resample(img, torch.tensor([[0.0, 2.5], [-3.0, 2.5]],
                                             dtype=torch.float64))
  assert torch.allclose(c[0], 0.5 * img[2, 0]) and (c[1] == 0).all()
  # A constant image resamples to itself wherever the edge clamps.
  const = torch.full((4, 4, 2), 0.25, dtype=torch.float64)
  loc = torch.as_tensor(rng.uniform(-2, 6, size=(50, 2)))
  assert torch.allclose(to.bilinear_resample(const, loc, "edge"),
                        torch.full((50, 2), 0.25, dtype=torch.float64))


@pytest.mark.parametrize("pad", ["constant", "edge"])
def test_oracle_finite_differences(pad):
  rng = np.random.default_rng(1)
  img = torch.as_tensor(rng.uniform(size=(6, 5, 2))).requires_grad_(True)
  loc = torch.as_tensor(rng.uniform(-1, 7, size=(40, 2))).requires_grad_(True)
  wgt = torch.as_tensor(rng.normal(size=(40, 2)))

  def f(i, l):
    return (to.bilinear_resample(i, l, pad) * wgt).sum()

  f(img, loc).backward()
  eps = 1e-6
  for k in range(12):
    n, d = int(rng.integers(0, 40)), int(rng.integers(0, 2))
    lp, lm = loc.detach().clone(), loc.detach().clone()
    lp[n, d] += eps
    lm[n, d] -= eps
    fd = (f(img.detach(), lp) - f(img.detach(), lm)).item() / (2 * eps)
    assert abs(fd - loc.grad[n, d].item()) <= 1e-6 * max(1.0, abs(fd))
  for k in range(12):
    idx = tuple(int(rng.integers(0, s)) for s in img.shape)
    ip, im = img.detach().clone(), img.detach().clone()
    ip[idx] += eps
    im[idx] -= eps
    fd = (f(ip, loc.detach()) - f(im, loc.detach())).item() / (2 * eps)
    assert abs(fd - img.grad[idx].item()) <= 1e-6 * max(1.0, abs(fd))


BOUNDARY_CASES = [  # image_test.py:62-87: (x, y) on a 30x10 ramp -> value
    ((0.5, 0.5), 1), ((0.1, 1.5), 0.6), ((-0.5, 2), 0), ((-1.3, 3), 0),
    ((-200, 3), 0), ((-np.inf, 4), 0), ((9.5, 0.5), 10), ((9.7, 0.5), 8.0),
    ((10.5, 1), 0), ((10.52, 2), 0), ((210, 3), 0), ((np.inf, 1), 0),
    ((0.5, -0.4), 0.1), ((2, -0.5), 0), ((3, -1.3), 0), ((2, -200), 0),
    ((1, -np.inf), 0), ((0.5, 29.5), 1), ((1.5, 29.8), 1.4), ((2, 30.5), 0),
    ((3, 30.52), 0), ((2, 230), 0), ((1, np.inf), 0)]


def boundary_inputs():
  """image_test.py:46-108 (HWC): a [30, 10, 3] ramp 1..10 along x sampled
  at its pixel centres, with the boundary cases in column 0."""
  img = np.repeat(np.arange(1, 11, dtype=np.float32)[None, :], 30, 0)
  img = np.repeat(img[..., None], 3, 2)
  x, y = np.meshgrid(np.arange(10), np.arange(30))
  loc = (np.stack((x, y), -1) + 0.5).astype(np.float32)
  want = img.copy()
  for k, (xy, value) in enumerate(BOUNDARY_CASES):
    loc[k, 0] = xy
    want[k, 0] = value
  return img, loc, want


def test_oracle_recovers_input():
  # image_test.py:25-44: sampling at every pixel centre returns the image.
  n = 64
  img = np.arange(3 * n * n, dtype=np.float32).reshape(n, n, 3)
  x, y = np.meshgrid(np.arange(n), np.arange(n))
  loc = (np.stack((x, y), -1) + 0.5).astype(np.float32)
  out = to.bilinear_resample(torch.as_tensor(img), torch.as_tensor(loc))
  assert np.array_equal(out.numpy(), img)


def test_oracle_boundary_cases():
  img, loc, want = boundary_inputs()
  out = to.bilinear_resample(torch.as_tensor(img), torch.as_tensor(loc))
  np.testing.assert_array_almost_equal(want, out.numpy(), decimal=5)


SPOT_EYE = (0.2, 0.3, -0.5)  # test_utils.py:51-53, world up (0, -1, 0)
SPOT_CASES = [("diffuse", "Spot_Textured.png"),
              ("point", "Spot_Textured_Pt_Light.png"),
              ("phong_point", "Spot_Textured_Phong_Point.png")]


def spot_scene():
  """shaders_test.py:111-123 for spot_triangulated.obj: positions scaled
  by 0.2 (test_utils.py:98-107), uvs, computed vertex normals, the 'spot'
  look-at with make_perspective_matrix(0)."""
  v, t = helpers.load_flat_obj("spot_triangulated.obj")
  v = v.copy()
  v[:, :3] *= np.float32(0.2)
  look = helpers.reference_look_at_eye(SPOT_EYE, (0.0, -1.0, 0.0))
  proj = (helpers.reference_perspective().astype(np.float32).astype(
      np.float64) @ look.astype(np.float32).astype(np.float64)).astype(
          np.float32)
  return v[:, :3].copy(), v[:, 3:5].copy(), t, proj


def spot_shade(kind, textured, normals, positions, light_fns):
  """shaders_test.py:125-147 with the library's (or the oracle's) lights."""
  diffuse, point = light_fns
  if kind == "diffuse":
    return diffuse(textured, normals)
  if kind == "point":
    return point(textured, normals, positions, SPOT_EYE)
  return point(textured, normals, positions, SPOT_EYE, SPOT_EYE, ks=1.0,
               ns=50.0)


def spot_oracle(kind):
  pos, uv, t, proj = spot_scene()
  nrm = to.vertex_normals(torch.as_tensor(pos), t)
  clip = so.transform(proj, pos, torch.float32).numpy()
  ids, _, b = raster_oracle.rasterize(clip, t, 640, 480, 1, 1)
  tex = to.load_texture(os.path.join(helpers.FIXTURES, "spot_texture.png"))
  attrs = [so.interpolate(torch.as_tensor(a), ids, b, t, (0.0,))[0]
           for a in (uv, pos, nrm)]
  textured = to.texture_map(attrs[0], torch.as_tensor(tex))
  lit = spot_shade(
      kind, textured, attrs[2], attrs[1],
      (to.diffuse_light,
       lambda d, n, p, l, e=(0.0, 0.0, 0.0), ks=0.0, ns=0.0:
       to.point_light(d, n, p, l, e, ks=ks, ns=ns)))
  return lit * torch.as_tensor(ids[0] >= 0, dtype=torch.float32)[..., None]


@pytest.mark.parametrize("kind,fixture", SPOT_CASES)
def test_oracle_spot_textured_matches_reference_png(kind, fixture):
  ok, msg = helpers.check_image(spot_oracle(kind).numpy(), fixture)
  assert ok, msg
