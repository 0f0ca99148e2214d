"""The texture-resampling oracle (oracle/texture_oracle.py) pinned on the
CPU: the reference's Textured_Square.png through the oracle render path,
closed forms and float64 finite differences."""
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
