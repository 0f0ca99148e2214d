"""The differentiable-barycentrics oracle (oracle/diffbary_oracle.py) pinned
on the CPU: the reference's golden Jacobian and PNG fixtures, the
rasterizer's independently computed barycentrics, invariants and float64
finite differences."""
import numpy as np
import pytest
import torch

import helpers
import scenes as golden_scenes
from oracle import diffbary_oracle as dbo
from oracle import raster_oracle

# rasterize_test.py:216-224: d barycentrics[0, 24, 32, i] / d ndc[j, k] for
# the triangle below at 64x48 (identity camera, so clip = (ndc, 1)).
GOLDEN_NDC = np.array([[-0.5, -0.5, 0.8], [0.0, 0.5, 0.3], [0.5, -0.5, 0.3]],
                      np.float32)
GOLDEN_JACOBIAN = np.array([[[0.22395831, 0.11197916, -0.],
                             [0.5208334, 0.2604167, -0.],
                             [0.2552083, 0.12760416, -0.]],
                            [[0., -0.22395831, -0.],
                             [-0., -0.5208334, -0.],
                             [0., -0.2552083, -0.]],
                            [[-0.22395831, 0.11197916, -0.],
                             [-0.5208334, 0.2604167, -0.],
                             [-0.2552083, 0.12760416, -0.]]])
TRI = np.array([[0, 1, 2]], np.int32)


def _clip(ndc):
  return np.concatenate([ndc, np.ones((ndc.shape[0], 1), ndc.dtype)], 1)


def test_oracle_golden_jacobian():
  ids, _, _ = raster_oracle.rasterize(_clip(GOLDEN_NDC), TRI, 64, 48, 1, 0)
  assert ids[0, 24, 32] == 0

  def f(ndc):
    clip = torch.cat([ndc, torch.ones(3, 1, dtype=ndc.dtype)], 1)
    return dbo.barycentrics_torch(clip, TRI, ids, 64, 48)[0, 24, 32]

  jac = torch.autograd.functional.jacobian(
      f, torch.as_tensor(GOLDEN_NDC, dtype=torch.float64))
  np.testing.assert_array_almost_equal(jac.numpy(), GOLDEN_JACOBIAN)


@pytest.mark.parametrize("w,fixture", [
    ((1.0, 1.0, 1.0), "Simple_Triangle.png"),
    ((0.2, 0.5, 2.0), "Perspective_Corrected_Triangle.png"),
])
def test_oracle_matches_reference_png(w, fixture):
  # rasterize_test.py:67-98,119-148 with compute_diff_barys=True.
  v = golden_scenes.SIMPLE_TRIANGLE * np.asarray(w, np.float32)[:, None]
  ids, _, _ = raster_oracle.rasterize(v, TRI, 640, 480, 1, 0)
  b = dbo.barycentrics_f32(v, TRI, ids, 640, 480)
  ok, msg = helpers.check_image(b[0], fixture)
  assert ok, msg


def test_oracle_cube_matches_reference_png():
  # rasterize_test.py:150-165 with compute_diff_barys=True.
  v, t = golden_scenes.CUBE_VERTICES, golden_scenes.CUBE_TRIANGLES
  ids, _, _ = raster_oracle.rasterize(v, t, 640, 480, 1, 0)
  b = dbo.barycentrics_f32(v, t, ids, 640, 480)
  ok, msg = helpers.check_image(b[0], "Barycentrics_Cube.png",
                                add_transparency=False)
  assert ok, msg


@pytest.mark.parametrize("seed", range(4))
def test_oracle_agrees_with_rasterizer_barycentrics(seed):
  # The rasterizer's barycentrics come from 16.16 fixed-point edge functions
  # (rasterize_triangles_impl.cc:79-82), an independent computation of the
  # same perspective-correct coordinates.
  rng = np.random.default_rng(seed)
  v = np.concatenate([rng.uniform(-1, 1, (30, 2)), rng.uniform(-0.5, 0.5,
                                                             (30, 1)),
                      rng.uniform(0.8, 2.0, (30, 1))], 1).astype(np.float32)
  v[:, :3] *= v[:, 3:]
  t = rng.integers(0, 30, (40, 3)).astype(np.int32)
  ids, _, braster = raster_oracle.rasterize(v, t, 96, 64, 2, 0)
  b = dbo.barycentrics_f32(v, t, ids, 96, 64)
  fg = ids >= 0
  assert fg.any()
  # Fixed-point rounding of the vertices moves sliver triangles' coordinates
  # most; nearly every pixel agrees to 1e-3.
  d = np.abs(b[fg] - braster[fg]).max(-1)
  assert np.median(d) <= 2e-5
  assert np.mean(d > 1e-3) <= 1e-3
  assert (b[~fg] == 0).all()
  assert np.allclose(b[fg].sum(-1), 1.0, atol=1e-5)


def test_oracle_float32_matches_float64():
  v, t = golden_scenes.CUBE_VERTICES, golden_scenes.CUBE_TRIANGLES
  ids, _, _ = raster_oracle.rasterize(v, t, 64, 48, 2, 0)
  b32 = dbo.barycentrics_f32(v, t, ids, 64, 48)
  b64 = dbo.barycentrics_torch(torch.as_tensor(v, dtype=torch.float64), t,
                               ids, 64, 48).numpy()
  assert np.abs(b32 - b64).max() <= 1e-5


def test_oracle_vertex_coordinates():
  # At a vertex the coordinates are the unit vector of that vertex.
  v = np.array([[-1.0, -1.0, 0.0, 1.0], [3.0, -1.0, 0.0, 1.0],
                [-1.0, 3.0, 0.0, 1.0]], np.float32)
  # 2x2 image: pixel (0, 0) is at ndc (-0.5, -0.5).
  ids = np.zeros((1, 2, 2), np.int32)
  b = dbo.barycentrics_f32(v, TRI, ids, 2, 2)
  np.testing.assert_allclose(b[0, 0, 0], [0.75, 0.125, 0.125], atol=1e-7)


def test_oracle_finite_differences():
  rng = np.random.default_rng(7)
  v = golden_scenes.CUBE_VERTICES.astype(np.float64)
  t = golden_scenes.CUBE_TRIANGLES
  ids, _, _ = raster_oracle.rasterize(v.astype(np.float32), t, 32, 24, 2, 0)
  wgt = torch.as_tensor(rng.normal(size=(2, 24, 32, 3)))

  def loss(c):
    return (dbo.barycentrics_torch(c, t, ids, 32, 24) * wgt).sum()

  c = torch.as_tensor(v, dtype=torch.float64).requires_grad_(True)
  loss(c).backward()
  eps = 1e-6
  for _ in range(12):
    i, j = int(rng.integers(0, v.shape[0])), int(rng.choice([0, 1, 3]))
    cp = torch.as_tensor(v).clone()
    cm = torch.as_tensor(v).clone()
    cp[i, j] += eps
    cm[i, j] -= eps
    fd = (loss(cp) - loss(cm)).item() / (2 * eps)
    assert abs(fd - c.grad[i, j].item()) <= 1e-6 * max(1.0, abs(fd))
  assert (c.grad[:, 2] == 0).all()
