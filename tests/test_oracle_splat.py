"""Pins of the interpolate / splat / over oracle (oracle/splat_oracle.py).

  * reference PNG fixtures at the reference's tolerance (0.04, 0.5%
    outliers: compare_images.py:26-70): Two_Triangles_Splat_* (splat_test.py:
    47-82) and Unlit_Cube_* (interpolate_test.py:51-75);
  * closed forms of the splat; `over` identities;
  * float64 central finite differences (eps 1e-3 as the reference tests use
    for optimisation, rasterization frozen) of the oracle's autodiff gradient.
"""
import math

import numpy as np
import pytest
import torch

import helpers
from diffren_amd import camera as cam_lib
from oracle import raster_oracle
from oracle import splat_oracle as so


def test_splat_coefficient_closed_form():
  w = 1 + 4 * math.exp(-2) + 4 * math.exp(-4)
  assert abs(so.splat_coefficient() - 1.05 / w) < 1e-15
  assert abs(so.splat_coefficient() - 0.6503143820850361) < 1e-15


def _flat(H, W, L=1, z=0.5, color=(0.2, 0.4, 0.6, 1.0)):
  ys, xs = np.meshgrid(np.arange(H), np.arange(W), indexing="ij")
  xyz = np.stack([xs + 0.5, ys + 0.5, np.full(xs.shape, z)], -1)
  xyz = torch.tensor(np.broadcast_to(xyz, (L, H, W, 3)).copy(),
                     dtype=torch.float64)
  col = torch.tensor(color, dtype=torch.float64).expand(L, H, W, 4).clone()
  return xyz, col


def test_flat_surface_normalises_to_its_colour():
  xyz, col = _flat(7, 9)
  out, acc, norm = so.splat_at_pixel_centers(xyz, col,
                                             torch.ones(1, 7, 9, 1,
                                                        dtype=torch.float64))
  # Interior: weight 1.05 -> normalised back to the colour.
  assert torch.allclose(acc[1, 1:-1, 1:-1, 3], torch.tensor(1.05,
                                                           dtype=torch.float64))
  assert torch.allclose(out[1:-1, 1:-1], col[0, 1:-1, 1:-1])
  # Borders are under-covered and not normalised.
  assert acc[1, 0, 0, 3] < 1.0
  # A single layer never lands in the fg / bg accumulators.
  assert acc[0].abs().max() == 0 and acc[2].abs().max() == 0


def test_over_identities():
  x = torch.rand(1, 4, 5, 4, dtype=torch.float64)
  assert torch.equal(so.over(x), x[0])
  a = torch.rand(3, 4, 5, 4, dtype=torch.float64)
  a[0, ..., 3] = 1.0  # opaque front layer hides everything behind it
  assert torch.allclose(so.over(a), a[0])


def test_two_triangle_splat_matches_reference_pngs():
  v = np.array([[-0.2, -0.2, 0, 1], [0.5, -0.2, 0, 1], [0.5, 0.5, 0, 1],
                [0.2, -0.2, 0.5, 1], [-0.5, -0.2, 0.5, 1],
                [-0.5, 0.5, 0.5, 1]], np.float32)
  t = np.array([[0, 1, 2], [3, 5, 4]], np.int32)
  colors = np.array([[0, 1, 0, 1]] * 3 + [[1, 0, 0, 1]] * 3, np.float32)
  ids, _, b = raster_oracle.rasterize(v, t, 32, 32, 2, 1)
  shaded = so.interpolate(torch.as_tensor(colors), ids, b, t, (0.0,))
  out, _, norm = so.splat_shaded_pixels(torch.as_tensor(v), ids, b, t,
                                        shaded)
  ok, msg = helpers.check_image(out.numpy(),
                                "Two_Triangles_Splat_Composite.png",
                                resize_to=(512, 512))
  assert ok, msg
  for i in range(3):
    ok, msg = helpers.check_image(norm[i].numpy(),
                                  f"Two_Triangles_Splat_Layer_{i}.png",
                                  resize_to=(512, 512))
    assert ok, (i, msg)


def _test_perspective():
  """test_utils.make_perspective_matrix(0) (test_utils.py:76-96)."""
  fov, aspect, near, far = 40.0, 640 / 480, 0.01, 10.0
  f = 1.0 / np.tan(fov * (math.pi / 360.0))
  a = -(far + near) / (far - near)
  b = -2.0 * (far * near / (far - near))
  return np.array([[f / aspect, 0, 0, 0], [0, f, 0, 0], [0, 0, a, b],
                   [0, 0, -1, 0]])


@pytest.mark.parametrize("layers", [1, 2, 3])
def test_unlit_cube_matches_reference_pngs(layers):
  pos = np.array([[-1, -1, 1], [-1, -1, -1], [-1, 1, -1], [-1, 1, 1],
                  [1, -1, 1], [1, -1, -1], [1, 1, -1], [1, 1, 1]], np.float32)
  tris = np.array([[0, 1, 2], [2, 3, 0], [3, 2, 6], [6, 7, 3], [7, 6, 5],
                   [5, 4, 7], [4, 5, 1], [1, 0, 4], [5, 6, 2], [2, 1, 5],
                   [7, 4, 0], [0, 3, 7]], np.int32)
  rgba = np.concatenate([pos * 0.5 + 0.5, np.ones((8, 1), np.float32)], 1)
  eyes = [(2.0, 3.0, 6.0), (-3.0, 1.0, 6.0)]
  for view, eye in enumerate(eyes):
    proj = _test_perspective() @ cam_lib.look_at(eye, (0, 0, 0), (0, 1, 0))
    clip = so.transform(proj.astype(np.float32), pos, torch.float32).numpy()
    ids, _, b = raster_oracle.rasterize(clip, tris, 640, 480, layers, 0)
    img = so.interpolate(torch.as_tensor(rgba), ids, b, tris, (0.0,))
    img = img * torch.as_tensor(ids >= 0, dtype=torch.float32)[..., None]
    for l in range(layers):
      ok, msg = helpers.check_image(img[l].numpy(),
                                    f"Unlit_Cube_{view}_{l}.png")
      assert ok, (view, l, msg)


@pytest.mark.parametrize("num_components", [4, 2])
def test_oracle_optimizes_single_triangle(num_components):
  """splat_test.py:84-168 on the oracle: the plain-raster loss has no vertex
  gradient, the splat loss has one, and 15 unit-rate descent steps on the
  float64 gradient halve the splat loss."""
  init = np.array([[0, 0, 0, 1], [0.5, 0, 0, 1], [0.5, 0.5, 0, 1]],
                  np.float32)
  target_v = np.array([[-0.25, 0, 0, 1], [0.25, 0, 0, 1], [0.25, 0.5, 0, 1]],
                      np.float32)
  tris = np.array([[0, 1, 2]], np.int32)
  colors = torch.tensor([[0, 0., 0.8, 1.0], [0, 0.8, 0, 1.0],
                         [0, 0., 0.8, 1.0]],
                        dtype=torch.float64)[:, -num_components:]

  def draw(v, do_splat):
    ids, _, b = raster_oracle.rasterize(
        v.detach().numpy().astype(np.float32), tris, 32, 32, 1, 1)
    shaded = so.interpolate(colors, ids, b, tris, (0.0,))
    if do_splat:
      return so.splat_shaded_pixels(v, ids, b, tris, shaded)[0]
    return so.over(shaded * torch.as_tensor(ids >= 0,
                                            dtype=torch.float64)[..., None])

  target = draw(torch.as_tensor(target_v, dtype=torch.float64), False)
  v = torch.as_tensor(init, dtype=torch.float64)
  initial = None
  for i in range(15):
    vv = v.clone().requires_grad_(True)
    raster_loss = ((draw(vv, False) - target) ** 2).mean()
    assert not raster_loss.requires_grad
    loss = ((draw(vv, True) - target) ** 2).mean()
    loss.backward()
    if i == 0:
      assert torch.linalg.norm(vv.grad).item() > 0.01
      initial = loss.item()
    v = v - vv.grad
  final = ((draw(v, True) - target) ** 2).mean().item()
  assert final < 0.5 * initial, (initial, final)


def spike_scene():
  """splat_test.py:170-196: sphere.obj with vertex 32 scaled by an offset
  (2.0), view_1 projection, white RGBA colours, 640x480, one layer, BACK
  culling; rasterization of the spiked mesh, frozen (stop_gradient)."""
  v, t = helpers.load_flat_obj("sphere.obj")
  proj = (helpers.reference_perspective().astype(np.float32).astype(
      np.float64) @ helpers.reference_look_at("view_1").astype(
          np.float32).astype(np.float64)).astype(np.float32)
  spiked = v.copy()
  spiked[32] *= 2.0
  clip = so.transform(proj, spiked, torch.float32).numpy()
  ids, _, b = raster_oracle.rasterize(clip, t, 640, 480, 1, 1)
  return v, t, proj, ids, b


def spike_visualisation(image, jacobian):
  """splat_test.py:201-209: the normalised JVP over a faint image."""
  nj = (jacobian / jacobian.abs().max()) * 0.5 + 0.5
  return (nj * 0.95 + image * 0.05)[..., :3]


def test_oracle_spike_derivative_matches_reference_png():
  # splat_test.py:170-224: the forward-mode derivative of the image with
  # respect to the spike offset, at pixel error threshold 0.
  v, t, proj, ids, b = spike_scene()
  base = torch.as_tensor(v, dtype=torch.float64)
  cols = torch.ones((v.shape[0], 4), dtype=torch.float64)

  def f(off):
    verts = torch.cat([base[:32], base[32:33] * off, base[33:]], 0)
    return so.render_splat(verts, cols,
                           torch.as_tensor(proj, dtype=torch.float64), ids, b,
                           t)

  img, jac = torch.autograd.functional.jvp(
      f, (torch.tensor([2.0], dtype=torch.float64),),
      (torch.ones(1, dtype=torch.float64),))
  ok, msg = helpers.check_image(spike_visualisation(img, jac).numpy(),
                                "Spike_Derivative.png", exact=True)
  assert ok, msg


def test_sphere_disparity_matches_reference_png():
  # render_test.py:100-141: sphere.obj, OVER compositing, BACK culling, the
  # shader maps world position to normalised disparity -1 / z_camera.
  v, t = helpers.load_flat_obj("sphere.obj")
  look = helpers.reference_look_at("view_1")
  proj = (helpers.reference_perspective() @ look).astype(np.float32)
  clip = so.transform(proj, v, torch.float32).numpy()
  ids, _, b = raster_oracle.rasterize(clip, t, 640, 480, 1, 1)
  pos = so.interpolate(torch.as_tensor(v), ids, b, t, (0.0,))
  cam = so.transform(look.astype(np.float32), pos.reshape(-1, 3),
                     torch.float32)
  disp = -1.0 / cam[:, 2]
  disp = (disp - helpers.DISPARITY_MIN) / (helpers.DISPARITY_MAX -
                                           helpers.DISPARITY_MIN)
  img = disp.reshape(ids.shape) * torch.as_tensor(ids >= 0,
                                                  dtype=torch.float32)
  ok, msg = helpers.check_image(img[0, :, :, None].numpy(),
                                "Sphere_Disparity.png")
  assert ok, msg


def test_gradient_matches_finite_differences():
  """d loss / d vertices and colours of the float64 oracle vs central
  differences (eps 1e-3) with the rasterization frozen."""
  from diffren_amd import scenes
  s = scenes.icosphere_scene(seed=2, subdivisions=1)
  W = H = 24
  clip = so.transform(s["camera"], s["vertices"], torch.float32).numpy()
  ids, _, b = raster_oracle.rasterize(clip, s["triangles"], W, H, 2, 1)
  rng = np.random.default_rng(0)
  target = torch.as_tensor(rng.uniform(size=(H, W, 4)))

  def loss_fn(v, c):
    out = so.render_splat(v, c, s["camera"], ids, b, s["triangles"])
    return ((out - target) ** 2).mean()

  v = torch.as_tensor(s["vertices"], dtype=torch.float64).requires_grad_(True)
  c = torch.as_tensor(s["colors"], dtype=torch.float64).requires_grad_(True)
  loss_fn(v, c).backward()
  eps = 1e-3
  for idx in rng.choice(v.numel(), 16, replace=False):
    vp = v.detach().clone().reshape(-1)
    vm = vp.clone()
    vp[idx] += eps
    vm[idx] -= eps
    fd = (loss_fn(vp.reshape(v.shape), c.detach()) -
          loss_fn(vm.reshape(v.shape), c.detach())) / (2 * eps)
    g = v.grad.reshape(-1)[idx]
    assert abs(fd - g) <= 1e-6 + 1e-3 * abs(g), (idx, fd, g)
  for idx in rng.choice(c.numel(), 16, replace=False):
    cp = c.detach().clone().reshape(-1)
    cm = cp.clone()
    cp[idx] += eps
    cm[idx] -= eps
    fd = (loss_fn(v.detach(), cp.reshape(c.shape)) -
          loss_fn(v.detach(), cm.reshape(c.shape))) / (2 * eps)
    g = c.grad.reshape(-1)[idx]
    assert abs(fd - g) <= 1e-7 + 1e-4 * abs(g), (idx, fd, g)
