"""render.render_triangles in its non-splat compositing modes (OVER, NONE)
through the GPU kernels: the reference's PNG fixtures at the reference's
tolerance, and the CPU oracle within 1e-5 relative.

Scenes: the unlit cube of interpolate_test.py:30-75 and render_test.py:
57-98, and the sphere disparity map of render_test.py:100-141.
"""
import numpy as np
import pytest
import torch

import helpers
from diffren_amd import constants
from diffren_amd import render
from diffren_amd import transforms
from oracle import raster_oracle
from oracle import splat_oracle as so

pytestmark = pytest.mark.gpu

CUBE = np.array([[-1, -1, 1], [-1, -1, -1], [-1, 1, -1], [-1, 1, 1],
                 [1, -1, 1], [1, -1, -1], [1, 1, -1], [1, 1, 1]], np.float32)
CUBE_TRIS = np.array([[0, 1, 2], [2, 3, 0], [3, 2, 6], [6, 7, 3], [7, 6, 5],
                      [5, 4, 7], [4, 5, 1], [1, 0, 4], [5, 6, 2], [2, 1, 5],
                      [7, 4, 0], [0, 3, 7]], np.int32)
RGBA = np.concatenate([CUBE * 0.5 + 0.5, np.ones((8, 1), np.float32)], 1)


def _projection(view):
  return (helpers.reference_perspective() @
          helpers.reference_look_at(view)).astype(np.float32)


def _oracle_layers(proj, layers, cull):
  clip = so.transform(proj, CUBE, torch.float32).numpy()
  ids, _, b = raster_oracle.rasterize(clip, CUBE_TRIS, 640, 480, layers, cull)
  img = so.interpolate(torch.as_tensor(RGBA), ids, b, CUBE_TRIS, (0.0,))
  return img * torch.as_tensor(ids >= 0, dtype=torch.float32)[..., None]


@pytest.mark.parametrize("layers", [1, 2, 3])
def test_unlit_cube_layers(layers, device):
  # interpolate_test.py:51-75: CompositingMode.NONE returns every layer.
  for view_index, view in enumerate(("view_1", "view_2")):
    proj = _projection(view)
    img = render.render_triangles(
        torch.as_tensor(CUBE, device=device),
        {"rgba": torch.as_tensor(RGBA, device=device)},
        torch.as_tensor(CUBE_TRIS, device=device),
        torch.as_tensor(proj, device=device), 640, 480, lambda a: a["rgba"],
        num_layers=layers, face_culling_mode=constants.FaceCullingMode.NONE,
        compositing_mode=constants.CompositingMode.NONE)
    assert tuple(img.shape) == (layers, 480, 640, 4)
    want = _oracle_layers(proj, layers, 0)
    assert helpers.rel_err(img.cpu(), want) <= 1e-5
    for l in range(layers):
      ok, msg = helpers.check_image(img[l].cpu().numpy(),
                                    f"Unlit_Cube_{view_index}_{l}.png")
      assert ok, (view, l, msg)


def test_unlit_cube_over(device):
  # render_test.py:57-98: OVER of one layer, both views.
  for view_index, view in enumerate(("view_1", "view_2")):
    proj = _projection(view)
    img = render.render_triangles(
        torch.as_tensor(CUBE, device=device),
        {"rgba": torch.as_tensor(RGBA, device=device)},
        torch.as_tensor(CUBE_TRIS, device=device),
        torch.as_tensor(proj, device=device), 640, 480, lambda a: a["rgba"],
        face_culling_mode=constants.FaceCullingMode.NONE,
        compositing_mode=constants.CompositingMode.OVER)
    assert tuple(img.shape) == (480, 640, 4)
    assert helpers.rel_err(img.cpu(), _oracle_layers(proj, 1, 0)[0]) <= 1e-5
    ok, msg = helpers.check_image(img.cpu().numpy(),
                                  f"Unlit_Cube_{view_index}_0.png")
    assert ok, (view, msg)


def test_over_composites_layers_front_to_back(device):
  # composite.py:20-37 over three layers of the cube with alpha 0.5.
  rgba = RGBA.copy()
  rgba[:, 3] = 0.5
  proj = _projection("view_1")
  img = render.render_triangles(
      torch.as_tensor(CUBE, device=device),
      {"rgba": torch.as_tensor(rgba, device=device)},
      torch.as_tensor(CUBE_TRIS, device=device),
      torch.as_tensor(proj, device=device), 640, 480, lambda a: a["rgba"],
      num_layers=3, face_culling_mode=constants.FaceCullingMode.NONE,
      compositing_mode=constants.CompositingMode.OVER)
  clip = so.transform(proj, CUBE, torch.float32).numpy()
  ids, _, b = raster_oracle.rasterize(clip, CUBE_TRIS, 640, 480, 3, 0)
  lay = so.interpolate(torch.as_tensor(rgba), ids, b, CUBE_TRIS, (0.0,))
  lay = lay * torch.as_tensor(ids >= 0, dtype=torch.float32)[..., None]
  want = lay[2]
  for l in (1, 0):
    want = lay[l] + (1.0 - lay[l][..., 3:]) * want
  assert helpers.rel_err(img.cpu(), want) <= 1e-5
  # Two faces overlap at most pixels the cube covers: alpha 0.75 there.
  assert (img[..., 3] == 0.75).float().mean().item() > 0.1


def test_sphere_disparity(device):
  # render_test.py:100-141: OVER, BACK culling, a shader that maps the
  # interpolated world position to normalised disparity.
  v, t = helpers.load_flat_obj("sphere.obj")
  look = torch.as_tensor(helpers.reference_look_at("view_1").astype(
      np.float32), device=device)
  proj = _projection("view_1")

  def shader(a):
    pos = a["world_position"]
    cam = transforms.transform_homogeneous(look, pos.reshape(-1, 3))
    disp = -1.0 / cam[:, 2]
    disp = (disp - helpers.DISPARITY_MIN) / (helpers.DISPARITY_MAX -
                                             helpers.DISPARITY_MIN)
    disp = disp.reshape(pos.shape[:-1])
    return torch.stack([disp, torch.ones_like(disp)], -1)

  vt = torch.as_tensor(v, device=device)
  img = render.render_triangles(
      vt, {"world_position": vt}, torch.as_tensor(t, device=device),
      torch.as_tensor(proj, device=device), 640, 480, shader,
      compositing_mode=constants.CompositingMode.OVER)
  ok, msg = helpers.check_image(img[..., 0:1].cpu().numpy(),
                                "Sphere_Disparity.png")
  assert ok, msg
