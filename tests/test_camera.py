"""Camera matrix builders (camera_test.py): host logic, no GPU."""
import numpy as np
import pytest

import helpers
from diffren_amd import camera


def test_perspective():
  # camera_test.py:29-49 against test_utils.make_perspective_matrix.
  got = camera.perspective(640 / 480, 40.0, 0.01, 10.0)
  np.testing.assert_allclose(helpers.reference_perspective(), got, rtol=1e-6)
  f = 1.0 / np.tan(60.0 * (np.pi / 360.0))
  a, b = -(20.0 + 0.1) / 19.9, -2.0 * (20.0 * 0.1 / 19.9)
  want = np.array([[f / (640 / 480), 0, 0, 0], [0, f, 0, 0], [0, 0, a, b],
                   [0, 0, -1, 0]])
  np.testing.assert_allclose(want, camera.perspective(640 / 480, 60.0, 0.1,
                                                      20.0), rtol=1e-6)


def test_perspective_from_intrinsics_center_of_projection():
  # camera_test.py:51-79.
  m = camera.perspective_from_intrinsics(100.0, 100.0, 50.0, 25.0, 10.0,
                                         100.0, 100, 50)
  v = np.array([[0.0, 0.0, -100.0, 1.0], [-50.0, -25.0, -100.0, 1.0],
                [-50.0, -50.0, -100.0, 1.0], [-50.0, -50.0, -100.0, 1.0],
                [-5.0, -5.0, -10.0, 1.0]])
  p = v @ m.T
  np.testing.assert_allclose(
      p[:, :3] / p[:, 3:4],
      np.array([[1.0, 1.0, 1.0], [0.0, 0.0, 1.0], [0.0, -1.0, 1.0],
                [0.0, -1.0, 1.0], [0.0, -1.0, -1.0]]), rtol=1e-6, atol=1e-12)


def test_ortho():
  # camera_test.py:81-93.
  m = camera.ortho(-50.0, 50.0, -25.0, 25.0, 10.0, 100.0)
  np.testing.assert_allclose(m @ np.array([50.0, 25.0, -100.0, 1.0]),
                             np.ones(4), rtol=1e-6)


@pytest.mark.parametrize("msg,eye,center,up", [
    ("up and gaze are close", (0.0, 0.0, 1.0), (0.0, 0.0, 0.0),
     (0.0, 0.0, 1.0)),
    ("eye and center are close", (0.0, 0.0, 0.0), (0.0, 0.0, 0.0),
     (0.0, 1.0, 0.0))])
def test_degenerate_look_at(msg, eye, center, up):
  # camera_test.py:95-111.
  with pytest.raises(ValueError, match=msg):
    camera.look_at(eye, center, up)


def test_look_at_matches_test_views():
  # test_utils.py:55-77 view_1.
  np.testing.assert_allclose(camera.look_at((2.0, 3.0, 6.0), (0, 0, 0),
                                            (0, 1, 0)),
                             helpers.reference_look_at("view_1"), rtol=1e-12)
