"""Camera matrix builders used to set up scenes (numpy, float64 math).

Restates diffren/jax/camera.py: `perspective` (:23-44),
`perspective_from_intrinsics` (:47-92), `ortho` (:95-123) and `look_at`
(:126-184), gluPerspective / glOrtho / gluLookAt conventions.  Matrices are
returned as float64 [4, 4] for column vectors (clip = M @ [x, y, z, 1]);
cast to float32 before handing them to the render path.
"""
import math

import numpy as np


def ortho(left, right, bottom, top, near, far):
  depth = far - near
  return np.array(
      [[2.0 / (right - left), 0.0, 0.0, -(right + left) / (right - left)],
       [0.0, 2.0 / (top - bottom), 0.0, -(top + bottom) / (top - bottom)],
       [0.0, 0.0, -2.0 / depth, -(far + near) / depth],
       [0.0, 0.0, 0.0, 1.0]], dtype=np.float64)


def perspective_from_intrinsics(focal_x, focal_y, center_offset_x,
                                center_offset_y, near_clip, far_clip,
                                image_width, image_height):
  a = near_clip + far_clip
  b = near_clip * far_clip
  persp = np.array([[focal_x, 0.0, -center_offset_x, 0.0],
                    [0.0, focal_y, -center_offset_y, 0.0],
                    [0.0, 0.0, a, b],
                    [0.0, 0.0, -1.0, 0.0]], dtype=np.float64)
  ndc = ortho(image_width * -0.5, image_width * 0.5, image_height * -0.5,
              image_height * 0.5, near_clip, far_clip)
  return ndc @ persp


def perspective(aspect_ratio, fov_y, near_clip, far_clip):
  """gluPerspective; fov_y in degrees."""
  focal_y = 1.0 / math.tan(fov_y * (math.pi / 360.0))
  return perspective_from_intrinsics(focal_y / aspect_ratio, focal_y, 0.0, 0.0,
                                     near_clip, far_clip, 2, 2)


def look_at(eye, center, world_up):
  """gluLookAt: world -> eye space, right handed."""
  eye = np.asarray(eye, np.float64)
  center = np.asarray(center, np.float64)
  world_up = np.asarray(world_up, np.float64)
  forward = center - eye
  n = np.linalg.norm(forward)
  if n < 1e-6:
    raise ValueError(
        "Camera matrix is degenerate because eye and center are close.")
  forward = forward / n
  side = np.cross(forward, world_up)
  n = np.linalg.norm(side)
  if n < 1e-6:
    raise ValueError("Camera matrix is degenerate because up and gaze are "
                     "close or up is degenerate.")
  side = side / n
  up = np.cross(side, forward)
  rot = np.eye(4)
  rot[0, :3], rot[1, :3], rot[2, :3] = side, up, -forward
  trans = np.eye(4)
  trans[:3, 3] = -eye
  return rot @ trans
