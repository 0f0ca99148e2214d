"""Shared test helpers: image tolerance semantics of the reference tests."""
import os

import numpy as np

FIXTURES = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden",
                        "reference_fixtures")


def pil_formatted(image):
  """compare_images.py:213-262: [0,1] float image -> uint8 RGBA."""
  image = np.asarray(image, np.float32)
  if image.ndim == 2:
    image = image[..., None]
  if image.shape[2] == 1:
    image = np.tile(image, [1, 1, 3])
  if image.shape[2] == 3:
    image = np.concatenate([image, np.ones(image.shape[:2] + (1,),
                                           np.float32)], 2)
  return np.clip(255.0 * image, 0.0, 255.0).astype(np.uint8)


def load_fixture(name):
  from PIL import Image
  return np.array(Image.open(os.path.join(FIXTURES, name)))


def images_are_near(baseline, result, max_outlier_fraction=0.005,
                    pixel_error_threshold=0.04):
  """compare_images.py:26-70."""
  if baseline.shape != result.shape:
    return False, f"shapes {baseline.shape} vs {result.shape}"
  diff = np.abs(baseline.astype(float) / 255.0 - result.astype(float) / 255.0)
  out = diff > pixel_error_threshold
  if baseline.ndim > 2:
    out = np.any(out, axis=2)
  frac = np.count_nonzero(out) / np.prod(baseline.shape[:2])
  return frac <= max_outlier_fraction, f"{frac:.5f} of pixels are outliers"


def check_image(image, fixture, resize_to=None, exact=False,
                add_transparency=True):
  """test_utils.py:143-172 semantics; returns (ok, message)."""
  from PIL import Image
  if add_transparency:
    img = pil_formatted(image)
  else:
    img = np.clip(255.0 * np.asarray(image, np.float32), 0.0,
                  255.0).astype(np.uint8)
  if resize_to:
    img = np.array(Image.fromarray(img).resize(resize_to,
                                               Image.Resampling.NEAREST))
  base = load_fixture(fixture)
  return images_are_near(base, img,
                         pixel_error_threshold=0.0 if exact else 0.04)


def rel_err(a, b):
  """max |a - b| / max(max |b|, tiny): the relative error used in tests."""
  a = np.asarray(a, np.float64)
  b = np.asarray(b, np.float64)
  scale = max(float(np.abs(b).max()) if b.size else 0.0, 1e-30)
  return float(np.abs(a - b).max()) / scale if b.size else 0.0


def same_bits(a, b):
  a = np.ascontiguousarray(a)
  b = np.ascontiguousarray(b)
  return a.shape == b.shape and np.array_equal(a.view(np.uint8),
                                               b.view(np.uint8))


# rasterize_test_utils.h:24-27
COLOR_CYCLE = np.array([255, 0, 0, 0, 255, 0, 0, 0, 255, 255, 0, 255, 0, 255,
                        255, 255, 255, 0, 128, 0, 0, 0, 128, 0, 0, 0, 128,
                        128, 0, 128, 0, 128, 128, 128, 128, 0],
                       np.uint8).reshape(12, 3)


def encode_barycentrics(barys, layer=0):
  """rasterize_test_utils.cc:157-170: uint8(b * 255), alpha 255."""
  b = np.asarray(barys, np.float32)[layer]
  img = np.empty(b.shape[:2] + (4,), np.uint8)
  img[..., :3] = (b * np.float32(255.0)).astype(np.uint8)
  img[..., 3] = 255
  return img


def encode_ids(ids, layer=0):
  """rasterize_test_utils.cc:172-193; background (-1 here, 0 in the C++
  tests' zero-filled buffers) is encoded as id 0."""
  i = np.asarray(ids)[layer]
  i = np.where(i < 0, 0, i)
  img = np.empty(i.shape + (4,), np.uint8)
  img[..., :3] = COLOR_CYCLE[i % 12]
  img[..., 3] = 255
  return img


def cpp_images_are_near(baseline, result):
  """rasterize_test_utils.cc:195-230: outlier channels (> 10/255) must not
  exceed 0.5% of the pixel count."""
  if baseline.ndim == 3 and baseline.shape[2] == 3:
    # lodepng decodes to RGBA with an opaque alpha.
    baseline = np.concatenate(
        [baseline, np.full(baseline.shape[:2] + (1,), 255, np.uint8)], 2)
  if baseline.shape != result.shape:
    return False, f"shapes {baseline.shape} vs {result.shape}"
  n = np.count_nonzero(np.abs(baseline.astype(int) - result.astype(int)) > 10)
  limit = int(0.005 * baseline.shape[0] * baseline.shape[1])
  return n <= limit, f"{n} outlier channels, max {limit}"


def load_flat_obj(name):
  """obj_loader.py:24-100 for the fixture .obj files: v / vt / vn / f
  lines; each distinct face-vertex name becomes one flattened vertex in
  order of first use; triangles are 0-based.  Returns (vertices [N, 3 + 2
  if vt + 3 if vn] float32, triangles [M, 3] int32)."""
  kinds = ("v", "vt", "vn")
  lists = {k: [] for k in kinds}
  flat, index, tris = [], {}, []
  used = {k: False for k in kinds}
  with open(os.path.join(FIXTURES, name)) as f:
    for line in f:
      tok = line.split()
      if not tok:
        continue
      if tok[0] in kinds:
        lists[tok[0]].append([float(x) for x in tok[1:]])
      elif tok[0] == "f":
        tri = []
        for name_ in tok[1:4]:
          if name_ not in index:
            parts = (name_.split("/") + ["", ""])[:3]
            vert = []
            for k, i in zip(kinds, parts):
              if i:
                vert += lists[k][int(i) - 1]
                used[k] = True
              else:
                vert += [0.0, 0.0] if k == "vt" else [0.0, 0.0, 0.0]
            index[name_] = len(flat)
            flat.append(vert)
          tri.append(index[name_])
        tris.append(tri)
  arr = np.asarray(flat, np.float32)
  out = [arr[:, :3]]
  if used["vt"]:
    out.append(arr[:, 3:5])
  if used["vn"]:
    out.append(arr[:, -3:])
  return np.concatenate(out, 1), np.asarray(tris, np.int32)


def reference_look_at(view):
  """test_utils.py:38-77: the test views' look-at matrices (float64)."""
  eye = {"view_1": (2.0, 3.0, 6.0), "view_2": (-3.0, 1.0, 6.0)}[view]
  return reference_look_at_eye(eye, (0.0, 1.0, 0.0))


def reference_look_at_eye(eye, world_up):
  """test_utils.py:55-77 for an eye position looking at the origin."""
  eye = np.array(eye)
  fwd = -eye / np.linalg.norm(eye)
  side = np.cross(fwd, world_up)
  side = side / np.linalg.norm(side)
  up = np.cross(side, fwd)
  rot = np.eye(4)
  rot[0, :3], rot[1, :3], rot[2, :3] = side, up, -fwd
  tr = np.eye(4)
  tr[:3, 3] = -eye
  return rot @ tr


def reference_perspective():
  """test_utils.py:80-96 with index 0: fov 40, 640/480, near 0.01, far 10."""
  f = 1.0 / np.tan(40.0 * (np.pi / 360.0))
  a = -(10.0 + 0.01) / (10.0 - 0.01)
  b = -2.0 * (10.0 * 0.01 / (10.0 - 0.01))
  return np.array([[f / (640 / 480), 0, 0, 0], [0, f, 0, 0], [0, 0, a, b],
                   [0, 0, -1, 0]])


# render_test.py:113-118: the disparity range Sphere_Disparity.png spans.
DISPARITY_MIN, DISPARITY_MAX = 0.12523349, 0.16625337
