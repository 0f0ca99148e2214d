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
