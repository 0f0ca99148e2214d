"""TEST INFRASTRUCTURE ONLY: CPU oracle for differentiable barycentrics.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may import
this module; nothing in diffren_amd/ does.

Restates diffren/jax/internal/rasterize.py:127-256 (differentiable_barycentrics,
normalized_pixel_coordinates, compute_triangle_matrix,
compute_barycentric_coordinates):
  * `barycentrics_f32` -- numpy float32, every expression in the jnp source's
    left-to-right order (no fused multiply-add), for exact comparison with the
    GPU kernel;
  * `barycentrics_torch` -- the same function on torch tensors in any dtype;
    in float64 its torch autograd is the gradient oracle ("autodiff of the
    restatement": jax is not installed, so the reference's jax.jacfwd cannot
    run).
Pins (tests/test_oracle_diffbary.py): the golden Jacobian of
rasterize_test.py:216-224 (64x48 image, pixel (32, 24), 6 decimals), the
reference PNG fixtures Simple_Triangle.png, Perspective_Corrected_Triangle.png
and Barycentrics_Cube.png at the reference's tolerance
(rasterize_test.py:67-98,150-165), the rasterizer's own fixed-point
barycentrics (an independent computation) on every covered pixel, the
invariants (coordinates sum to 1, a vertex has (1, 0, 0)), and float64
central finite differences.  The exact float32 output of the reference's XLA
fusion (its summation order) is parity unpinned below those tolerances.
"""
import numpy as np
import torch


def pixel_coordinates_f32(width, height):
  """rasterize.py:184-189: (2i + 1) / size - 1 in float32, [H, W] each."""
  x = (2 * np.arange(width) + 1).astype(np.float32) / np.float32(width)
  y = (2 * np.arange(height) + 1).astype(np.float32) / np.float32(height)
  x = x - np.float32(1)
  y = y - np.float32(1)
  return np.meshgrid(x, y)


def _corners(clip, triangles, ids):
  """Clip vertices (x, y, w) of each pixel-layer's triangle; id -1 wraps to
  the last triangle like the jnp gather (masked afterwards)."""
  t = np.asarray(ids).astype(np.int64)
  t = np.where(t < 0, t + len(triangles), t)
  vid = np.asarray(triangles)[t]                       # [L,H,W,3]
  c = np.asarray(clip)[vid]                            # [L,H,W,3,4]
  return c[..., 0], c[..., 1], c[..., 3]               # x, y, w [L,H,W,3]


def barycentrics_f32(clip, triangles, ids, width, height):
  """[L,H,W,3] float32 (rasterize.py:127-256 in jnp's evaluation order)."""
  f = np.float32
  clip = np.asarray(clip, f)
  ids = np.asarray(ids)
  x, y, w = _corners(clip, triangles, ids)
  x0, x1, x2 = x[..., 0], x[..., 1], x[..., 2]
  y0, y1, y2 = y[..., 0], y[..., 1], y[..., 2]
  w0, w1, w2 = w[..., 0], w[..., 1], w[..., 2]
  # xyw[r, k]: row r in (x, y, w), column k = vertex (rasterize.py:208-210)
  d11 = y1 * w2 - y2 * w1
  d21 = y2 * w0 - y0 * w2
  d31 = y0 * w1 - y1 * w0
  d12 = w1 * x2 - w2 * x1
  d22 = w2 * x0 - w0 * x2
  d32 = w0 * x1 - w1 * x0
  d13 = x1 * y2 - x2 * y1
  d23 = x2 * y0 - x0 * y2
  d33 = x0 * y1 - x1 * y0
  det = (x0 * d11 + y0 * d12) + w0 * d13
  sign = np.sign(det) + (det == 0).astype(f)
  px, py = pixel_coordinates_f32(width, height)
  rows = ((d11, d12, d13), (d21, d22, d23), (d31, d32, d33))
  with np.errstate(all="ignore"):
    r = [((sign * a) * px + (sign * b) * py) + sign * c for a, b, c in rows]
    s = (r[0] + r[1]) + r[2]
    b = np.stack([np.nan_to_num(rk / s) for rk in r], -1).astype(f)
  mask = (ids >= 0)[..., None]
  return np.where(mask, b, f(0))


def barycentrics_torch(clip, triangles, ids, width, height):
  """Differentiable [L,H,W,3] from clip [V,4] (torch, any float dtype)."""
  dt = clip.dtype
  t = torch.as_tensor(np.asarray(ids)).long()
  tris = torch.as_tensor(np.asarray(triangles)).long()
  vid = tris[torch.where(t < 0, t + tris.shape[0], t)]          # [L,H,W,3]
  c = clip[vid]                                                 # [L,H,W,3,4]
  x, y, w = c[..., 0], c[..., 1], c[..., 3]
  x0, x1, x2 = x[..., 0], x[..., 1], x[..., 2]
  y0, y1, y2 = y[..., 0], y[..., 1], y[..., 2]
  w0, w1, w2 = w[..., 0], w[..., 1], w[..., 2]
  d11 = y1 * w2 - y2 * w1
  d21 = y2 * w0 - y0 * w2
  d31 = y0 * w1 - y1 * w0
  d12 = w1 * x2 - w2 * x1
  d22 = w2 * x0 - w0 * x2
  d32 = w0 * x1 - w1 * x0
  d13 = x1 * y2 - x2 * y1
  d23 = x2 * y0 - x0 * y2
  d33 = x0 * y1 - x1 * y0
  det = (x0 * d11 + y0 * d12) + w0 * d13
  sign = (torch.sign(det) + (det == 0).to(dt)).detach()
  xs = (2 * torch.arange(width) + 1).to(dt) / width - 1
  ys = (2 * torch.arange(height) + 1).to(dt) / height - 1
  py, px = torch.meshgrid(ys, xs, indexing="ij")
  rows = ((d11, d12, d13), (d21, d22, d23), (d31, d32, d33))
  r = [((sign * a) * px + (sign * b) * py) + sign * cc for a, b, cc in rows]
  s = (r[0] + r[1]) + r[2]
  b = torch.stack([torch.nan_to_num(rk / s) for rk in r], -1)
  mask = (t >= 0).to(dt)[..., None]
  return b * mask


def barycentrics_torch_at(clip, triangles, ids, width, height, values):
  """`values` (the float32 coordinates, e.g. from `barycentrics_f32`) carrying
  the derivative of `barycentrics_torch(clip, ...)`: a gradient oracle at the
  same forward point as a float32 evaluation, so that comparing gradients
  does not also compare float32 and float64 forward values."""
  b = barycentrics_torch(clip, triangles, ids, width, height)
  return torch.as_tensor(np.asarray(values), dtype=b.dtype) + (b - b.detach())
