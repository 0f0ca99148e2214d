"""TEST INFRASTRUCTURE ONLY: CPU oracle for bilinear texture resampling.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may import
this module; nothing in diffren_amd/ does.

Restates image.bilinear_resample (diffren/jax/utils/image.py:21-118) and
shaders.texture_map (diffren/jax/utils/shaders.py:47-72) in torch CPU tensor
ops in the reference's own form: pad the image by one texel, take the 2x2
footprint at the clipped floor, weight it by the clipped fractional offsets.
In float32 it mirrors the reference arithmetic (each product and sum as the
source writes it); in float64 its autograd is the gradient oracle.  The clip
is `minimum(maximum(t, 0), 1)`, as jnp.clip is written, so its derivative at
a bound is 1/2 in torch as in JAX.
Also restates the shaders the reference's textured-object tests apply
after the lookup (shaders.py:75-173 diffuse_light / point_light,
mesh.py:22-52 vertex normals, transforms.py:26-37 l2_normalize) and the
texture decode (shaders.py:25-34).
Pins (tests/test_oracle_texture.py): the reference PNG fixtures
Textured_Square.png (shaders_test.py:33-71) and Spot_Textured*.png
(shaders_test.py:75-172), each through the oracle render path, closed forms (a texel centre returns the texel, a midpoint the mean
of the two texels, edge padding clamps, constant padding fades to 0 over
the outer half texel, a constant image resamples to itself), and float64
central finite differences.
"""
import numpy as np
import torch


def _clip01(t):
  zero = torch.zeros((), dtype=t.dtype)
  one = torch.ones((), dtype=t.dtype)
  return torch.minimum(torch.maximum(t, zero), one)


def bilinear_resample(image, locations, pad_mode="constant"):
  """image [H, W, C], locations [..., 2] (x, y) -> [..., C] (torch)."""
  img = torch.as_tensor(image)
  loc = torch.as_tensor(locations).to(img.dtype)
  H, W, C = img.shape
  if pad_mode == "constant":
    pad = torch.zeros((H + 2, W + 2, C), dtype=img.dtype)
    pad[1:-1, 1:-1] = img
  else:
    ii = torch.clamp(torch.arange(-1, H + 1), 0, H - 1)
    jj = torch.clamp(torch.arange(-1, W + 1), 0, W - 1)
    pad = img[ii][:, jj]
  flat = loc.reshape(-1, 2)
  # locations - 0.5, flip to (y, x), + 1.0 (image.py:59-64,111)
  l0 = (flat[:, 1] - 0.5) + 1.0
  l1 = (flat[:, 0] - 0.5) + 1.0
  f0 = torch.clamp(torch.floor(l0), 0, H).detach()
  f1 = torch.clamp(torch.floor(l1), 0, W).detach()
  c0 = _clip01(l0 - f0)
  c1 = _clip01(l1 - f1)
  w0 = 1.0 - c0
  w1 = 1.0 - c1
  i0, j0 = f0.long(), f1.long()
  p00 = pad[i0, j0]
  p10 = pad[i0 + 1, j0]
  p11 = pad[i0 + 1, j0 + 1]
  p01 = pad[i0, j0 + 1]
  ll = (p00 * w0[:, None]) * w1[:, None]
  ul = (p10 * c0[:, None]) * w1[:, None]
  ur = (p11 * c0[:, None]) * c1[:, None]
  lr = (p01 * w0[:, None]) * c1[:, None]
  out = ((ll + ul) + ur) + lr
  return out.reshape(tuple(loc.shape[:-1]) + (C,))


def texture_map(uv_image, texture):
  """shaders.py:47-72: edge-clamped lookup at uv * (W, H)."""
  tex = torch.as_tensor(texture)
  uv = torch.as_tensor(uv_image).to(tex.dtype)
  scale = torch.tensor([tex.shape[1], tex.shape[0]], dtype=tex.dtype)
  return bilinear_resample(tex, uv * scale, "edge")


def load_texture(path):
  """shaders.py:25-34: uint8 image / 255, rows flipped (lower-left origin)."""
  from PIL import Image
  with Image.open(path) as im:
    return np.ascontiguousarray(np.flipud(np.asarray(im).astype(np.float32) /
                                          255.0))


def l2_normalize(x, eps=1e-12):
  return x * torch.rsqrt((x * x).sum(-1, keepdim=True) + eps)


def vertex_normals(vertices, triangles):
  """mesh.py:22-52: area-weighted, normalised."""
  v = torch.as_tensor(vertices)
  t = torch.as_tensor(np.asarray(triangles)).long()
  c = v[t]
  n = torch.linalg.cross(c[:, 1] - c[:, 0], c[:, 2] - c[:, 0])
  acc = torch.zeros_like(v)
  for i in range(3):
    acc = acc + torch.zeros_like(v).index_add_(0, t[:, i], n)
  return l2_normalize(acc)


def diffuse_light(diffuse, normals, light_direction=(0.5, 0.5, 1.0), ka=0.5,
                  kd=0.5):
  """shaders.py:75-97."""
  d = l2_normalize(torch.tensor(light_direction, dtype=normals.dtype))
  ndl = torch.clamp((normals * d).sum(-1, keepdim=True), 0.0, 1.0)
  return diffuse * ka + diffuse * kd * ndl


def point_light(diffuse, normals, positions, light, eye=(0.0, 0.0, 0.0),
                ka=0.3, kd=1.4, ks=0.0, ns=0.0, ac=1.0, al=0.7, aq=1.8):
  """shaders.py:100-173 (one light)."""
  light = torch.tensor(light, dtype=positions.dtype)
  eye = torch.tensor(eye, dtype=positions.dtype)
  ld = l2_normalize(light - positions)
  vd = l2_normalize(eye - positions)
  h = l2_normalize(ld + vd)
  nol = torch.clamp((normals * ld).sum(-1, keepdim=True), 0.0, 1.0)
  noh = torch.clamp((normals * h).sum(-1, keepdim=True), 0.0, 1.0)
  dist = torch.linalg.vector_norm(light - positions, dim=-1, keepdim=True)
  decay = 1.0 / (ac + al * dist + aq * dist * dist)
  return diffuse * (ka + kd * nol + ks * noh ** ns) * decay
