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
Pins (tests/test_oracle_texture.py): the reference PNG fixture
Textured_Square.png (shaders_test.py:33-71, through the oracle render
path), closed forms (a texel centre returns the texel, a midpoint the mean
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
