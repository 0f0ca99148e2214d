"""The C render-splat (oracle/splat_cpu.c, the CPU baseline's splat leg)
against the torch restatement: its float32 forward, and its hand-derived
adjoint against float64 autograd of the push formulation."""
import numpy as np
import pytest
import torch

from diffren_amd import scenes
from oracle import raster_oracle
from oracle import splat_cpu
from oracle import splat_oracle as so


def _scene(seed, w, h, layers, channels):
  s = scenes.icosphere_scene(seed=seed, subdivisions=2)
  rng = np.random.default_rng(seed + 10)
  colors = rng.uniform(size=(s["vertices"].shape[0], channels)).astype(
      np.float32)
  clip = so.transform(s["camera"], s["vertices"], torch.float32).numpy()
  # No culling, so the deeper layers hold the sphere's back faces.
  ids, _, b = raster_oracle.rasterize(clip, s["triangles"], w, h, layers, 0)
  target = rng.uniform(size=(h, w, channels)).astype(np.float32)
  return s, colors, ids, b, target


@pytest.mark.parametrize("w,h,layers,channels", [(40, 32, 1, 4),
                                                 (33, 47, 3, 4),
                                                 (48, 40, 2, 1),
                                                 (36, 36, 3, 7)])
def test_forward_and_gradients_match_the_torch_restatement(w, h, layers,
                                                           channels):
  s, colors, ids, b, target = _scene(1, w, h, layers, channels)
  scale = 1.0 / target.size
  img, loss, dp, dc = splat_cpu.render(s["vertices"], colors, s["camera"],
                                       s["triangles"], ids, b, target, scale,
                                       threads=4)
  assert (ids >= 0).sum() > 100
  # Forward against the float32 restatement (differences only from the
  # exponential's last bit).
  out32 = so.render_splat(torch.as_tensor(s["vertices"]),
                          torch.as_tensor(colors), s["camera"], ids, b,
                          s["triangles"]).numpy()
  assert np.abs(img - out32).max() <= 2e-6 * max(np.abs(out32).max(), 1.0)
  # Loss and gradients against float64 autograd.
  v64 = torch.as_tensor(s["vertices"], dtype=torch.float64).requires_grad_()
  c64 = torch.as_tensor(colors, dtype=torch.float64).requires_grad_()
  out64 = so.render_splat(v64, c64, s["camera"], ids, b, s["triangles"])
  loss64 = ((out64 - torch.as_tensor(target, dtype=torch.float64)) ** 2
            ).sum() * scale
  loss64.backward()
  assert abs(loss - loss64.item()) <= 1e-5 * loss64.item()
  gv, gc = v64.grad.numpy(), c64.grad.numpy()
  assert np.linalg.norm(dp - gv) <= 1e-4 * np.linalg.norm(gv)
  assert np.linalg.norm(dc - gc) <= 1e-5 * np.linalg.norm(gc)


def test_thread_count_does_not_change_bits():
  s, colors, ids, b, target = _scene(2, 64, 48, 3, 4)
  r1 = splat_cpu.render(s["vertices"], colors, s["camera"], s["triangles"],
                        ids, b, target, 0.5, threads=1)
  r4 = splat_cpu.render(s["vertices"], colors, s["camera"], s["triangles"],
                        ids, b, target, 0.5, threads=4)
  for a, c in zip((r1[0], r1[2], r1[3]), (r4[0], r4[2], r4[3])):
    assert np.array_equal(a, c)


def test_forward_only_and_bad_sizes():
  s, colors, ids, b, _ = _scene(3, 20, 20, 1, 4)
  img, loss, dp, dc = splat_cpu.render(s["vertices"], colors, s["camera"],
                                       s["triangles"], ids, b)
  assert loss is None and dp is None and dc is None
  assert np.isfinite(img).all()
  wide = np.zeros((colors.shape[0], 33), np.float32)
  with pytest.raises(ValueError):
    splat_cpu.render(s["vertices"], wide, s["camera"], s["triangles"], ids, b)
