"""Parity at the full sizes of the 1M-triangle workloads, one view each.

The 1,000,000-triangle Perlin sphere (500 x 1000 UV sphere, 500,002
vertices) at 1024x1024 with 3 layers and back-face culling, seen from the
first of the 8 seeded orbit cameras:
  * differentiation (RGBA colours, L2 loss against a target rendered from a
    copy with vertices + N(0, 0.01)): triangle ids and barycentrics
    bit-exact against the CPU rasterizer oracle, the fused and the plain
    image within 1e-5 of the float32 oracle restatement (the reference's
    float32 expression order, splat.py:103-104, 218-221), vertex and colour
    gradients within 1e-4 of the float64 oracle;
  * forward with normals + 8 features + alpha (C = 12): the same image bound
    for the fused and the plain path.
Relative error is max |a - b| / max |b|.
"""
import numpy as np
import pytest
import torch

import helpers
from diffren_amd import batch
from diffren_amd import mesh as mesh_lib
from diffren_amd import render
from diffren_amd import scenes
from oracle import raster_oracle
from oracle import splat_oracle as so

pytestmark = pytest.mark.gpu

R, L = 1024, 3


@pytest.fixture(scope="module")
def sphere():
  v, t = scenes.uv_sphere_mesh(500, 1000, seed=0)
  cam = scenes.orbit_cameras(8, seed=0)[0]
  clip32 = so.transform(cam, v, torch.float32).numpy()
  ids, _, b = raster_oracle.rasterize(clip32, t, R, R, L, 1)
  return dict(v=v, t=t, cam=cam, ids=ids, b=b)


def _check_raster(r, s):
  ids = r.ids[0].cpu().numpy()
  assert np.array_equal(ids, s["ids"])
  fg = s["ids"] >= 0
  # Sparse raster outputs: barycentrics are defined under triangles only.
  assert helpers.same_bits(r.barys[0].cpu().numpy()[fg], s["b"][fg])


def test_differentiation_view_parity(sphere, device):
  s = sphere
  rng = np.random.default_rng(1)
  V = s["v"].shape[0]
  colors = np.concatenate([rng.uniform(0, 1, (V, 3)), np.ones((V, 1))],
                          1).astype(np.float32)
  v_target = (s["v"] + rng.normal(0, 0.01, s["v"].shape)).astype(np.float32)
  tris = torch.as_tensor(s["t"], device=device)
  cam = torch.as_tensor(s["cam"], device=device)
  mesh = dict(vertices=torch.as_tensor(s["v"], device=device),
              triangles=tris, colors=torch.as_tensor(colors, device=device))
  tmesh = dict(vertices=torch.as_tensor(v_target, device=device),
               triangles=tris, colors=mesh["colors"])
  tr = batch.BatchedSplatRenderer([tmesh], [(0, cam, R, R)], L)
  tr.render()
  torch.cuda.synchronize()
  target = tr.images[0].clone()
  del tr
  scale = 1.0 / (R * R * 4)
  r = batch.BatchedSplatRenderer([mesh], [(0, cam, R, R)], L,
                                 loss_scale=scale)
  r.targets[0].copy_(target)
  r.step()
  torch.cuda.synchronize()
  _check_raster(r, s)
  fused = r.images[0].cpu()

  with torch.no_grad():
    plain = render.render_triangles(mesh["vertices"], {"c": mesh["colors"]},
                                    tris, cam, R, R, lambda a: a["c"],
                                    num_layers=L).cpu()
    o32 = so.render_splat(torch.as_tensor(s["v"]), torch.as_tensor(colors),
                          s["cam"], s["ids"], s["b"], s["t"])
  e_fused = helpers.rel_err(fused, o32)
  e_plain = helpers.rel_err(plain, o32)
  print(f"image vs float32 oracle: fused {e_fused:.2e}, plain {e_plain:.2e}")
  assert e_fused <= 1e-5
  assert e_plain <= 1e-5
  # Both GPU paths evaluate the same float32 expressions in the same order.
  assert torch.equal(fused, plain)

  v64 = torch.as_tensor(s["v"], dtype=torch.float64).requires_grad_(True)
  c64 = torch.as_tensor(colors, dtype=torch.float64).requires_grad_(True)
  o64 = so.render_splat(v64, c64, s["cam"], s["ids"], s["b"], s["t"])
  loss = ((o64 - target.cpu().double()) ** 2).sum() * scale
  loss.backward()
  e_v = helpers.rel_err(r.d_vertices[0].cpu(), v64.grad)
  e_c = helpers.rel_err(r.d_colors[0].cpu(), c64.grad)
  e_l = abs(r.loss.item() - loss.item()) / abs(loss.item())
  print(f"vs float64: d vertices {e_v:.2e}, d colours {e_c:.2e}, "
        f"loss {e_l:.2e}")
  assert e_v <= 1e-4
  assert e_c <= 1e-4
  assert e_l <= 1e-5


def test_forward_twelve_channels_parity(sphere, device):
  s = sphere
  vt = torch.as_tensor(s["v"], device=device)
  tris = torch.as_tensor(s["t"], device=device)
  normals = mesh_lib.compute_vertex_normals(vt, tris)
  rng = np.random.default_rng(2)
  feats = torch.as_tensor(
      rng.normal(size=(vt.shape[0], 8)).astype(np.float32), device=device)
  attrs = torch.cat([normals, feats, torch.ones_like(feats[:, :1])],
                    1).contiguous()
  cam = torch.as_tensor(s["cam"], device=device)
  r = batch.BatchedSplatRenderer(
      [dict(vertices=vt, triangles=tris, colors=attrs)], [(0, cam, R, R)], L)
  r.render()
  torch.cuda.synchronize()
  _check_raster(r, s)
  fused = r.images[0].cpu()
  with torch.no_grad():
    plain = render.render_triangles(vt, {"a": attrs}, tris, cam, R, R,
                                    lambda a: a["a"], num_layers=L).cpu()
    o32 = so.render_splat(torch.as_tensor(s["v"]), attrs.cpu(), s["cam"],
                          s["ids"], s["b"], s["t"])
  e_fused = helpers.rel_err(fused, o32)
  e_plain = helpers.rel_err(plain, o32)
  print(f"C = 12 image vs float32 oracle: fused {e_fused:.2e}, "
        f"plain {e_plain:.2e}")
  assert e_fused <= 1e-5
  assert e_plain <= 1e-5
  assert torch.equal(fused, plain)
