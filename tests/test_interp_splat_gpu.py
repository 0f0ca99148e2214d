"""Interpolation and splat kernels (forward and hand-written backward) vs the
CPU oracle (oracle/splat_oracle.py).

Tolerances: forward images within 1e-5 relative (max-abs error over the
max-abs value); gradients within 1e-4 relative of the float64 oracle.
"""
import numpy as np
import pytest
import torch

import helpers
from diffren_amd import interpolate
from diffren_amd import rasterize
from diffren_amd import render
from diffren_amd import scenes
from diffren_amd import splat
from diffren_amd import transforms
from oracle import raster_oracle
from oracle import splat_oracle as so

pytestmark = pytest.mark.gpu

FWD_TOL = 1e-5
GRAD_TOL = 1e-4


def scene(seed, w, h, layers, subdivisions=3, cull=1):
  s = scenes.icosphere_scene(seed=seed, subdivisions=subdivisions)
  clip = scenes.to_clip(s["vertices"], s["camera"])
  ids, z, b = raster_oracle.rasterize(clip, s["triangles"], w, h, layers,
                                      cull)
  return s, clip, ids, z, b


def test_transform_matches_oracle(device):
  s = scenes.icosphere_scene(seed=0)
  v = torch.as_tensor(s["vertices"], device=device)
  m = torch.as_tensor(s["camera"], device=device)
  got = transforms.transform_homogeneous(m, v).cpu()
  want = so.transform(s["camera"], s["vertices"], torch.float32)
  assert torch.equal(got, want)  # same order, no contraction: bit-exact


@pytest.mark.parametrize("channels", [1, 3, 4, 11])
def test_interpolate_forward_and_backward(channels, device):
  s, clip, ids, z, b = scene(1, 96, 64, 3, cull=0)
  rng = np.random.default_rng(channels)
  attr = rng.normal(size=(clip.shape[0], channels)).astype(np.float32)
  fbuf = rasterize.rasterize_triangles(
      torch.as_tensor(clip, device=device),
      torch.as_tensor(s["triangles"], device=device), None, 96, 64, 3, 0)
  a = torch.as_tensor(attr, device=device, dtype=torch.float32)
  a.requires_grad_(True)
  bg = rng.normal(size=channels).astype(np.float32)
  out = interpolate.interpolate_vertex_attribute(a, fbuf, bg)
  want = so.interpolate(torch.as_tensor(attr), ids, b, s["triangles"], bg)
  assert helpers.rel_err(out.detach().cpu(), want) <= FWD_TOL
  g = rng.normal(size=out.shape).astype(np.float32)
  out.backward(torch.as_tensor(g, device=device))
  a64 = torch.as_tensor(attr, dtype=torch.float64).requires_grad_(True)
  w64 = so.interpolate(a64, ids, b, s["triangles"], bg)
  w64.backward(torch.as_tensor(g, dtype=torch.float64))
  assert helpers.rel_err(a.grad.cpu(), a64.grad) <= GRAD_TOL


@pytest.mark.parametrize("channels", [2, 8, 16, 64])
def test_interpolate_backward_ragged_frame(channels, device):
  """The backward flushes C lanes per vertex record when C divides 64; a
  37x23x3 frame leaves the last wave partly past the end."""
  s, clip, ids, z, b = scene(2, 37, 23, 3, cull=0)
  rng = np.random.default_rng(100 + channels)
  attr = rng.normal(size=(clip.shape[0], channels)).astype(np.float32)
  fbuf = rasterize.rasterize_triangles(
      torch.as_tensor(clip, device=device),
      torch.as_tensor(s["triangles"], device=device), None, 37, 23, 3, 0)
  a = torch.as_tensor(attr, device=device).requires_grad_(True)
  out = interpolate.interpolate_vertex_attribute(a, fbuf, None)
  g = rng.normal(size=out.shape).astype(np.float32)
  out.backward(torch.as_tensor(g, device=device))
  a64 = torch.as_tensor(attr, dtype=torch.float64).requires_grad_(True)
  w64 = so.interpolate(a64, ids, b, s["triangles"], None)
  w64.backward(torch.as_tensor(g, dtype=torch.float64))
  assert a64.grad.abs().sum() > 0
  assert helpers.rel_err(a.grad.cpu(), a64.grad) <= GRAD_TOL


def _splat_inputs(seed, w, h, layers, channels):
  s, clip, ids, z, b = scene(seed, w, h, layers)
  cbuf = so.interpolate(torch.as_tensor(clip), ids, b, s["triangles"],
                        (0, 0, 1, 1))
  xyz = so.viewport_positions(cbuf, w, h).numpy()
  rng = np.random.default_rng(seed)
  color = rng.uniform(0, 1, size=(layers, h, w, channels)).astype(np.float32)
  mask = (ids >= 0).astype(np.float32)[..., None]
  return xyz.astype(np.float32), color, mask


# 17 and 21: padded to the next instantiated count; 40 and 70: split into
# groups of 31 colour channels plus alpha.
@pytest.mark.parametrize("layers,channels", [(1, 4), (2, 4), (3, 4), (3, 12),
                                             (4, 2), (2, 17), (3, 21), (2, 40),
                                             (1, 70)])
def test_splat_forward(layers, channels, device):
  xyz, color, mask = _splat_inputs(layers + channels, 80, 72, layers,
                                   channels)
  out, acc, norm = splat.splat_at_pixel_centers(
      torch.as_tensor(xyz, device=device),
      torch.as_tensor(color, device=device),
      torch.as_tensor(mask, device=device))
  w_out, w_acc, w_norm = so.splat_at_pixel_centers(
      torch.as_tensor(xyz), torch.as_tensor(color), torch.as_tensor(mask))
  assert helpers.rel_err(out.cpu(), w_out) <= FWD_TOL
  assert helpers.rel_err(acc.cpu(), w_acc) <= FWD_TOL
  assert helpers.rel_err(norm.cpu(), w_norm) <= FWD_TOL


@pytest.mark.parametrize("layers,channels", [(1, 4), (3, 4), (3, 12),
                                             (2, 17), (2, 40), (1, 70)])
def test_splat_backward_vs_float64_autodiff(layers, channels, device):
  xyz, color, mask = _splat_inputs(7 * layers + channels, 48, 40, layers,
                                   channels)
  x = torch.as_tensor(xyz, device=device).requires_grad_(True)
  c = torch.as_tensor(color, device=device).requires_grad_(True)
  out, _, _ = splat.splat_at_pixel_centers(x, c,
                                           torch.as_tensor(mask,
                                                           device=device))
  g = np.random.default_rng(0).normal(size=out.shape).astype(np.float32)
  out.backward(torch.as_tensor(g, device=device))
  x64 = torch.as_tensor(xyz, dtype=torch.float64).requires_grad_(True)
  c64 = torch.as_tensor(color, dtype=torch.float64).requires_grad_(True)
  o64, _, _ = so.splat_at_pixel_centers(x64, c64,
                                        torch.as_tensor(mask,
                                                        dtype=torch.float64))
  o64.backward(torch.as_tensor(g, dtype=torch.float64))
  assert helpers.rel_err(c.grad.cpu(), c64.grad) <= GRAD_TOL
  assert helpers.rel_err(x.grad.cpu()[..., :2], x64.grad[..., :2]) <= GRAD_TOL


def _deep_stack_inputs(seed, w, h, layers, channels):
  """Pixel-layer samples with many layers: positions jittered around the
  pixel centres, z increasing front to back, each pixel's stack cut at a
  random depth (z 1 and mask 0 below it, as a rasterizer leaves them)."""
  rng = np.random.default_rng(seed)
  xs, ys = np.meshgrid(np.arange(w) + 0.5, np.arange(h) + 0.5)
  xyz = np.empty((layers, h, w, 3), np.float32)
  xyz[..., 0] = xs + rng.uniform(-0.3, 0.3, size=(layers, h, w))
  xyz[..., 1] = ys + rng.uniform(-0.3, 0.3, size=(layers, h, w))
  xyz[..., 2] = np.sort(rng.uniform(0.1, 0.9, size=(layers, h, w)), axis=0)
  depth = rng.integers(0, layers + 1, size=(h, w))
  mask = (np.arange(layers)[:, None, None] < depth[None]).astype(np.float32)
  xyz[..., 2] = np.where(mask > 0, xyz[..., 2], 1.0)
  color = rng.uniform(0, 1, size=(layers, h, w, channels)).astype(np.float32)
  return xyz, color, mask[..., None]


# 8 layers keep the depths in registers; 9 and 12 take the kernels that read
# them from memory (any number of layers, as splat.py:26-175 takes).
@pytest.mark.parametrize("layers,channels", [(8, 4), (9, 3), (12, 4)])
def test_splat_deep_stacks(layers, channels, device):
  xyz, color, mask = _deep_stack_inputs(layers, 26, 19, layers, channels)
  x = torch.as_tensor(xyz, device=device).requires_grad_(True)
  c = torch.as_tensor(color, device=device).requires_grad_(True)
  out, acc, norm = splat.splat_at_pixel_centers(
      x, c, torch.as_tensor(mask, device=device))
  g = np.random.default_rng(1).normal(size=out.shape).astype(np.float32)
  out.backward(torch.as_tensor(g, device=device))
  w_out, w_acc, w_norm = so.splat_at_pixel_centers(
      torch.as_tensor(xyz), torch.as_tensor(color), torch.as_tensor(mask))
  assert helpers.rel_err(out.detach().cpu(), w_out) <= FWD_TOL
  assert helpers.rel_err(acc.cpu(), w_acc) <= FWD_TOL
  assert helpers.rel_err(norm.cpu(), w_norm) <= FWD_TOL
  x64 = torch.as_tensor(xyz, dtype=torch.float64).requires_grad_(True)
  c64 = torch.as_tensor(color, dtype=torch.float64).requires_grad_(True)
  o64, a64, _ = so.splat_at_pixel_centers(
      x64, c64, torch.as_tensor(mask, dtype=torch.float64))
  # The deeper layers reach every class: fg, surface and bg all get weight.
  assert (a64[..., -1] > 0).flatten(1).any(1).all()
  o64.backward(torch.as_tensor(g, dtype=torch.float64))
  assert helpers.rel_err(c.grad.cpu(), c64.grad) <= GRAD_TOL
  assert helpers.rel_err(x.grad.cpu()[..., :2], x64.grad[..., :2]) <= GRAD_TOL


def test_render_more_layers_than_the_fused_kernels(device):
  """render_triangles with 10 layers: the splat takes the unfused
  composition (the fused kernels stop at 8), the result matches the
  oracle."""
  s = scenes.icosphere_scene(seed=4)
  w, h, L = 48, 40, 10
  img, v, col = _render_gpu(s, w, h, L, device)
  img.sum().backward()
  clip = so.transform(s["camera"], s["vertices"], torch.float32).numpy()
  ids, _, b = raster_oracle.rasterize(clip, s["triangles"], w, h, L, 1)
  v64 = torch.as_tensor(s["vertices"], dtype=torch.float64).requires_grad_()
  c64 = torch.as_tensor(s["colors"], dtype=torch.float64).requires_grad_()
  o64 = so.render_splat(v64, c64, s["camera"], ids, b, s["triangles"])
  o64.sum().backward()
  assert helpers.rel_err(img.detach().cpu(), o64.detach()) <= FWD_TOL
  assert helpers.rel_err(v.grad.cpu(), v64.grad) <= GRAD_TOL
  assert helpers.rel_err(col.grad.cpu(), c64.grad) <= GRAD_TOL


def _render_gpu(s, w, h, layers, device, colors=None):
  v = torch.as_tensor(s["vertices"], device=device).requires_grad_(True)
  col = torch.as_tensor(s["colors"] if colors is None else colors,
                        device=device).requires_grad_(True)
  cam = torch.as_tensor(s["camera"], device=device)
  tris = torch.as_tensor(s["triangles"], device=device)
  img = render.render_triangles(v, {"c": col}, tris, cam, w, h,
                                lambda a: a["c"], num_layers=layers)
  return img, v, col


@pytest.mark.parametrize("layers", [1, 3])
def test_render_splat_forward_and_gradients(layers, device):
  w = h = 128
  s = scenes.icosphere_scene(seed=4)
  img, v, col = _render_gpu(s, w, h, layers, device)
  target = torch.rand_like(img)
  loss = ((img - target) ** 2).mean()
  loss.backward()
  # Oracle: rasterize on the CPU from the same clip-space vertices the GPU
  # rasterized (the kernel transform is bit-exact to so.transform float32).
  clip32 = so.transform(s["camera"], s["vertices"], torch.float32).numpy()
  ids, z, b = raster_oracle.rasterize(clip32, s["triangles"], w, h, layers, 1)
  v64 = torch.as_tensor(s["vertices"], dtype=torch.float64).requires_grad_(
      True)
  c64 = torch.as_tensor(s["colors"], dtype=torch.float64).requires_grad_(True)
  o64 = so.render_splat(v64, c64, s["camera"], ids, b, s["triangles"])
  assert helpers.rel_err(img.detach().cpu(), o64.detach()) <= FWD_TOL
  l64 = ((o64 - target.cpu().double()) ** 2).mean()
  l64.backward()
  assert helpers.rel_err(col.grad.cpu(), c64.grad) <= GRAD_TOL
  assert helpers.rel_err(v.grad.cpu(), v64.grad) <= GRAD_TOL


def test_recompute_rasterization(device):
  """render_triangles(recompute_rasterization=True): the backward
  rasterizes again instead of keeping the framebuffer; image and gradients
  are the stored path's, and less memory stays allocated between forward
  and backward."""
  w, h, L = 512, 384, 3
  s = scenes.icosphere_scene(seed=6, subdivisions=4)
  out = []
  for recompute in (False, True):
    v = torch.as_tensor(s["vertices"], device=device).requires_grad_(True)
    col = torch.as_tensor(s["colors"], device=device).requires_grad_(True)
    cam = torch.as_tensor(s["camera"], device=device)
    tris = torch.as_tensor(s["triangles"], device=device)
    torch.cuda.synchronize()
    base = torch.cuda.memory_allocated()
    img = render.render_triangles(v, {"c": col}, tris, cam, w, h,
                                  lambda a: a["c"], num_layers=L,
                                  recompute_rasterization=recompute)
    held = torch.cuda.memory_allocated() - base
    loss = (img * img).sum()
    loss.backward()
    torch.cuda.synchronize()
    out.append((img.detach(), v.grad, col.grad, held))
  (i0, v0, c0, m0), (i1, v1, c1, m1) = out
  print(f"memory held between forward and backward: stored {m0 / 2**20:.1f} "
        f"MiB, recomputed {m1 / 2**20:.1f} MiB")
  assert torch.equal(i0, i1)
  assert helpers.rel_err(v1.cpu(), v0.cpu()) <= 1e-5
  assert helpers.rel_err(c1.cpu(), c0.cpu()) <= 1e-5
  assert m1 < m0 / 4


@pytest.mark.parametrize("num_components", [4, 2])
def test_optimize_single_triangle(num_components, device):
  """splat_test.py:84-168: gradient descent on a triangle's clip-space
  vertices toward a translated copy.  Plain rasterization gives a zero
  vertex gradient; rasterize-then-splat gives a useful one, and 15 steps of
  unit-rate descent halve the splat loss."""
  from diffren_amd import constants
  w = h = 32
  init = torch.tensor([[0, 0, 0, 1], [0.5, 0, 0, 1], [0.5, 0.5, 0, 1]],
                      dtype=torch.float32, device=device)
  target_v = torch.tensor([[-0.25, 0, 0, 1], [0.25, 0, 0, 1],
                           [0.25, 0.5, 0, 1]], dtype=torch.float32,
                          device=device)
  tris = torch.tensor([[0, 1, 2]], dtype=torch.int32, device=device)
  colors = torch.tensor([[0, 0., 0.8, 1.0], [0, 0.8, 0, 1.0],
                         [0, 0., 0.8, 1.0]], dtype=torch.float32,
                        device=device)[:, -num_components:]

  def draw(v, do_splat):
    mode = (constants.CompositingMode.SPLAT_OVER if do_splat else
            constants.CompositingMode.OVER)
    return render.render_triangles(v, {"colors": colors}, tris, None, w, h,
                                   lambda a: a["colors"],
                                   compositing_mode=mode)

  target = draw(target_v, False).detach()

  def grads(v):
    out = []
    for do_splat in (False, True):
      vv = v.clone().requires_grad_(True)
      loss = ((draw(vv, do_splat) - target) ** 2).mean()
      # Rasterization is not differentiable: the plain image does not
      # depend on the vertices through any differentiable path.
      if loss.requires_grad:
        loss.backward()
      g = vv.grad if vv.grad is not None else torch.zeros_like(vv)
      out.append((loss.item(), g))
    return out

  v = init.clone()
  initial = None
  for i in range(15):
    (_, g_raster), (splat_loss, g_splat) = grads(v)
    if i == 0:
      assert torch.linalg.norm(g_raster).item() == 0.0
      assert torch.linalg.norm(g_splat).item() > 0.01
      initial = splat_loss
    v = v - g_splat
  final = ((draw(v, True) - target) ** 2).mean().item()
  print(f"splat loss {initial:.5f} -> {final:.5f}")
  assert final < 0.5 * initial


def test_spike_derivative(device):
  """splat_test.py:170-224 on the GPU path.  The hand-written backward is
  checked against the oracle's float64 forward-mode derivative J 1 (d image
  / d spike offset) through the adjoint identity <u, J 1> = (J^T u) . 1 for
  seeded u; the derivative image from central differences of the GPU
  forward pass (rasterization frozen, as stop_gradient freezes it) matches
  the oracle's and Spike_Derivative.png at pixel error threshold 0."""
  from diffren_amd import constants
  from test_oracle_splat import spike_scene, spike_visualisation
  v, t, proj, ids, b = spike_scene()
  base = torch.as_tensor(v, dtype=torch.float64)
  cols64 = torch.ones((v.shape[0], 4), dtype=torch.float64)
  p64 = torch.as_tensor(proj, dtype=torch.float64)

  def f64(off):
    verts = torch.cat([base[:32], base[32:33] * off, base[33:]], 0)
    return so.render_splat(verts, cols64, p64, ids, b, t)

  img64, jac64 = torch.autograd.functional.jvp(
      f64, (torch.tensor([2.0], dtype=torch.float64),),
      (torch.ones(1, dtype=torch.float64),))

  vt = torch.as_tensor(v, device=device)
  tt = torch.as_tensor(t, device=device)
  pt = torch.as_tensor(proj, device=device)
  cols = torch.ones((v.shape[0], 4), dtype=torch.float32, device=device)

  def spiked(off):
    return torch.cat([vt[:32], vt[32:33] * off, vt[33:]], 0)

  off = torch.tensor([2.0], device=device, requires_grad=True)
  img = render.render_triangles(spiked(off), {"c": cols}, tt, pt, 640, 480,
                                lambda a: a["c"])
  # Forward: against the float32 oracle (same arithmetic type).
  sp32 = torch.as_tensor(v).clone()
  sp32[32] *= 2.0
  img32 = so.render_splat(sp32, cols64.float(), torch.as_tensor(proj), ids, b,
                          t)
  assert helpers.rel_err(img.detach().cpu(), img32) <= FWD_TOL
  rng = np.random.default_rng(11)
  for k in range(3):
    u = rng.normal(size=tuple(img.shape))
    off.grad = None
    (img * torch.as_tensor(u, dtype=torch.float32, device=device)).sum(
    ).backward(retain_graph=True)
    want = float((torch.as_tensor(u) * jac64).sum())
    scale = float((torch.as_tensor(u).abs() * jac64.abs()).sum())
    assert abs(off.grad.item() - want) <= GRAD_TOL * scale, (k, want)

  fb = rasterize.rasterize_triangles(spiked(torch.tensor([2.0],
                                                         device=device)),
                                     tt, pt, 640, 480,
                                     face_culling_mode=constants.
                                     FaceCullingMode.BACK)
  shaded = interpolate.interpolate_vertex_attributes({"c": cols}, fb,
                                                     (0.0,))["c"]
  # Taps move between pixels as the offset changes (the splat is piecewise
  # smooth): a step of 3e-4 stays within one piece for nearly every pixel.
  h = 3e-4

  def f32(o):
    with torch.no_grad():
      return splat.splat_shaded_pixels(
          spiked(torch.tensor([o], device=device)), pt, fb, shaded).double()

  fd = ((f32(2.0 + h) - f32(2.0 - h)) / (2 * h)).cpu()
  assert helpers.rel_err(fd, jac64) <= 1e-2
  ok, msg = helpers.check_image(
      spike_visualisation(img.detach().cpu().double(), fd).numpy(),
      "Spike_Derivative.png", exact=True)
  assert ok, msg


@pytest.mark.parametrize("batched", [False, True])
@pytest.mark.parametrize("channels", [3, 4])
def test_transform_homogeneous_shapes(batched, channels, device):
  # transforms_test.py:31-49, plus values against the oracle.
  rng = np.random.default_rng(channels)
  shape = (3, 10, channels) if batched else (10, channels)
  v = rng.normal(size=shape).astype(np.float32)
  m = rng.normal(size=((3, 4, 4) if batched else (4, 4))).astype(np.float32)
  got = transforms.transform_homogeneous(torch.as_tensor(m, device=device),
                                         torch.as_tensor(v, device=device))
  assert tuple(got.shape) == shape[:-1] + (4,)
  for i in range(3 if batched else 1):
    mi, vi, gi = (m[i], v[i], got[i]) if batched else (m, v, got)
    want = so.transform(mi, vi, torch.float32).numpy()
    assert helpers.same_bits(gi.cpu().numpy(), want)
  with pytest.raises(ValueError):
    transforms.transform_homogeneous(torch.as_tensor(m, device=device),
                                     torch.as_tensor(v[..., :2],
                                                     device=device))
