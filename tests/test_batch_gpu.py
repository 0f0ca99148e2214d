"""Batched fused render (diffren_amd.batch) vs the unfused per-view path and
the float64 oracle; HIP-graph replays vs eager."""
import numpy as np
import pytest
import torch

import helpers
from diffren_amd import batch
from diffren_amd import render
from diffren_amd import scenes
from oracle import raster_oracle
from oracle import splat_oracle as so

pytestmark = pytest.mark.gpu


def _meshes(device):
  out = []
  for seed, sub in ((0, 3), (5, 2)):
    s = scenes.icosphere_scene(seed=seed, subdivisions=sub)
    out.append(dict(
        vertices=torch.as_tensor(s["vertices"], device=device),
        triangles=torch.as_tensor(s["triangles"], device=device),
        colors=torch.as_tensor(s["colors"], device=device),
        camera=torch.as_tensor(s["camera"], device=device), scene=s))
  return out


def _jobs(meshes):
  rot = torch.tensor(scenes.random_rotation(np.random.default_rng(9)),
                     dtype=torch.float32, device=meshes[0]["camera"].device)
  view = torch.eye(4, device=rot.device)
  view[:3, :3] = rot
  # Views of a mesh centred at z=-3: rotate about the mesh centre.
  shift = torch.eye(4, device=rot.device)
  shift[2, 3] = -3.0
  unshift = torch.eye(4, device=rot.device)
  unshift[2, 3] = 3.0
  cam2 = meshes[0]["camera"] @ shift @ view @ unshift
  return [(0, meshes[0]["camera"], 96, 64), (0, cam2.contiguous(), 128, 128),
          (1, meshes[1]["camera"], 40, 72)]


def _unfused(meshes, jobs, L, targets):
  """Per-view render_triangles + autograd (the plain path)."""
  vs = [m["vertices"].clone().requires_grad_(True) for m in meshes]
  cs = [m["colors"].clone().requires_grad_(True) for m in meshes]
  total = 0.0
  imgs = []
  npix = sum(w * h * meshes[mi]["colors"].shape[1] for mi, _, w, h in jobs)
  for (mi, cam, w, h), tgt in zip(jobs, targets):
    img = render.render_triangles(vs[mi], {"c": cs[mi]},
                                  meshes[mi]["triangles"], cam, w, h,
                                  lambda a: a["c"], num_layers=L)
    imgs.append(img)
    total = total + ((img - tgt) ** 2).sum() / npix
  total.backward()
  return imgs, total, [v.grad for v in vs], [c.grad for c in cs]


@pytest.mark.parametrize("L", [1, 3])
def test_fused_matches_unfused_and_oracle(L, device):
  meshes = _meshes(device)
  jobs = _jobs(meshes)
  r = batch.BatchedSplatRenderer(meshes, jobs, L)
  gen = torch.Generator(device="cpu").manual_seed(L)
  for t in r.targets:
    t.copy_(torch.rand(t.shape, generator=gen))
  r.step()
  torch.cuda.synchronize()
  imgs, loss, dv, dc = _unfused(meshes, jobs, L, r.targets)
  for a, b in zip(r.images, imgs):
    assert helpers.rel_err(a.cpu(), b.detach().cpu()) <= 1e-5
  assert abs(r.loss.item() - loss.item()) <= 1e-5 * abs(loss.item())
  for a, b in zip(r.d_vertices + r.d_colors, dv + dc):
    assert helpers.rel_err(a.cpu(), b.cpu()) <= 1e-4
  # Float64 oracle of the same loss, rasterization from the oracle.
  v64 = [torch.as_tensor(m["scene"]["vertices"],
                         dtype=torch.float64).requires_grad_(True)
         for m in meshes]
  c64 = [torch.as_tensor(m["scene"]["colors"],
                         dtype=torch.float64).requires_grad_(True)
         for m in meshes]
  total = 0.0
  for i, ((mi, cam, w, h), tgt) in enumerate(zip(jobs, r.targets)):
    s = meshes[mi]["scene"]
    cam_np = cam.cpu().numpy()
    clip32 = so.transform(cam_np, s["vertices"], torch.float32).numpy()
    ids, _, b = raster_oracle.rasterize(clip32, s["triangles"], w, h, L, 1)
    # The rasterizer is bit-exact, so the GPU's frame is the oracle's.
    assert np.array_equal(ids, r.ids[i].cpu().numpy())
    o = so.render_splat(v64[mi], c64[mi], cam_np, ids, b, s["triangles"])
    total = total + ((o - tgt.cpu().double()) ** 2).sum() * r.loss_scale
  total.backward()
  assert abs(r.loss.item() - total.item()) <= 1e-5 * abs(total.item())
  for a, b in zip(r.d_vertices + r.d_colors,
                  [v.grad for v in v64] + [c.grad for c in c64]):
    assert helpers.rel_err(a.cpu(), b) <= 1e-4


@pytest.mark.parametrize("streams", [1, 2, 4])
def test_graph_replays_match_eager(streams, device):
  meshes = _meshes(device)
  jobs = _jobs(meshes)
  r = batch.BatchedSplatRenderer(meshes, jobs, 3, streams=streams)
  assert len(r.groups) == min(streams, len(jobs))
  for t in r.targets:
    t.uniform_()
  r.step()
  torch.cuda.synchronize()
  want = [t.clone() for t in r.d_vertices + r.d_colors + r.images]
  want_loss = r.loss.item()
  r.capture()
  for k in range(4):
    r.replay()
    torch.cuda.synchronize()
    got = r.d_vertices + r.d_colors + r.images
    for a, b in zip(got, want):
      assert helpers.rel_err(a.cpu(), b.cpu()) <= 1e-5, k
    assert abs(r.loss.item() - want_loss) <= 1e-5 * abs(want_loss)
    # Eager work between replays.
    torch.ones(1000, device=device).sum().item()


def test_autotuned_group_count(device, tmp_path):
  """streams="auto": the step is timed with 1, 2 and 4 job groups (as many
  as there are jobs), the fastest count is kept (and cached), then the tile
  pass's block size likewise; the step's results equal a renderer built with
  that count and the other block size."""
  meshes = _meshes(device)
  jobs = _jobs(meshes)
  cache = str(tmp_path / "tune.json")
  r = batch.BatchedSplatRenderer(meshes, jobs, 3, streams="auto",
                                 autotune_cache=cache)
  # Three jobs: one, two (four groups would exceed the job count).
  assert set(r.streams_timings) == {"1", "2"}
  best = min(r.streams_timings, key=r.streams_timings.get)
  assert len(r.groups) == int(best)
  # Then the tile pass's block size, timed on the chosen step.
  assert set(r.tile_threads_timings) == {"128", "256"}
  threads = int(min(r.tile_threads_timings, key=r.tile_threads_timings.get))
  assert r.raster_config.tile_threads == threads
  f = batch.BatchedSplatRenderer(meshes, jobs, 3, streams=int(best),
                                 tile_threads=384 - threads)
  for ta, tb in zip(r.targets, f.targets):
    ta.uniform_()
    tb.copy_(ta)
  r.step()
  f.step()
  r.capture()
  r.replay()
  torch.cuda.synchronize()
  for a, b in zip(r.images + r.d_vertices + r.d_colors,
                  f.images + f.d_vertices + f.d_colors):
    assert helpers.rel_err(a.cpu(), b.cpu()) <= 1e-5
  assert abs(r.loss.item() - f.loss.item()) <= 1e-6 * abs(f.loss.item())
  # A second renderer of the same batch takes the cached choice.
  again = batch.BatchedSplatRenderer(meshes, jobs, 3, streams="auto",
                                     autotune_cache=cache)
  assert len(again.groups) == len(r.groups)
  assert again.streams_timings == r.streams_timings
  assert again.tile_threads_timings == r.tile_threads_timings
  assert again.raster_config.tile_threads == threads
  with pytest.raises(ValueError):
    batch.BatchedSplatRenderer(meshes, jobs, 3, tile_threads=64)


@pytest.mark.parametrize("streams", [1, 4])
def test_fused_transform_matches_clip_buffers(streams, device):
  """Vertex transform on load (no clip-space buffers) vs the transform
  launch: rasterization and images bit-identical, gradients to float
  summation order; graph replays too."""
  meshes = _meshes(device)
  jobs = _jobs(meshes)
  a = batch.BatchedSplatRenderer(meshes, jobs, 3, streams=streams)
  b = batch.BatchedSplatRenderer(meshes, jobs, 3, streams=streams,
                                 fused_transform=True)
  assert not b.clip_blocks
  for ta, tb in zip(a.targets, b.targets):
    ta.uniform_()
    tb.copy_(ta)
  a.step()
  b.step()
  torch.cuda.synchronize()
  for x, y in zip(a.ids + a.images, b.ids + b.images):
    assert torch.equal(x, y)
  for ia, za, ba, zb, bb in zip(a.ids, a.z, a.barys, b.z, b.barys):
    hit = ia >= 0
    assert helpers.same_bits(za[hit].cpu().numpy(), zb[hit].cpu().numpy())
    assert helpers.same_bits(ba[hit].cpu().numpy(), bb[hit].cpu().numpy())
  want = [t.clone() for t in a.d_vertices + a.d_colors]
  for x, y in zip(want, b.d_vertices + b.d_colors):
    assert helpers.rel_err(y.cpu(), x.cpu()) <= 1e-5
  assert abs(a.loss.item() - b.loss.item()) <= 1e-6 * abs(a.loss.item())
  b.capture()
  for k in range(3):
    b.replay()
    torch.cuda.synchronize()
    for x, y in zip(want, b.d_vertices + b.d_colors):
      assert helpers.rel_err(y.cpu(), x.cpu()) <= 1e-5, k
    for x, y in zip(a.images, b.images):
      assert torch.equal(x, y), k
    torch.ones(1000, device=device).sum().item()


@pytest.mark.parametrize("channels", [1, 9, 12, 15, 28])
def test_fused_channel_counts(channels, device):
  """The fused path at other channel counts (12: normals + 8 features +
  alpha, the forward-only workload's pass-through shader; 15: the
  interpolated channels of that workload with the clip xyzw; 28: the
  largest) vs the unfused path, forward and backward; forward-only graph
  replays too."""
  meshes = _meshes(device)
  gen = torch.Generator(device="cpu").manual_seed(channels)
  for m in meshes:
    V = m["vertices"].shape[0]
    m["colors"] = torch.rand((V, channels), generator=gen).to(device)
  jobs = _jobs(meshes)
  r = batch.BatchedSplatRenderer(meshes, jobs, 3)
  for t in r.targets:
    t.copy_(torch.rand(t.shape, generator=gen))
  r.step()
  torch.cuda.synchronize()
  imgs, loss, dv, dc = _unfused(meshes, jobs, 3, r.targets)
  for a, b in zip(r.images, imgs):
    assert helpers.rel_err(a.cpu(), b.detach().cpu()) <= 1e-5
  assert abs(r.loss.item() - loss.item()) <= 1e-5 * abs(loss.item())
  for a, b in zip(r.d_vertices + r.d_colors, dv + dc):
    assert helpers.rel_err(a.cpu(), b.cpu()) <= 1e-4
  want = [i.clone() for i in r.images]
  for i in r.images:
    i.zero_()
  r.capture(forward_only=True)
  r.replay()
  torch.cuda.synchronize()
  for a, b in zip(r.images, want):
    assert torch.equal(a, b)


def test_uneven_batch_autotuned_matches_per_view(device):
  """The batching configuration at reduced scale: meshes of uneven triangle
  counts and resolutions in one batch, autotuned tiles, vs per-view renders
  of the plain path."""
  from diffren_amd import scenes as sc
  work = sc.batching_workload(num_meshes=6, t_min=500, t_max=20_000, seed=3,
                              resolutions=(64, 96, 128, 200))
  meshes = [dict(vertices=torch.as_tensor(m["vertices"], device=device),
                 triangles=torch.as_tensor(m["triangles"], device=device),
                 colors=torch.as_tensor(m["colors"], device=device))
            for m in work]
  jobs = [(i, torch.as_tensor(m["camera"], device=device), m["resolution"],
           m["resolution"]) for i, m in enumerate(work)]
  r = batch.BatchedSplatRenderer(meshes, jobs, 3, tile="auto")
  assert r.autotune_timings
  for t in r.targets:
    t.uniform_()
  r.step()
  torch.cuda.synchronize()
  imgs, loss, dv, dc = _unfused(meshes, jobs, 3, r.targets)
  for a, b in zip(r.images, imgs):
    assert helpers.rel_err(a.cpu(), b.detach().cpu()) <= 1e-5
  assert abs(r.loss.item() - loss.item()) <= 1e-5 * abs(loss.item())
  for a, b in zip(r.d_vertices + r.d_colors, dv + dc):
    assert helpers.rel_err(a.cpu(), b.cpu()) <= 1e-4
  need, overflow = r.bin_stats()
  assert not overflow and need > 0


def test_large_frames_match_per_view(device):
  """Frames large enough that the raster, target and source kernels take
  their tiles in XCD-ordered runs over many full groups plus a ragged tail
  (1000x700 and 648x520: 2772 + 1353 raster tiles of 16x16, 2800 + 1430
  splat tiles of 64x4), vs per-view renders of the plain path."""
  s = scenes.icosphere_scene(seed=4, subdivisions=4)
  mesh = dict(vertices=torch.as_tensor(s["vertices"], device=device),
              triangles=torch.as_tensor(s["triangles"], device=device),
              colors=torch.as_tensor(s["colors"], device=device))
  cam = torch.as_tensor(s["camera"], device=device)
  jobs = [(0, cam, 1000, 700), (0, cam, 648, 520)]
  r = batch.BatchedSplatRenderer([mesh], jobs, 3, tile=(16, 16), streams=2)
  for t in r.targets:
    t.uniform_()
  r.step()
  torch.cuda.synchronize()
  clip = so.transform(s["camera"], s["vertices"], torch.float32).numpy()
  for (mi, c, w, h), ids in zip(jobs, r.ids):
    want, _, _ = raster_oracle.rasterize(clip, s["triangles"], w, h, 3, 1)
    assert np.array_equal(ids.cpu().numpy(), want)
  imgs, loss, dv, dc = _unfused([mesh], jobs, 3, r.targets)
  # Both paths form the splat offsets floor(x) + 0.5 - x from float32
  # viewport coordinates in the reference's order (splat.py:103-104,
  # 218-221) and splat in the same order: the images are identical.
  for a, b in zip(r.images, imgs):
    assert torch.equal(a, b.detach())
  assert abs(r.loss.item() - loss.item()) <= 1e-5 * abs(loss.item())
  # Gradients against the float64 oracle.  The fused backward evaluates the
  # tap derivatives at the exact splat offset (compensated product); the
  # plain backward, like the reference's float32 autodiff, at the offset
  # formed from the rounded float32 viewport x (up to 6e-5 px off at
  # x ~ 1000), so it is held to 3e-4 here.
  v64 = torch.as_tensor(s["vertices"], dtype=torch.float64).requires_grad_()
  c64 = torch.as_tensor(s["colors"], dtype=torch.float64).requires_grad_()
  total = 0.0
  for (mi, c, w, h), tgt in zip(jobs, r.targets):
    ids, _, b = raster_oracle.rasterize(clip, s["triangles"], w, h, 3, 1)
    o = so.render_splat(v64, c64, s["camera"], ids, b, s["triangles"])
    total = total + ((o - tgt.cpu().double()) ** 2).sum() * r.loss_scale
  total.backward()
  for a, p, want in ((r.d_vertices[0], dv[0], v64.grad),
                     (r.d_colors[0], dc[0], c64.grad)):
    assert helpers.rel_err(a.cpu(), want) <= 1e-4
    assert helpers.rel_err(p.cpu(), want) <= 3e-4


def test_vertex_gradient_matches_central_differences(device):
  # Central differences of the fused path's own loss (rasterization frozen)
  # for 64 seeded vertices x 3 coordinates, at eps 1e-3 (0.055 px of
  # fragment motion on this 192x192 frame) and 3e-4.  Entries whose renders
  # move a pixel's accumulated weight across the normalisation kink at 1 are
  # counted and held out of the kink-free error.
  import bench
  v, t = scenes.uv_sphere_mesh(60, 120, seed=2)
  rng = np.random.default_rng(4)
  col = np.concatenate([rng.uniform(0, 1, (v.shape[0], 3)),
                        np.ones((v.shape[0], 1))], 1).astype(np.float32)
  mesh = dict(vertices=torch.as_tensor(v, device=device),
              triangles=torch.as_tensor(t, device=device),
              colors=torch.as_tensor(col, device=device))
  cam = torch.as_tensor(scenes.orbit_cameras(1, seed=5)[0], device=device)
  R = 192
  r = batch.BatchedSplatRenderer([mesh], [(0, cam, R, R)], 3,
                                 loss_scale=1.0 / (R * R * 4),
                                 return_accum_buffers=True)
  target = torch.as_tensor(rng.uniform(size=(R, R, 4)).astype(np.float32),
                           device=device)
  r.targets[0].copy_(target)
  r.step()
  torch.cuda.synchronize()
  for eps, tol in ((1e-3, 3e-3), (3e-4, 3e-3)):
    fd = bench.finite_differences(r, mesh["vertices"], target,
                                  1.0 / (R * R * 4), eps=eps)
    print(f"eps {eps}: {fd}")
    assert fd["crossing_kink"] < fd["entries"] // 4
    assert fd["rel_err_kink_free"] <= tol


def test_fused_accum_buffers_match_plain(device):
  """return_accum_buffers on the fused path: the accumulated and normalised
  fg / surface / bg layers equal the plain path's (render.py:130-132), and
  the accumulated weights equal the alpha accumulation for alpha = 1."""
  meshes = _meshes(device)
  jobs = _jobs(meshes)
  r = batch.BatchedSplatRenderer(meshes, jobs, 3, return_accum_buffers=True)
  r.render()
  torch.cuda.synchronize()
  for i, (mi, cam, w, h) in enumerate(jobs):
    m = meshes[mi]
    with torch.no_grad():
      img, accum, norm = render.render_triangles(
          m["vertices"], {"c": m["colors"]}, m["triangles"], cam, w, h,
          lambda a: a["c"], num_layers=3, return_accum_buffers=True)
    assert torch.equal(r.images[i], img)
    assert torch.equal(r.accum[i], accum)
    assert torch.equal(r.norm[i], norm)
    assert helpers.rel_err(r.accum_weights[i].cpu(),
                           accum[..., -1].cpu()) <= 1e-6


def test_camera_gradient(device):
  """d loss / d camera on the fused path (through the splat's transform,
  splat.py:209-212; rasterization held fixed) against the plain path's
  autograd and the float64 oracle."""
  meshes = _meshes(device)
  jobs = _jobs(meshes)
  r = batch.BatchedSplatRenderer(meshes, jobs, 3, camera_grads=True)
  gen = torch.Generator(device="cpu").manual_seed(7)
  for t in r.targets:
    t.copy_(torch.rand(t.shape, generator=gen))
  r.step()
  torch.cuda.synchronize()
  for i, (mi, cam, w, h) in enumerate(jobs):
    m = meshes[mi]
    cc = cam.clone().requires_grad_(True)
    img = render.render_triangles(m["vertices"], {"c": m["colors"]},
                                  m["triangles"], cc, w, h, lambda a: a["c"],
                                  num_layers=3)
    (((img - r.targets[i]) ** 2).sum() * r.loss_scale).backward()
    assert helpers.rel_err(r.d_cameras[i].cpu(), cc.grad.cpu()) <= 1e-4
    s = m["scene"]
    cam_np = cam.cpu().numpy()
    clip32 = so.transform(cam_np, s["vertices"], torch.float32).numpy()
    ids, _, b = raster_oracle.rasterize(clip32, s["triangles"], w, h, 3, 1)
    c64 = torch.as_tensor(cam_np, dtype=torch.float64).requires_grad_(True)
    o = so.render_splat(torch.as_tensor(s["vertices"], dtype=torch.float64),
                        torch.as_tensor(s["colors"], dtype=torch.float64),
                        c64, ids, b, s["triangles"])
    ((o - r.targets[i].cpu().double()) ** 2).sum().mul(r.loss_scale).backward()
    assert helpers.rel_err(r.d_cameras[i].cpu(), c64.grad) <= 1e-4
    # z row: the splat takes z from the rasterizer, so it gets no gradient.
    assert r.d_cameras[i][2].abs().max().item() == 0.0


def test_deterministic_gradients(device):
  """Deterministic mode: vertex, colour and camera gradients accumulate as
  exact 64-bit fixed-point sums, so eager steps and graph replays on four
  concurrent job groups give bit-identical results; they agree with the
  float-atomic mode to float32 rounding."""
  meshes = _meshes(device)
  jobs = _jobs(meshes)
  r = batch.BatchedSplatRenderer(meshes, jobs, 3, streams=4,
                                 camera_grads=True, deterministic=True)
  f = batch.BatchedSplatRenderer(meshes, jobs, 3, streams=4,
                                 camera_grads=True)
  gen = torch.Generator(device="cpu").manual_seed(11)
  for a, b in zip(r.targets, f.targets):
    a.copy_(torch.rand(a.shape, generator=gen))
    b.copy_(a)
  r.step()
  f.step()
  torch.cuda.synchronize()
  want = [t.clone() for t in r.records + [r.d_cameras]]
  want_loss = r.loss.clone()
  for a, b in zip(want, f.records + [f.d_cameras]):
    assert helpers.rel_err(a.cpu(), b.cpu()) <= 1e-5
  r.capture()
  for k in range(3):
    r.replay() if k else r.step()
    torch.cuda.synchronize()
    for a, b in zip(r.records + [r.d_cameras], want):
      assert torch.equal(a, b), k
    assert torch.equal(r.loss, want_loss), k
    torch.ones(1000, device=device).sum().item()


def test_invalid_batches_raise(device):
  meshes = _meshes(device)
  jobs = _jobs(meshes)
  bad = dict(meshes[0], vertices=meshes[0]["vertices"].double())
  with pytest.raises(ValueError):
    batch.BatchedSplatRenderer([bad], jobs[:1], 1)
  with pytest.raises(ValueError):
    batch.BatchedSplatRenderer(meshes, [(5, jobs[0][1], 16, 16)], 1)
  with pytest.raises(ValueError):
    batch.BatchedSplatRenderer(meshes, [(0, jobs[0][1], 0, 16)], 1)


def test_jobs_without_gradient_inputs(device):
  """A job with no target (grad_mode 1) or d_image (grad_mode 2) renders its
  image only; the source pass skips jobs with no d_records."""
  import ctypes
  from diffren_amd import _native
  meshes = _meshes(device)
  jobs = _jobs(meshes)[:1]
  r = batch.BatchedSplatRenderer(meshes, jobs, 3)
  r.targets[0].uniform_()
  r.step()
  torch.cuda.synchronize()
  want = r.images[0].clone()
  lib = _native.lib()
  table = (_native.SplatJob * 1)()
  mi, _, w, h = jobs[0]
  m = meshes[mi]
  j = table[0]
  j.clip_vertices = r.clip[0].data_ptr()
  j.triangles = m["triangles"].data_ptr()
  j.colors = m["colors"].data_ptr()
  j.camera = r.cameras[0].data_ptr()
  j.triangle_ids = r.ids[0].data_ptr()
  j.z_buffer = r.z[0].data_ptr()
  j.barycentrics = r.barys[0].data_ptr()
  image = torch.zeros_like(want)
  j.image = image.data_ptr()
  j.grad_scratch = r.scratch[0].data_ptr()
  j.num_vertices = m["vertices"].shape[0]
  j.num_triangles = m["triangles"].shape[0]
  j.width, j.height = w, h
  tiles = ctypes.c_int64()
  _native.check(lib.dr_splat_jobs_prepare(table, 1, r.C, ctypes.byref(tiles)),
                "prepare")
  dev_table = _native.table_to_device(table, device)
  ws = _native.workspace(lib.dr_render_splat_workspace_size(tiles.value),
                         device)
  loss = torch.zeros(1, device=device)
  for mode in (1, 2):
    image.zero_()
    _native.check(lib.dr_render_splat_forward(
        _native.ptr(dev_table), 1, tiles.value, 3, r.C, mode, 1.0,
        _native.ptr(loss), _native.ptr(ws), ws.numel(),
        _native.stream_handle()), "forward")
    _native.check(lib.dr_render_splat_backward(
        _native.ptr(dev_table), 1, tiles.value, 3, r.C, 0,
        _native.stream_handle()), "backward")
    torch.cuda.synchronize()
    assert torch.equal(image, want)
    assert loss.item() == 0.0


def test_bin_overflow_regrows(device):
  """A mesh grown past the autotuned bin capacity, in a captured graph:
  the overflowing replay is still bit-exact (its overflowing tiles read
  every triangle record of their job), the next replay grows the bins and
  captures the graph again, and from then on nothing overflows."""
  v, t = scenes.uv_sphere_mesh(100, 200, seed=3)
  rng = np.random.default_rng(8)
  col = np.concatenate([rng.uniform(0, 1, (v.shape[0], 3)),
                        np.ones((v.shape[0], 1))], 1).astype(np.float32)
  # Tuned while the mesh is small on screen (0.25 x: a few thousand
  # fragments), rendered at full size (about 20k fragments, each a bin
  # entry of a small triangle).
  full = torch.as_tensor(v, device=device)
  mesh = dict(vertices=full * 0.25,
              triangles=torch.as_tensor(t, device=device),
              colors=torch.as_tensor(col, device=device))
  cam_np = scenes.orbit_cameras(1, seed=2)[0]
  cam = torch.as_tensor(cam_np, device=device)
  R, L = 256, 3
  r = batch.BatchedSplatRenderer([mesh], [(0, cam, R, R)], L, tile="auto")
  cap = r.raster_config.bin_capacity
  r.capture(forward_only=True)
  r.replay()
  torch.cuda.synchronize()
  need0, over0 = r.bin_stats()
  assert not over0 and need0 <= cap

  def check_ids():
    clip = so.transform(cam_np, mesh["vertices"].cpu().numpy(),
                        torch.float32).numpy()
    want, _, _ = raster_oracle.rasterize(clip, t, R, R, L, 1)
    assert np.array_equal(r.ids[0].cpu().numpy(), want)

  mesh["vertices"].copy_(full)
  r.replay()
  torch.cuda.synchronize()
  need1, over1 = r.bin_stats()
  assert over1 and need1 > cap, (need0, need1, cap)
  check_ids()
  assert r.regrows == 0
  r.replay()
  torch.cuda.synchronize()
  assert r.regrows == 1
  assert r.raster_config.bin_capacity >= 2 * need1
  need2, over2 = r.bin_stats()
  assert not over2 and need2 == need1
  check_ids()
  want = r.images[0].clone()
  r.replay()
  torch.cuda.synchronize()
  assert r.regrows == 1
  assert torch.equal(r.images[0], want)


_LIGHT = {}


def _point_light(attrs):
  """RGBA point-light shader on interpolated colour / normal / position
  (shaders.point_light, shaders.py:100-173), alpha interpolated.  Its
  constants live on the device (created once, outside any graph capture)."""
  from diffren_amd import shading
  dev = attrs["color"].device
  if dev not in _LIGHT:
    _LIGHT[dev] = (torch.tensor([1.0, 2.5, 3.0], device=dev),
                   torch.zeros(3, device=dev))
  light, eye = _LIGHT[dev]
  rgb = shading.point_light(attrs["color"][..., :3],
                            shading.l2_normalize(attrs["normal"]),
                            attrs["position"], light, eye_position=eye,
                            ks=0.5, ns=8.0)
  return torch.cat([rgb, attrs["color"][..., 3:]], -1)


def _shaded_meshes(device):
  from diffren_amd import mesh
  meshes = _meshes(device)
  for m in meshes:
    m["attributes"] = {
        "position": m["vertices"].clone().requires_grad_(True),
        "normal": mesh.compute_vertex_normals(m["vertices"],
                                              m["triangles"]),
        "color": m["colors"].clone().requires_grad_(True)}
  return meshes


@pytest.mark.parametrize("streams", [1, 4])
def test_batched_shader_matches_per_view(streams, device):
  """BatchedSplatRenderer with a point-light shading_function against
  render_triangles per view with the same shader (splat on the fused
  kernels) and against the unfused per-view path: images, loss, the
  position gradient through the splat and the attribute gradients through
  the shader; then graph replays against the eager step."""
  meshes = _shaded_meshes(device)
  jobs = _jobs(meshes)
  L = 3
  r = batch.BatchedSplatRenderer(meshes, jobs, L, streams=streams,
                                 shading_function=_point_light, channels=4)
  assert r.d_colors is None
  gen = torch.Generator(device="cpu").manual_seed(11)
  for t in r.targets:
    t.copy_(torch.rand(t.shape, generator=gen))
  r.step()
  torch.cuda.synchronize()
  npix = sum(w * h * 4 for _, _, w, h in jobs)
  for fused in (None, False):
    vs = [m["vertices"].clone().requires_grad_(True) for m in meshes]
    leaves = [{k: a.detach().clone().requires_grad_(a.requires_grad)
               for k, a in m["attributes"].items()} for m in meshes]
    total = 0.0
    for i, ((mi, cam, w, h), tgt) in enumerate(zip(jobs, r.targets)):
      img = render.render_triangles(vs[mi], leaves[mi],
                                    meshes[mi]["triangles"], cam, w, h,
                                    _point_light, num_layers=L,
                                    fused_splat=fused)
      tol = 0.0 if fused is None else 1e-5
      assert helpers.rel_err(r.images[i].cpu(), img.detach().cpu()) <= tol
      total = total + ((img - tgt) ** 2).sum() / npix
    total.backward()
    assert abs(r.loss.item() - total.item()) <= 1e-5 * abs(total.item())
    for mi in range(len(meshes)):
      assert helpers.rel_err(r.d_vertices[mi].cpu(), vs[mi].grad.cpu()) <= 1e-4
      for k in ("position", "color"):
        assert helpers.rel_err(r.d_attributes[mi][k].cpu(),
                               leaves[mi][k].grad.cpu()) <= 1e-4, (mi, k)
      assert "normal" not in r.d_attributes[mi]
  want = [t.clone() for t in r.images + r.d_vertices] + [
      d[k].clone() for d in r.d_attributes for k in sorted(d)]
  want_loss = r.loss.item()
  r.capture()
  for k in range(3):
    r.replay()
    torch.cuda.synchronize()
    got = r.images + r.d_vertices + [d[n] for d in r.d_attributes
                                     for n in sorted(d)]
    for a, b in zip(got, want):
      assert helpers.rel_err(a.cpu(), b.cpu()) <= 1e-5, k
    assert abs(r.loss.item() - want_loss) <= 1e-5 * abs(want_loss)
    torch.ones(1000, device=device).sum().item()


def test_batched_shader_forward_only(device):
  """render() with a shading_function: images equal the step's."""
  meshes = _shaded_meshes(device)
  jobs = _jobs(meshes)
  r = batch.BatchedSplatRenderer(meshes, jobs, 2,
                                 shading_function=_point_light, channels=4)
  r.render()
  torch.cuda.synchronize()
  imgs = [t.clone() for t in r.images]
  r.step()
  torch.cuda.synchronize()
  for a, b in zip(imgs, r.images):
    assert torch.equal(a, b)
  with pytest.raises(ValueError):
    batch.BatchedSplatRenderer(meshes, jobs, 2, shading_function=_point_light)
