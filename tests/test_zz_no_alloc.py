"""dr_rasterize_triangles_noalloc (static scratch, no workspace argument)
against dr_rasterize_triangles (caller workspace), through the C ABI: two
calls in a row, the second on a smaller input, each bit-identical in ids, z
and barycentrics; and a call that needs more than the static scratch is
refused before any launch."""
import ctypes

import numpy as np
import pytest
import torch

from diffren_amd import _native
from diffren_amd import scenes
from oracle import splat_oracle as so

pytestmark = pytest.mark.gpu


def _clip_mesh(rows, cols, seed, device):
  v, t = scenes.uv_sphere_mesh(rows, cols, seed=seed)
  cam = scenes.orbit_cameras(1, seed=seed)[0]
  clip = so.transform(cam, v, torch.float32).numpy()
  return (torch.as_tensor(clip, device=device).contiguous(),
          torch.as_tensor(t, device=device).contiguous())


def _outputs(L, h, w, device):
  return (torch.full((L, h, w), 7, dtype=torch.int32, device=device),
          torch.full((L, h, w), 7.0, dtype=torch.float32, device=device),
          torch.full((L, h, w, 3), 7.0, dtype=torch.float32, device=device))


def _with_workspace(clip, tris, w, h, L, cfg, device):
  lib = _native.lib()
  ids, z, b = _outputs(L, h, w, device)
  tiles = ((w + cfg.tile_w - 1) // cfg.tile_w) * \
      ((h + cfg.tile_h - 1) // cfg.tile_h)
  ws = _native.workspace(lib.dr_raster_workspace_size(
      tris.shape[0], tiles, ctypes.byref(cfg)), device)
  _native.check(lib.dr_rasterize_triangles(
      _native.ptr(clip), clip.shape[0], _native.ptr(tris), tris.shape[0], w,
      h, L, 1, _native.ptr(ids), _native.ptr(z), _native.ptr(b),
      ctypes.byref(cfg), _native.ptr(ws), ws.numel(),
      _native.stream_handle()), "dr_rasterize_triangles")
  return ids, z, b


def _no_alloc(clip, tris, w, h, L, cfg, device):
  lib = _native.lib()
  ids, z, b = _outputs(L, h, w, device)
  _native.check(lib.dr_rasterize_triangles_noalloc(
      _native.ptr(clip), clip.shape[0], _native.ptr(tris), tris.shape[0], w,
      h, L, 1, _native.ptr(ids), _native.ptr(z), _native.ptr(b),
      ctypes.byref(cfg), _native.stream_handle()),
      "dr_rasterize_triangles_noalloc")
  return ids, z, b


def _same(a, b):
  for x, y in zip(a, b):
    assert x.dtype == y.dtype and x.shape == y.shape
    assert np.array_equal(x.cpu().numpy().view(np.uint8),
                          y.cpu().numpy().view(np.uint8))


def test_no_alloc_matches_workspace_entry_point(device):
  cfg = _native.RasterConfig(16, 16, 0, 0, 0)
  # First call: 120k triangles at 512x384, 3 layers.
  clip, tris = _clip_mesh(200, 300, 1, device)
  want = _with_workspace(clip, tris, 512, 384, 3, cfg, device)
  got = _no_alloc(clip, tris, 512, 384, 3, cfg, device)
  torch.cuda.synchronize()
  assert (want[0] >= 0).any()
  _same(got, want)
  # Second call in a row on a smaller input: 7k triangles at 96x80, 2
  # layers; the scratch still holds the first call's bins and counts.
  clip2, tris2 = _clip_mesh(30, 120, 2, device)
  want2 = _with_workspace(clip2, tris2, 96, 80, 2, cfg, device)
  got2 = _no_alloc(clip2, tris2, 96, 80, 2, cfg, device)
  torch.cuda.synchronize()
  assert (want2[0] >= 0).any()
  _same(got2, want2)


def test_no_alloc_refuses_calls_beyond_the_static_scratch(device):
  lib = _native.lib()
  cfg = _native.RasterConfig(16, 16, 0, 0, 0)
  # 12M degenerate triangles (every index 0): their workspace exceeds the
  # 64 MiB scratch.  The buffers are real, so even a launch would stay in
  # bounds; the call must be refused before it.
  n = 12_000_000
  tris = torch.zeros((n, 3), dtype=torch.int32, device=device)
  clip = torch.zeros((1, 4), dtype=torch.float32, device=device)
  tiles = (64 // 16) * (64 // 16)
  assert lib.dr_raster_workspace_size(n, tiles, ctypes.byref(cfg)) > (64 << 20)
  ids, z, b = _outputs(1, 64, 64, device)
  rc = lib.dr_rasterize_triangles_noalloc(
      _native.ptr(clip), 1, _native.ptr(tris), n, 64, 64, 1, 1,
      _native.ptr(ids), _native.ptr(z), _native.ptr(b), ctypes.byref(cfg),
      _native.stream_handle())
  torch.cuda.synchronize()
  assert rc == 2  # DR_WORKSPACE_TOO_SMALL
  assert b"static scratch" in lib.dr_last_error()
  assert bool((ids == 7).all())
