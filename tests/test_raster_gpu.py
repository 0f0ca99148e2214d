"""GPU tile rasterizer vs the oracle: triangle ids, z and barycentrics must be
bit-identical (rasterize_triangles_impl.cc:27-118 semantics)."""
import hashlib

import numpy as np
import pytest
import torch

import helpers
import scenes as golden_scenes
from diffren_amd import _native
from diffren_amd import rasterize
from diffren_amd import scenes
from oracle import raster_oracle

pytestmark = pytest.mark.gpu

GOLDEN = np.load(helpers.os.path.join(helpers.os.path.dirname(__file__),
                                      "golden", "raster_golden.npz"))
CASES = sorted({k.split("/")[0] for k in GOLDEN.files})
TILES = [(8, 8), (16, 16), (32, 8), (32, 16), (32, 32), (64, 8)]


def _digest(a):
  return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def gpu_raster(v, t, w, h, l, cull, device, **kw):
  print(f"raster V={len(v)} T={len(t)} {w}x{h} L={l} cull={cull} {kw}",
        flush=True)
  vt = torch.as_tensor(np.ascontiguousarray(v, np.float32), device=device)
  tt = torch.as_tensor(np.ascontiguousarray(t, np.int32), device=device)
  ids, z, b = rasterize.rasterize_triangles_primitive(vt, tt, w, h, l, cull,
                                                      **kw)
  torch.cuda.synchronize()
  return ids.cpu().numpy(), z.cpu().numpy(), b.cpu().numpy()


def assert_same(got, want):
  for g, w, name in zip(got, want, ("ids", "z", "barycentrics")):
    if not helpers.same_bits(g, w):
      bad = np.argwhere(g.view(np.uint32) != w.view(np.uint32)
                        if g.dtype != np.int32 else g != w)
      raise AssertionError(f"{name} differ at {len(bad)} places, first "
                           f"{bad[:4].tolist()}")


@pytest.mark.parametrize("name", CASES)
def test_golden_vectors_bit_exact(name, device):
  v = GOLDEN[f"{name}/vertices"]
  t = GOLDEN[f"{name}/triangles"]
  w, h, l, cull = (int(x) for x in GOLDEN[f"{name}/params"])
  ids, z, b = gpu_raster(v, t, w, h, l, cull, device)
  assert [_digest(ids), _digest(z), _digest(b)] == list(
      GOLDEN[f"{name}/sha256"])


@pytest.mark.parametrize("tile", TILES)
def test_tile_shapes_bit_exact(tile, device):
  for seed in range(3):
    v, t = golden_scenes.adversarial_scene(200 + seed, 60, 200)
    w, h = [(64, 48), (33, 17), (100, 7)][seed]
    l, cull = 1 + seed, seed
    assert_same(gpu_raster(v, t, w, h, l, cull, device, tile=tile),
                raster_oracle.rasterize(v, t, w, h, l, cull))


@pytest.mark.parametrize("threads", [128, 256])
@pytest.mark.parametrize("tile", [(8, 8), (32, 8), (32, 32)])
def test_tile_block_sizes_bit_exact(tile, threads, device):
  # dr_raster_config.tile_threads: the tile pass in 128- or 256-thread
  # blocks, on scenes with large triangles (wave-walked queues), dense
  # overdraw (entry loops of several passes) and exact depth ties.
  rng = np.random.default_rng(threads + tile[0])
  scenes_ = [golden_scenes.adversarial_scene(300 + threads, 80, 400),
             _stack(60, rng, coplanar=True), _stack(90, rng, coplanar=False)]
  for k, (v, t) in enumerate(scenes_):
    w, h, l = [(96, 80, 3), (48, 40, 8), (64, 56, 4)][k]
    cfg = rasterize._config(tile, 0, 0, l)
    cfg.tile_threads = threads
    vt = torch.as_tensor(np.ascontiguousarray(v, np.float32), device=device)
    tt = torch.as_tensor(np.ascontiguousarray(t, np.int32), device=device)
    print(f"tile {tile} threads {threads} scene {k}", flush=True)
    ids, z, b = rasterize._rasterize_one(vt, tt, w, h, l, 0, cfg)
    torch.cuda.synchronize()
    assert_same((ids.cpu().numpy(), z.cpu().numpy(), b.cpu().numpy()),
                raster_oracle.rasterize(v, t, w, h, l, 0))


@pytest.mark.parametrize("seed", range(10))
def test_random_scenes_bit_exact(seed, device):
  rng = np.random.default_rng(seed)
  v, t = golden_scenes.adversarial_scene(500 + seed, 50, 300)
  w = int(rng.choice([1, 5, 16, 33, 128]))
  h = int(rng.choice([1, 7, 16, 64]))
  l = int(rng.integers(1, 5))
  cull = int(rng.integers(0, 3))
  assert_same(gpu_raster(v, t, w, h, l, cull, device),
              raster_oracle.rasterize(v, t, w, h, l, cull))


def test_bin_overflow_falls_back_exactly(device):
  s = scenes.icosphere_scene(seed=3)
  clip = scenes.to_clip(s["vertices"], s["camera"])
  want = raster_oracle.rasterize(clip, s["triangles"], 96, 80, 2, 0)
  for cap in (1, 100, 1000):
    assert_same(gpu_raster(clip, s["triangles"], 96, 80, 2, 0, device,
                           bin_capacity=cap), want)


def test_icosphere_workload_bit_exact(device):
  s = scenes.icosphere_scene(seed=0)
  clip = scenes.to_clip(s["vertices"], s["camera"])
  assert_same(gpu_raster(clip, s["triangles"], 256, 256, 1, 1, device),
              raster_oracle.rasterize(clip, s["triangles"], 256, 256, 1, 1))


def test_million_triangle_sphere_bit_exact(device):
  v, t = scenes.uv_sphere_mesh(500, 1000)
  cam = scenes.orbit_cameras(1, seed=0)[0]
  clip = scenes.to_clip(v, cam)
  want = raster_oracle.rasterize(clip, t, 1024, 1024, 3, 1)
  assert_same(gpu_raster(clip, t, 1024, 1024, 3, 1, device), want)
  assert (want[0][0] >= 0).mean() > 0.2


def test_uv_sphere_ids_match_reference_png(device):
  # rasterize_triangles_impl_cuda_test.cc:171-187 (230x230 UV sphere at
  # 1600x1600, no culling): bit-exact to the oracle and within the C++
  # tests' tolerance of UV_Sphere_Ids.png.
  v, t = golden_scenes.reference_uv_sphere(230, 230, 0.9)
  got = gpu_raster(v, t, 1600, 1600, 1, 0, device)
  assert_same(got, raster_oracle.rasterize(v, t, 1600, 1600, 1, 0))
  ok, msg = helpers.cpp_images_are_near(
      helpers.load_fixture("UV_Sphere_Ids.png"), helpers.encode_ids(got[0], 0))
  assert ok, msg


def test_large_triangles_cross_many_tiles(device):
  # Few triangles covering the whole frame: the wave-per-triangle path.
  v = np.array([[-1, -1, 0.2, 1], [1, -1, 0.2, 1], [1, 1, -0.3, 1],
                [-1, 1, 0.5, 1], [0, 0, 0.9, 1]], np.float32)
  t = np.array([[0, 1, 2], [0, 2, 3], [0, 1, 4], [1, 2, 4], [2, 3, 4]],
               np.int32)
  for cull in (0, 1, 2):
    assert_same(gpu_raster(v, t, 300, 200, 3, cull, device),
                raster_oracle.rasterize(v, t, 300, 200, 3, cull))


def test_batched_jobs_match_single_jobs(device):
  jobs = []
  for i, (w, h) in enumerate([(64, 64), (128, 32), (17, 200), (256, 256)]):
    s = scenes.icosphere_scene(seed=10 + i, subdivisions=2 + i % 2)
    jobs.append((scenes.to_clip(s["vertices"], s["camera"]), s["triangles"],
                 w, h))
  vs = [torch.as_tensor(j[0], device=device) for j in jobs]
  ts = [torch.as_tensor(j[1], device=device) for j in jobs]
  out = rasterize.rasterize_batched(vs, ts, [(j[2], j[3]) for j in jobs], 3,
                                    1)
  for (v, t, w, h), (ids, z, b) in zip(jobs, out):
    assert_same((ids.cpu().numpy(), z.cpu().numpy(), b.cpu().numpy()),
                raster_oracle.rasterize(v, t, w, h, 3, 1))


def test_no_barycentrics_and_empty_mesh(device):
  s = scenes.icosphere_scene(seed=1)
  clip = scenes.to_clip(s["vertices"], s["camera"])
  vt = torch.as_tensor(clip, device=device)
  tt = torch.as_tensor(s["triangles"], device=device)
  ids, z, _ = rasterize._rasterize_one(vt, tt, 64, 64, 2, 1,
                                       rasterize._config(), want_barys=False)
  want = raster_oracle.rasterize(clip, s["triangles"], 64, 64, 2, 1)
  assert helpers.same_bits(ids.cpu().numpy(), want[0])
  assert helpers.same_bits(z.cpu().numpy(), want[1])
  ids, z, b = gpu_raster(clip, np.zeros((0, 3), np.int32), 16, 8, 2, 0, device)
  assert (ids == -1).all() and (z == 1).all() and (b == 0).all()


def test_invalid_arguments_raise(device):
  vt = torch.zeros((3, 4), device=device)
  tt = torch.zeros((1, 3), dtype=torch.int32, device=device)
  with pytest.raises(ValueError):
    rasterize.rasterize_triangles_primitive(vt, tt, 0, 4, 1, 0)
  with pytest.raises(ValueError):
    rasterize.rasterize_triangles_primitive(vt, tt, 4, 4, 65, 0)
  with pytest.raises(ValueError):
    rasterize.rasterize_triangles_primitive(vt, tt, 4, 4, 1, 3)


def _stack(num, rng, coplanar):
  """`num` overlapping random triangles around the image centre; coplanar
  ones all lie at z/w = 0.25 (every fragment ties), others at random z."""
  v = []
  for i in range(num):
    c = rng.uniform(-0.3, 0.3, 2)
    z = 0.25 if coplanar else rng.uniform(-0.9, 0.9)
    w = rng.uniform(0.5, 2.0)
    for k in range(3):
      a = 2 * np.pi * (k / 3.0 + rng.uniform(-0.1, 0.1))
      p = c + rng.uniform(0.3, 0.8) * np.array([np.cos(a), np.sin(a)])
      v.append([p[0] * w, p[1] * w, z * w, w])
  t = np.arange(3 * num, dtype=np.int32).reshape(num, 3)
  return np.array(v, np.float32), t


@pytest.mark.parametrize("layers", [1, 2, 4, 8, 16, 40])
def test_exact_depth_ties_follow_reference_order(layers, device):
  # Every fragment has the same z: the kept ids and their order are those of
  # the reference's id-ordered insertion with strict '<', not simply the
  # lowest ids.
  rng = np.random.default_rng(layers)
  v, t = _stack(40, rng, coplanar=True)
  for cull in (0, 1):
    assert_same(gpu_raster(v, t, 48, 40, layers, cull, device),
                raster_oracle.rasterize(v, t, 48, 40, layers, cull))


def test_eight_layer_stack(device):
  rng = np.random.default_rng(11)
  v, t = _stack(60, rng, coplanar=False)
  got = gpu_raster(v, t, 64, 64, 8, 0, device)
  assert_same(got, raster_oracle.rasterize(v, t, 64, 64, 8, 0))
  assert (got[0][7] >= 0).any()  # the deepest layer is populated


@pytest.mark.parametrize("layers,tile", [(24, None), (64, None), (32, (16, 16))])
def test_deep_layer_stacks(layers, tile, device):
  # More layers than a splat uses: 24 and 32 fit the default and 16x16
  # tiles' key lists, 64 switches to 8x8 tiles (the LDS key budget).
  rng = np.random.default_rng(layers)
  v, t = _stack(90, rng, coplanar=False)
  got = gpu_raster(v, t, 64, 56, layers, 0, device, tile=tile)
  assert_same(got, raster_oracle.rasterize(v, t, 64, 56, layers, 0))
  assert (got[0][min(layers, 40) - 1] >= 0).any()


def _scaled_scene(seed, n, width):
  """`n` triangles of 0.5-12 pixels at random places and depths, each with
  its clip coordinates scaled by w = 2^k, k in [-3, 16], half of them wound
  clockwise.  Scaling by w leaves the screen triangle but multiplies the
  fixed-point matrix by w and its adjugate by w^2, so covering triangles
  fall on both sides of 2^33 (where the kernels switch between exact double
  and int64 edge values), and those with coordinates of 2^15 and more take
  the general fixed-point path (64-bit entries, wrapping products)."""
  rng = np.random.default_rng(seed)
  v = []
  for _ in range(n):
    c = rng.uniform(-1.0, 1.0, 2)
    size = rng.uniform(0.5, 12.0) * 2.0 / width
    w = 2.0 ** int(rng.integers(-3, 17))
    z = rng.uniform(-0.9, 0.9)
    for k in range(3):
      a = 2 * np.pi * (k / 3.0 + rng.uniform(-0.12, 0.12))
      p = c + size * rng.uniform(0.5, 1.0) * np.array([np.cos(a), np.sin(a)])
      v.append([p[0] * w, p[1] * w, (z + rng.uniform(-0.05, 0.05)) * w, w])
  t = np.arange(3 * n, dtype=np.int32).reshape(n, 3)
  t[::2] = t[::2][:, [0, 2, 1]]
  return np.array(v, np.float32), t


def _adjugate_max(v, t):
  """Largest adjugate entry magnitude of each triangle's fixed-point matrix
  (columns (x, y, w) * 2^16 of its vertices; the adjugate's rows are the
  columns' cross products), in float64 -- enough to tell 2^32 from 2^34."""
  m = np.trunc(v[t][:, :, [0, 1, 3]].astype(np.float64) * 65536.0)
  r = [np.cross(m[:, 1], m[:, 2]), np.cross(m[:, 2], m[:, 0]),
       np.cross(m[:, 0], m[:, 1])]
  return np.abs(np.concatenate(r, 1)).max(1)


@pytest.mark.parametrize("cull", [0, 1, 2])
def test_edge_value_paths_bit_exact(cull, device):
  # Triangles whose adjugates lie below and above 2^33, in the count pass's
  # inline path (boxes of at most 4 x 4 pixels) and the tile pass's.
  v, t = _scaled_scene(40 + cull, 600, 96)
  want = raster_oracle.rasterize(v, t, 96, 80, 3, cull)
  assert_same(gpu_raster(v, t, 96, 80, 3, cull, device), want)
  hit = np.unique(want[0][want[0] >= 0])
  rmax = _adjugate_max(v, t)[hit]
  assert (rmax < 2.0 ** 32).sum() > 20 and (rmax > 2.0 ** 34).sum() > 20
  assert (np.abs(v[t[hit]]).max((1, 2)) >= 32768).any()


def test_prefix_mode_writes_up_to_the_first_empty_layer(device):
  # DR_RASTER_SPARSE | DR_RASTER_PREFIX: a pixel's layers are written up to
  # and including its first empty one, exactly as the oracle has them; the
  # layers behind it, and background z / barycentrics, keep what the buffers
  # held before.
  s = scenes.icosphere_scene(seed=5)
  clip = scenes.to_clip(s["vertices"], s["camera"])
  L, W, H = 4, 96, 80
  want = raster_oracle.rasterize(clip, s["triangles"], W, H, L, 0)
  vt = torch.as_tensor(clip, device=device)
  tt = torch.as_tensor(s["triangles"], device=device)
  ids = torch.full((L, H, W), 12345, dtype=torch.int32, device=device)
  z = torch.full((L, H, W), 7.0, device=device)
  b = torch.full((L, H, W, 3), 7.0, device=device)
  cfg = rasterize._config(
      flags=_native.RASTER_SPARSE | _native.RASTER_PREFIX)
  rasterize._rasterize_one(vt, tt, W, H, L, 0, cfg, out=(ids, z, b))
  ids, z, b = ids.cpu().numpy(), z.cpu().numpy(), b.cpu().numpy()
  frag = want[0] >= 0
  written = np.ones((L, H, W), bool)
  written[1:] = np.cumprod(frag[:-1], axis=0).astype(bool)
  assert np.array_equal(ids[written], want[0][written])
  assert (ids[~written] == 12345).all()
  assert helpers.same_bits(z[frag], want[1][frag])
  assert helpers.same_bits(b[frag], want[2][frag])
  assert (z[~frag] == 7.0).all() and (b[~frag] == 7.0).all()
  assert (~written).any() and (written & ~frag).any() and frag[1].any()


@pytest.mark.parametrize("w,h", [(1, 1), (2048, 160), (160, 2048), (31, 1)])
def test_extreme_image_shapes(w, h, device):
  v, t = golden_scenes.adversarial_scene(900 + w, 50, 200)
  assert_same(gpu_raster(v, t, w, h, 3, 0, device),
              raster_oracle.rasterize(v, t, w, h, 3, 0))


@pytest.mark.parametrize("batched", [False, True])
@pytest.mark.parametrize("axes", [(None, None), (None, 0), (0, None), (0, 0)])
def test_primitive_batch_axes_simple_triangle(axes, batched, device):
  # rasterize_triangles_xla_test.py:35-85: the primitive with and without
  # leading batch axes on vertices and / or triangles renders
  # Simple_Triangle.png.
  v = torch.tensor([[-0.5, -0.5, 0.8, 1.0], [0.0, 0.5, 0.3, 1.0],
                    [0.5, -0.5, 0.3, 1.0]], device=device)
  t = torch.tensor([[0, 1, 2]], dtype=torch.int32, device=device)
  if axes[0] is not None:
    v = v.unsqueeze(0)
  if axes[1] is not None:
    t = t.unsqueeze(0)
  ids, z, b = rasterize.rasterize_triangles_primitive(v, t, 640, 480, 1, 0,
                                                      batched=batched)
  lead = 1 if axes != (None, None) else 0
  assert tuple(b.shape) == (1,) * lead + (1, 480, 640, 3)
  b = b.reshape(480, 640, 3).cpu().numpy()
  rgba = np.concatenate([b, np.ones((480, 640, 1), np.float32)], 2)
  ok, msg = helpers.check_image(rgba, "Simple_Triangle.png")
  assert ok, msg
  want = raster_oracle.rasterize(v.reshape(3, 4).cpu().numpy(),
                                 t.reshape(1, 3).cpu().numpy(), 640, 480, 1,
                                 0)
  assert_same((ids.reshape(1, 480, 640).cpu().numpy(),
               z.reshape(1, 480, 640).cpu().numpy(),
               b.reshape(1, 480, 640, 3)), want)

