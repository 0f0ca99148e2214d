"""The rasterizer oracle (oracle/raster_oracle.c) pinned on the CPU.

Pins, in order of strength:
  1. bit-identity with the reference's own CPU rasterizer compiled from its
     sources (oracle/_ref, built only where the reference is present);
  2. the committed golden vectors written from that compiled reference
     (tests/golden/raster_golden.npz, SHA-256 of the exact output bytes);
  3. the reference's PNG fixtures with the C++ tests' tolerance
     (rasterize_triangles_impl_test.cc:125-294);
  4. invariants of the C++ tests: barycentrics sum to 1 on covered pixels and
     0 elsewhere, edge pixels covered, degenerate triangles never covered,
     culling empties the second cube layer.
"""
import hashlib

import numpy as np
import pytest

import helpers
import scenes as golden_scenes
from oracle import raster_oracle

GOLDEN = np.load(helpers.os.path.join(helpers.os.path.dirname(__file__),
                                      "golden", "raster_golden.npz"))
CASES = sorted({k.split("/")[0] for k in GOLDEN.files})


def _digest(a):
  return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


@pytest.mark.parametrize("name", CASES)
def test_oracle_matches_golden_vectors(name):
  v = GOLDEN[f"{name}/vertices"]
  t = GOLDEN[f"{name}/triangles"]
  w, h, l, cull = (int(x) for x in GOLDEN[f"{name}/params"])
  ids, z, b = raster_oracle.rasterize(v, t, w, h, l, cull)
  want = list(GOLDEN[f"{name}/sha256"])
  assert [_digest(ids), _digest(z), _digest(b)] == want
  if f"{name}/ids" in GOLDEN.files:
    assert helpers.same_bits(ids, GOLDEN[f"{name}/ids"])
    assert helpers.same_bits(z, GOLDEN[f"{name}/z"])
    assert helpers.same_bits(b, GOLDEN[f"{name}/barys"])


@pytest.mark.skipif(not raster_oracle.ref_available(),
                    reason="reference sources not built here (oracle/_ref)")
@pytest.mark.parametrize("seed", range(40))
def test_oracle_bit_identical_to_compiled_reference(seed):
  rng = np.random.default_rng(1000 + seed)
  v, t = golden_scenes.adversarial_scene(1000 + seed)
  W = int(rng.choice([1, 2, 7, 16, 33, 64, 100]))
  H = int(rng.choice([1, 3, 16, 31, 64]))
  L = int(rng.integers(1, 4))
  cull = int(rng.integers(0, 3))
  a = raster_oracle.rasterize(v, t, W, H, L, cull)
  b = raster_oracle.ref_rasterize(v, t, W, H, L, cull)
  for x, y in zip(a, b):
    assert helpers.same_bits(x, y)


@pytest.mark.parametrize("case,fixture", [
    ("simple_triangle", "Simple_Triangle.png"),
    ("perspective_triangle", "Perspective_Corrected_Triangle.png"),
    ("external_triangle", "External_Triangle.png"),
    ("inside_box", "Inside_Box.png"),
    ("tetrahedron", "Simple_Tetrahedron.png"),
])
def test_oracle_barycentrics_match_reference_png(case, fixture):
  cases = {c[0]: c for c in golden_scenes.reference_cases()}
  _, v, t, w, h, l, cull = cases[case]
  _, _, b = raster_oracle.rasterize(v, t, w, h, l, cull)
  ok, msg = helpers.cpp_images_are_near(helpers.load_fixture(fixture),
                                        helpers.encode_barycentrics(b))
  assert ok, msg


def test_oracle_cube_layers_match_reference_png():
  v, t = golden_scenes.CUBE_VERTICES, golden_scenes.CUBE_TRIANGLES
  ids, _, b = raster_oracle.rasterize(v, t, 640, 480, 2, 0)
  for layer, idf, bf in [(0, "Ids_Cube.png", "Barycentrics_Cube.png"),
                         (1, "Ids_Cube_Back.png",
                          "Barycentrics_Cube_Back.png")]:
    ok, msg = helpers.cpp_images_are_near(helpers.load_fixture(idf),
                                          helpers.encode_ids(ids, layer))
    assert ok, msg
    ok, msg = helpers.cpp_images_are_near(
        helpers.load_fixture(bf), helpers.encode_barycentrics(b, layer))
    assert ok, msg


@pytest.mark.parametrize("cull,fixture", [(1, "Ids_Cube_Back.png"),
                                          (2, "Ids_Cube.png")])
def test_oracle_culling(cull, fixture):
  v, t = golden_scenes.CUBE_VERTICES, golden_scenes.CUBE_TRIANGLES
  ids, _, _ = raster_oracle.rasterize(v, t, 640, 480, 2, cull)
  ok, msg = helpers.cpp_images_are_near(helpers.load_fixture(fixture),
                                        helpers.encode_ids(ids, 0))
  assert ok, msg
  assert (ids[1] == -1).all()


def test_oracle_uv_sphere_ids_match_reference_png():
  # rasterize_triangles_impl_cuda_test.cc:171-187: 230x230 UV sphere of
  # radius 0.9 at 1600x1600, one layer, no culling, ids encoded by colour.
  v, t = golden_scenes.reference_uv_sphere(230, 230, 0.9)
  assert t.shape == (2 * 230 + 2 * 229 * 230, 3)
  ids, _, _ = raster_oracle.rasterize(v, t, 1600, 1600, 1, 0)
  ok, msg = helpers.cpp_images_are_near(helpers.load_fixture("UV_Sphere_Ids.png"),
                                        helpers.encode_ids(ids, 0))
  assert ok, msg


def test_oracle_invariants():
  # Edge pixel (rasterize_triangles_impl_test.cc:170-198), both windings.
  for tri in ([[0, 1, 2]], [[2, 1, 0]]):
    _, _, b = raster_oracle.rasterize(golden_scenes.EDGE_TRIANGLE,
                                      np.array(tri, np.int32), 641, 480, 1, 0)
    assert abs(b[0, 5, 320].sum() - 1.0) < 1e-6
  # Image corners (:200-211).
  _, _, b = raster_oracle.rasterize(golden_scenes.FULL_QUAD,
                                    np.array([[0, 1, 2], [0, 2, 3]], np.int32),
                                    640, 480, 1, 0)
  for y, x in [(0, 0), (0, 639), (479, 639), (479, 0)]:
    assert abs(b[0, y, x].sum() - 1.0) < 1e-6
  # Degenerate triangle (:213-225).
  ids, _, b = raster_oracle.rasterize(golden_scenes.DEGENERATE,
                                      np.array([[0, 1, 2]], np.int32), 1, 1,
                                      1, 0)
  assert ids[0, 0, 0] == -1 and b.sum() == 0.0
  # Sums: 1 on foreground, 0 on background; z in [-1, 1).
  ids, z, b = raster_oracle.rasterize(golden_scenes.CUBE_VERTICES,
                                      golden_scenes.CUBE_TRIANGLES, 64, 48, 2,
                                      0)
  fg = ids >= 0
  assert np.allclose(b.sum(-1)[fg], 1.0, atol=1e-6)
  assert (b[~fg] == 0).all() and (z[~fg] == 1.0).all()
  assert (z[fg] >= -1).all() and (z[fg] < 1).all()
  # Layers are sorted by (z, id).
  both = fg[0] & fg[1]
  assert ((z[0][both] < z[1][both]) |
          ((z[0][both] == z[1][both]) & (ids[0][both] < ids[1][both]))).all()


def test_oracle_brute_force_tiny():
  """Tiny scenes against a direct per-pixel restatement in Python: the L
  smallest (z, id) pairs over all covering triangles."""
  for seed in range(8):
    rng = np.random.default_rng(seed)
    v = np.concatenate([rng.uniform(-1.2, 1.2, (9, 2)),
                        rng.uniform(-0.9, 0.9, (9, 1)),
                        np.ones((9, 1))], 1).astype(np.float32)
    t = rng.integers(0, 9, (12, 3)).astype(np.int32)
    W, H, L = 5, 4, 3
    ids, z, _ = raster_oracle.rasterize(v, t, W, H, L, 0)
    # Per-pixel candidates from single-triangle runs of the oracle.
    cand = [[[] for _ in range(W)] for _ in range(H)]
    for k in range(t.shape[0]):
      i1, z1, _ = raster_oracle.rasterize(v, t[k:k + 1], W, H, 1, 0)
      for y in range(H):
        for x in range(W):
          if i1[0, y, x] == 0:
            cand[y][x].append((float(z1[0, y, x]), k))
    for y in range(H):
      for x in range(W):
        best = sorted(cand[y][x])[:L]
        got = [(float(z[l, y, x]), int(ids[l, y, x])) for l in range(L)
               if ids[l, y, x] >= 0]
        assert got == best
