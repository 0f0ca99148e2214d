"""Fixed test scenes shared by the golden-vector script and the tests.

Clip-space vertex data below are the reference test fixtures:
  CUBE_VERTICES / CUBE_TRIANGLES     rasterize_test_utils.h:33-45
  LARGE_CUBE_*                       rasterize_test_utils.h:47-54
  single-triangle scenes             rasterize_triangles_impl_test.cc:125-225
"""
import numpy as np

CUBE_VERTICES = np.array([
    -2.60648608, -3.22707772, 6.85085106, 6.85714293,
    -1.30324292, -0.992946863, 8.56856918, 8.5714283,
    -1.30324292, 3.97178817, 7.70971, 7.71428585,
    -2.60648608, 1.73765731, 5.991992, 6,
    1.30324292, -3.97178817, 6.27827835, 6.28571415,
    2.60648608, -1.73765731, 7.99599648, 8,
    2.60648608, 3.22707772, 7.13713741, 7.14285707,
    1.30324292, 0.992946863, 5.41941929, 5.4285717], np.float32).reshape(-1, 4)
CUBE_TRIANGLES = np.array([0, 1, 2, 2, 3, 0, 3, 2, 6, 6, 7, 3, 7, 6, 5, 5, 4,
                           7, 4, 5, 1, 1, 0, 4, 5, 6, 2, 2, 1, 5, 7, 4, 0, 0,
                           3, 7], np.int32).reshape(-1, 3)
LARGE_CUBE_VERTICES = np.array([
    -1.0, -1.0, 1.98, 2.0, 1.0, -1.0, 1.98, 2.0, 1.0, 1.0, 1.98, 2.0, -1.0,
    1.0, 1.98, 2.0, -1.0, -1.0, -2.02, -2.0, 1.0, -1.0, -2.02, -2.0, 1.0, 1.0,
    -2.02, -2.0, -1.0, 1.0, -2.02, -2.0], np.float32).reshape(-1, 4)
LARGE_CUBE_TRIANGLES = np.array([0, 1, 2, 0, 2, 3, 4, 5, 6, 4, 6, 7, 2, 3, 7,
                                 2, 7, 6, 1, 0, 4, 1, 4, 5, 0, 3, 7, 0, 7, 4,
                                 1, 2, 6, 1, 6, 5], np.int32).reshape(-1, 3)

SIMPLE_TRIANGLE = np.array([[-0.5, -0.5, 0.8, 1.0], [0.0, 0.5, 0.3, 1.0],
                            [0.5, -0.5, 0.3, 1.0]], np.float32)
EXTERNAL_TRIANGLE = np.array([[-0.5, -0.5, 0.99, 1.0], [0.0, -0.5, -1.01, -1.0],
                              [0.5, -0.5, 0.99, 1.0]], np.float32)
TETRAHEDRON = np.array([[-0.5, -0.5, 0.8, 1.0], [0.0, 0.5, 0.3, 1.0],
                        [0.5, -0.5, 0.3, 1.0], [0.0, 0.0, 0.0, 1.0]],
                       np.float32)
TETRAHEDRON_TRIANGLES = np.array([[0, 2, 1], [0, 1, 3], [1, 2, 3], [2, 0, 3]],
                                 np.int32)
EDGE_TRIANGLE = np.array([[0.0, -1.0, 0.5, 1.0], [0.0, 1.0, 0.5, 1.0],
                          [0.5, -1.0, 0.5, 1.0]], np.float32)
FULL_QUAD = np.array([[-1, -1, 0, 1], [1, -1, 0, 1], [1, 1, 0, 1],
                      [-1, 1, 0, 1]], np.float32)
DEGENERATE = np.array([[-1, -1, 0, 1], [1, 1, 0, 1], [0, 0, 0, 1]], np.float32)


def adversarial_scene(seed, num_vertices=40, num_triangles=80):
  """Random clip-space soup with near-plane crossings, tiny/zero/negative w,
  huge x/y and vertices exactly on the near plane."""
  rng = np.random.default_rng(seed)
  v = rng.normal(size=(num_vertices, 4)).astype(np.float32)
  v[:, 3] = rng.uniform(-0.5, 3, size=num_vertices)
  k = rng.integers(0, 6, size=num_vertices)
  v[k == 0, 3] = rng.choice([1e-30, -1e-30, 0.0, 1e-6], size=(k == 0).sum())
  v[k == 1, :2] *= rng.choice([1e3, 1e5, 1e8, 1e12],
                              size=((k == 1).sum(), 1))
  v[k == 2, 2] = -v[k == 2, 3]
  t = rng.integers(0, num_vertices, size=(num_triangles, 3)).astype(np.int32)
  return v, t


def reference_uv_sphere(n_rows, n_cols, radius):
  """UV sphere with the vertex and triangle order of
  rasterize_test_utils.cc:34-118: top pole, n_rows rings of n_cols vertices
  (phi = pi (r + 1) / (n_rows + 1), theta = 2 pi c / n_cols), bottom pole;
  top fan, two triangles per quad, bottom fan.  Returns clip-space
  vertices [V, 4] (w = 1, AddHomogeneousCoord :120-132) and triangles."""
  r = np.arange(n_rows)[:, None]
  c = np.arange(n_cols)[None, :]
  phi = np.broadcast_to((np.pi * (r + 1) / (n_rows + 1)).astype(np.float32),
                        (n_rows, n_cols))
  theta = np.broadcast_to((2.0 * np.pi * c / n_cols).astype(np.float32),
                          (n_rows, n_cols))
  rad = np.float32(radius)
  ring = np.stack([rad * np.sin(phi) * np.sin(theta), rad * np.cos(phi),
                   rad * np.sin(phi) * np.cos(theta)], -1).reshape(-1, 3)
  xyz = np.concatenate([[[0.0, radius, 0.0]], ring,
                        [[0.0, -radius, 0.0]]]).astype(np.float32)
  n_verts = xyz.shape[0]
  nxt = (np.arange(n_cols) + 1) % n_cols
  cols = np.arange(n_cols)
  tris = [np.stack([np.zeros(n_cols, int), 1 + cols, 1 + nxt], 1)]
  for ri in range(n_rows - 1):
    tl = 1 + ri * n_cols + cols
    tr = 1 + ri * n_cols + nxt
    bl = 1 + (ri + 1) * n_cols + cols
    br = 1 + (ri + 1) * n_cols + nxt
    quad = np.stack([tl, bl, br, tr, tl, br], 1).reshape(-1, 3)
    tris.append(quad)
  last = n_verts - 1 - n_cols
  tris.append(np.stack([last + cols, np.full(n_cols, n_verts - 1),
                        last + nxt], 1))
  v = np.concatenate([xyz, np.ones((n_verts, 1), np.float32)], 1)
  return v, np.concatenate(tris).astype(np.int32)


def reference_cases():
  """(name, vertices[V,4], triangles[T,3], W, H, L, cull) for every C++ test
  scene of rasterize_triangles_impl_test.cc:125-294."""
  tri = np.array([[0, 1, 2]], np.int32)
  cases = [
      ("simple_triangle", SIMPLE_TRIANGLE, tri, 640, 480, 1, 0),
      ("perspective_triangle",
       SIMPLE_TRIANGLE * np.array([[0.2], [0.5], [2.0]], np.float32), tri,
       640, 480, 1, 0),
      ("external_triangle", EXTERNAL_TRIANGLE, tri, 640, 480, 1, 0),
      ("inside_box", LARGE_CUBE_VERTICES, LARGE_CUBE_TRIANGLES, 640, 480, 1, 0),
      ("tetrahedron", TETRAHEDRON, TETRAHEDRON_TRIANGLES, 640, 480, 1, 0),
      ("edge_pixel", EDGE_TRIANGLE, tri, 641, 480, 1, 0),
      ("edge_pixel_reversed", EDGE_TRIANGLE, tri[:, ::-1].copy(), 641, 480, 1,
       0),
      ("image_edges", FULL_QUAD, np.array([[0, 1, 2], [0, 2, 3]], np.int32),
       640, 480, 1, 0),
      ("degenerate", DEGENERATE, tri, 1, 1, 1, 0),
      ("cube_two_layers", CUBE_VERTICES, CUBE_TRIANGLES, 640, 480, 2, 0),
      ("cube_cull_back", CUBE_VERTICES, CUBE_TRIANGLES, 640, 480, 2, 1),
      ("cube_cull_front", CUBE_VERTICES, CUBE_TRIANGLES, 640, 480, 2, 2),
  ]
  return cases
