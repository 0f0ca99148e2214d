"""Seeded synthetic scenes for the rasterize-then-splat workloads.

Every scene is built with numpy.random.default_rng(seed) so tests, the bench
and the CPU baseline see identical inputs.  Workloads:

  icosphere_scene   icosphere (3 subdivisions: 642 vertices, 1280 triangles),
                    radial noise U(+-5%), random rotation, 3 units in front of
                    a 60 degree perspective camera, RGB colours + alpha 1.
  uv_sphere_mesh    UV sphere with `rows` x `cols` vertices plus 2 poles,
                    2*rows*cols triangles (rows=500, cols=1000 gives 1,000,000
                    triangles and 500,002 vertices), displaced by 4-octave
                    seeded Perlin noise of amplitude 0.1.
  orbit_cameras     perspective x look-at matrices on a radius-3 orbit.

The sphere topology follows the UV-sphere fixture of the reference tests
(diffren/common/kernels/rasterize_test_utils.cc:34-118): one vertex at each
pole, `rows` rings of `cols` vertices, counter-clockwise triangles.
"""
import math

import numpy as np

from diffren_amd import camera

NEAR_CLIP = 0.1
FAR_CLIP = 10.0
FOV_Y = 60.0


# ---------------------------------------------------------------------------
# Meshes


def icosphere(subdivisions):
  """Unit icosphere; returns (vertices [V,3] f64, triangles [T,3] i32)."""
  t = (1.0 + math.sqrt(5.0)) / 2.0
  verts = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t),
           (0, 1, t), (0, -1, -t), (0, 1, -t), (t, 0, -1), (t, 0, 1),
           (-t, 0, -1), (-t, 0, 1)]
  faces = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11),
           (1, 5, 9), (5, 11, 4), (11, 10, 2), (10, 7, 6), (7, 1, 8),
           (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5),
           (2, 4, 11), (6, 2, 10), (8, 6, 7), (9, 8, 1)]
  verts = [np.array(v, np.float64) / np.linalg.norm(v) for v in verts]
  for _ in range(subdivisions):
    cache = {}
    new_faces = []

    def mid(a, b):
      key = (min(a, b), max(a, b))
      if key not in cache:
        m = verts[a] + verts[b]
        verts.append(m / np.linalg.norm(m))
        cache[key] = len(verts) - 1
      return cache[key]

    for a, b, c in faces:
      ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
      new_faces += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
    faces = new_faces
  return np.array(verts), np.array(faces, np.int32)


def uv_sphere(rows, cols):
  """Unit UV sphere; (vertices [rows*cols+2, 3] f64, triangles [2*rows*cols, 3])."""
  ri = np.arange(rows)
  ci = np.arange(cols)
  phi = (math.pi * (ri + 1)) / (rows + 1)
  theta = 2.0 * math.pi * ci / cols
  sp, cp = np.sin(phi)[:, None], np.cos(phi)[:, None]
  st, ct = np.sin(theta)[None, :], np.cos(theta)[None, :]
  ring = np.stack([sp * st, np.broadcast_to(cp, (rows, cols)), sp * ct],
                  axis=-1).reshape(-1, 3)
  verts = np.concatenate([[[0.0, 1.0, 0.0]], ring, [[0.0, -1.0, 0.0]]], 0)
  nv = verts.shape[0]
  cnext = (ci + 1) % cols
  top = np.stack([np.zeros(cols, np.int64), 1 + ci, 1 + cnext], -1)
  r = np.arange(rows - 1)[:, None]
  tl = 1 + r * cols + ci[None, :]
  tr = 1 + r * cols + cnext[None, :]
  bl = 1 + (r + 1) * cols + ci[None, :]
  br = 1 + (r + 1) * cols + cnext[None, :]
  quads = np.stack([np.stack([tl, bl, br], -1),
                    np.stack([tr, tl, br], -1)], axis=2).reshape(-1, 3)
  last = nv - 1 - cols
  bottom = np.stack([last + ci, np.full(cols, nv - 1), last + cnext], -1)
  tris = np.concatenate([top, quads, bottom], 0).astype(np.int32)
  return verts, tris


def _perlin_perm(rng):
  p = rng.permutation(256)
  return np.concatenate([p, p]).astype(np.int64)


def _grad(h, x, y, z):
  h = h & 15
  u = np.where(h < 8, x, y)
  v = np.where(h < 4, y, np.where((h == 12) | (h == 14), x, z))
  return np.where(h & 1, -u, u) + np.where(h & 2, -v, v)


def perlin3(p, perm):
  """Improved Perlin noise (K. Perlin, SIGGRAPH 2002) at points p [N,3]."""
  fl = np.floor(p)
  i = fl.astype(np.int64) & 255
  f = p - fl
  fade = f * f * f * (f * (f * 6 - 15) + 10)
  X, Y, Z = i[:, 0], i[:, 1], i[:, 2]
  x, y, z = f[:, 0], f[:, 1], f[:, 2]
  A = perm[X] + Y
  AA, AB = perm[A] + Z, perm[A + 1] + Z
  B = perm[X + 1] + Y
  BA, BB = perm[B] + Z, perm[B + 1] + Z
  u, v, w = fade[:, 0], fade[:, 1], fade[:, 2]
  lerp = lambda t, a, b: a + t * (b - a)
  return lerp(w,
              lerp(v, lerp(u, _grad(perm[AA], x, y, z),
                           _grad(perm[BA], x - 1, y, z)),
                   lerp(u, _grad(perm[AB], x, y - 1, z),
                        _grad(perm[BB], x - 1, y - 1, z))),
              lerp(v, lerp(u, _grad(perm[AA + 1], x, y, z - 1),
                           _grad(perm[BA + 1], x - 1, y, z - 1)),
                   lerp(u, _grad(perm[AB + 1], x, y - 1, z - 1),
                        _grad(perm[BB + 1], x - 1, y - 1, z - 1))))


def fbm(p, perm, octaves=4, base_frequency=2.0):
  """Sum of `octaves` Perlin octaves, frequency x2 and amplitude x0.5 each."""
  out = np.zeros(p.shape[0])
  for o in range(octaves):
    out += perlin3(p * (base_frequency * 2**o), perm) * 0.5**o
  return out


def vertex_normals(verts, tris):
  """Area-weighted vertex normals, unit length (float64)."""
  v0, v1, v2 = verts[tris[:, 0]], verts[tris[:, 1]], verts[tris[:, 2]]
  fn = np.cross(v1 - v0, v2 - v0)
  n = np.zeros_like(verts)
  for k in range(3):
    for c in range(3):
      n[:, c] += np.bincount(tris[:, k], weights=fn[:, c],
                             minlength=verts.shape[0])
  return n / np.maximum(np.linalg.norm(n, axis=1, keepdims=True), 1e-20)


def uv_sphere_mesh(rows=500, cols=1000, amplitude=0.1, seed=0):
  """Perlin-displaced UV sphere: (vertices f32 [V,3], triangles i32 [T,3])."""
  rng = np.random.default_rng(seed)
  perm = _perlin_perm(rng)
  verts, tris = uv_sphere(rows, cols)
  r = 1.0 + amplitude * fbm(verts, perm)
  return (verts * r[:, None]).astype(np.float32), tris


def random_rotation(rng):
  """Uniform random rotation from a uniform unit quaternion."""
  q = rng.normal(size=4)
  q /= np.linalg.norm(q)
  w, x, y, z = q
  return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w),
                    2 * (x * z + y * w)],
                   [2 * (x * y + z * w), 1 - 2 * (x * x + z * z),
                    2 * (y * z - x * w)],
                   [2 * (x * z - y * w), 2 * (y * z + x * w),
                    1 - 2 * (x * x + y * y)]])


# ---------------------------------------------------------------------------
# Cameras


def perspective_camera(aspect=1.0, fov_y=FOV_Y, near=NEAR_CLIP, far=FAR_CLIP):
  return camera.perspective(aspect, fov_y, near, far)


def orbit_cameras(n, radius=3.0, seed=0, aspect=1.0):
  """n seeded poses on a sphere of `radius` looking at the origin; [n,4,4]."""
  rng = np.random.default_rng(seed)
  proj = perspective_camera(aspect)
  mats = []
  for _ in range(n):
    d = rng.normal(size=3)
    d /= np.linalg.norm(d)
    up = np.array([0.0, 1.0, 0.0])
    if abs(d[1]) > 0.99:
      up = np.array([1.0, 0.0, 0.0])
    mats.append(proj @ camera.look_at(d * radius, np.zeros(3), up))
  return np.stack(mats).astype(np.float32)


def to_clip(vertices, cam):
  """[V,3] -> [V,4] clip space, float32 result of a float64 product."""
  v = np.concatenate([np.asarray(vertices, np.float64),
                      np.ones((vertices.shape[0], 1))], 1)
  return (v @ np.asarray(cam, np.float64).T).astype(np.float32)


# ---------------------------------------------------------------------------
# Workload instances


def icosphere_scene(seed=0, subdivisions=3):
  """Small scene: returns dict(vertices [V,3], triangles, colors [V,4], camera)."""
  rng = np.random.default_rng(seed)
  v, t = icosphere(subdivisions)
  v = v * (1.0 + rng.uniform(-0.05, 0.05, size=(v.shape[0], 1)))
  v = v @ random_rotation(rng).T
  v = v + np.array([0.0, 0.0, -3.0])
  cam = perspective_camera()
  rgb = rng.uniform(0.0, 1.0, size=(v.shape[0], 3))
  colors = np.concatenate([rgb, np.ones((v.shape[0], 1))], 1)
  return dict(vertices=v.astype(np.float32), triangles=t,
              colors=colors.astype(np.float32), camera=cam.astype(np.float32))


def sphere_features(vertices, triangles, channels=8, seed=0):
  """Normals [V,3] plus `channels` N(0,1) features, float32."""
  rng = np.random.default_rng(seed + 1)
  n = vertex_normals(vertices.astype(np.float64), triangles)
  f = rng.normal(size=(vertices.shape[0], channels))
  return n.astype(np.float32), f.astype(np.float32)


def batching_workload(num_meshes=64, t_min=1_000, t_max=500_000, seed=0,
                      resolutions=(128, 256, 512, 1024)):
  """Independent noise-displaced spheres with log-uniform triangle counts,
  each with its own random orbit camera and square resolution.

  Returns a list of dicts: vertices [V,3], triangles [T,3], colors [V,4]
  (RGB uniform, alpha 1), camera [4,4], resolution.  Triangle counts are
  2 * rows * cols with cols = 2 * rows, the UV sphere closest to the draw.
  """
  rng = np.random.default_rng(seed)
  out = []
  for i in range(num_meshes):
    t = math.exp(rng.uniform(math.log(t_min), math.log(t_max)))
    rows = max(2, int(round(math.sqrt(t / 4.0))))
    v, tris = uv_sphere_mesh(rows, 2 * rows, seed=seed * 1000 + i)
    col = np.concatenate([rng.uniform(0, 1, (v.shape[0], 3)),
                          np.ones((v.shape[0], 1))], 1).astype(np.float32)
    cam = orbit_cameras(1, seed=seed * 1000 + i)[0]
    res = int(rng.choice(resolutions))
    out.append(dict(vertices=v, triangles=tris, colors=col, camera=cam,
                    resolution=res))
  return out
