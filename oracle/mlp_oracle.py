"""TEST INFRASTRUCTURE ONLY: CPU oracle for the learned deferred shader.

Only tests/, __graft_entry__.smoke() and bench.py may import this module.

The network (no reference counterpart: the reference has no learned shader)
is an MLP 12 -> 64 -> 64 -> 64 -> 64 -> 3 with ReLU on the hidden layers,
bfloat16 weights and activations as matrix-core operands and float32
accumulation.  This oracle quantises to bfloat16 (round to nearest even) at
exactly the points the GPU kernels do and accumulates in float64, so GPU and
oracle differ only by float32 accumulation order (and, rarely, by a bfloat16
rounding flip that this causes).  Its own pins: float64 central finite
differences of the unquantised network (tests/test_mlp_oracle.py), and the
Phong restatement against a direct single-light evaluation of
shaders.point_light (diffren/jax/utils/shaders.py:100-173).

Flat float32 parameter vector: W1[12,64] b1[64] W2[64,64] b2 W3 b3 W4 b4
W5[64,3] b5[3] (W[k_in, unit_out]).
"""
import numpy as np

SIZES = [12, 64, 64, 64, 64, 3]


def param_slices():
  out, o = [], 0
  for i in range(5):
    k, n = SIZES[i], SIZES[i + 1]
    out.append(((o, o + k * n), (k, n), (o + k * n, o + k * n + n)))
    o += k * n + n
  return out, o


def num_params():
  return param_slices()[1]


def unpack(params):
  p = np.asarray(params)
  if p.dtype != np.float64:
    p = p.astype(np.float32)
  sl, _ = param_slices()
  return [(p[a:b].reshape(shape), p[c:d]) for (a, b), shape, (c, d) in sl]


def init_params(seed=0):
  rng = np.random.default_rng(seed)
  p = np.zeros(num_params(), np.float32)
  sl, _ = param_slices()
  for (a, b), (k, n), (c, d) in sl:
    p[a:b] = (rng.normal(size=k * n) * np.sqrt(2.0 / k)).astype(np.float32)
  return p


def bf16(x):
  """float32 -> bfloat16 (round to nearest even) -> float32."""
  u = np.ascontiguousarray(x, np.float32).view(np.uint32).astype(np.uint64)
  u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
  return u.astype(np.uint32).view(np.float32)


def normalized(x, norm):
  """(x - shift) * scale in float32 (norm = [shift 12 | scale 12] or None)."""
  x = np.asarray(x, np.float32)
  if norm is None:
    return x
  norm = np.asarray(norm, np.float32)
  return ((x - norm[:12]) * norm[12:]).astype(np.float32)


def forward(params, x, keep=False, norm=None, modulate=False):
  """x [N,12] f32 -> y [N,3] f32 (and per-layer activations if keep).
  With modulate the image is albedo x[:, 9:12] * network output."""
  layers = unpack(params)
  a = bf16(normalized(x, norm)).astype(np.float64)
  acts, pre = [a], []
  for i, (W, b) in enumerate(layers):
    z = a @ bf16(W).astype(np.float64) + b.astype(np.float64)
    z = z.astype(np.float32)
    pre.append(z)
    if i < 4:
      a = bf16(np.maximum(z, 0.0)).astype(np.float64)
      acts.append(a)
  y = pre[-1]
  if modulate:
    y = (np.asarray(x, np.float32)[:, 9:12] * y).astype(np.float32)
  return (y, acts, pre) if keep else y


def loss_and_grad(params, x, target, mask, scale, norm=None, modulate=False):
  """L = scale * sum(mask * (y - t)^2); returns (L, dL/dparams flat f32)."""
  layers = unpack(params)
  y, acts, pre = forward(params, x, keep=True, norm=norm, modulate=modulate)
  m = np.asarray(mask, np.float32)[:, None]
  d = (y - target).astype(np.float64)
  loss = float(scale * np.sum(m * d * d))
  delta = (2.0 * d * scale * m).astype(np.float32)
  if modulate:
    delta = (delta * np.asarray(x, np.float32)[:, 9:12]).astype(np.float32)
  grads = [None] * 5
  for i in range(4, -1, -1):
    W, _ = layers[i]
    dq = bf16(delta).astype(np.float64)
    gW = acts[i].T @ dq
    gb = delta.astype(np.float64).sum(0)
    grads[i] = (gW.astype(np.float32), gb.astype(np.float32))
    if i > 0:
      da = dq @ bf16(W).astype(np.float64).T
      delta = (da * (pre[i - 1] > 0)).astype(np.float32)
  flat = np.concatenate([np.concatenate([g.reshape(-1), b]) for g, b in grads])
  return loss, flat.astype(np.float32)


def forward_f64(params, x):
  """Unquantised float64 network (for finite-difference pins)."""
  a = np.asarray(x, np.float64)
  for i, (W, b) in enumerate(unpack(params)):
    a = a @ W.astype(np.float64) + b
    if i < 4:
      a = np.maximum(a, 0.0)
  return a


def loss_and_grad_f64(params, x, target, mask, scale):
  layers = [(W.astype(np.float64), b.astype(np.float64))
            for W, b in unpack(params)]
  acts, pre = [np.asarray(x, np.float64)], []
  a = acts[0]
  for i, (W, b) in enumerate(layers):
    z = a @ W + b
    pre.append(z)
    a = np.maximum(z, 0.0) if i < 4 else z
    acts.append(a)
  m = np.asarray(mask, np.float64)[:, None]
  d = a - target
  loss = scale * np.sum(m * d * d)
  delta = 2.0 * d * scale * m
  grads = [None] * 5
  for i in range(4, -1, -1):
    grads[i] = (acts[i].T @ delta, delta.sum(0))
    if i > 0:
      delta = (delta @ layers[i][0].T) * (pre[i - 1] > 0)
  return loss, np.concatenate([np.concatenate([g.reshape(-1), b])
                               for g, b in grads])


def adam(params, grad, m, v, step, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
  """Adam (Kingma & Ba 2015) on float32 master weights; returns new arrays."""
  g = grad.astype(np.float32)
  m = (b1 * m + (1 - b1) * g).astype(np.float32)
  v = (b2 * v + (1 - b2) * g * g).astype(np.float32)
  mh = m / np.float32(1 - b1 ** step)
  vh = v / np.float32(1 - b2 ** step)
  return (params - lr * mh / (np.sqrt(vh) + eps)).astype(np.float32), m, v


def phong(gbuf, lights, ka=0.3, kd=1.4, ks=0.5, ns=16.0, att=(1.0, 0.7, 1.8)):
  """Many-light Phong in float64: gbuf [N,12] (pos, normal, view, albedo),
  lights [K,4] (xyz, intensity) -> [N,3].  Ambient counted once."""
  g = np.asarray(gbuf, np.float64)
  pos, n, vd, alb = g[:, 0:3], g[:, 3:6], g[:, 6:9], g[:, 9:12]
  lum = np.zeros(g.shape[0])
  for L in np.asarray(lights, np.float64):
    l = L[:3] - pos
    d2 = np.sum(l * l, 1)
    d = np.sqrt(d2)
    u = l / np.sqrt(d2 + 1e-12)[:, None]
    nol = np.clip(np.sum(n * u, 1), 0, 1)
    h = u + vd
    h = h / np.sqrt(np.sum(h * h, 1) + 1e-12)[:, None]
    noh = np.clip(np.sum(n * h, 1), 0, 1)
    decay = 1.0 / (att[0] + att[1] * d + att[2] * d2)
    lum += L[3] * decay * (kd * nol + ks * noh ** ns)
  return alb * (ka + lum)[:, None]


def point_light_single(diffuse, normals, positions, light, eye, ka, kd, ks, ns,
                       att=(1.0, 0.7, 1.8)):
  """Direct float64 evaluation of shaders.point_light (shaders.py:100-173)
  for one light and intensity 1, for pinning `phong` with K = 1:
  point_light = diffuse * (ka + kd nol + ks noh^ns) * decay."""
  l = light - positions
  ld = l / np.sqrt(np.sum(l * l, -1, keepdims=True) + 1e-12)
  v = eye - positions
  vd = v / np.sqrt(np.sum(v * v, -1, keepdims=True) + 1e-12)
  h = ld + vd
  h = h / np.sqrt(np.sum(h * h, -1, keepdims=True) + 1e-12)
  nol = np.clip(np.sum(normals * ld, -1, keepdims=True), 0, 1)
  noh = np.clip(np.sum(normals * h, -1, keepdims=True), 0, 1)
  dist = np.linalg.norm(l, axis=-1, keepdims=True)
  decay = 1.0 / (att[0] + att[1] * dist + att[2] * dist * dist)
  return diffuse * (ka + kd * nol + ks * noh ** ns) * decay
