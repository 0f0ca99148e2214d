"""Pins of the learned-shader oracle (oracle/mlp_oracle.py) on the CPU."""
import numpy as np

from oracle import mlp_oracle as mo


def test_param_layout():
  sl, n = mo.param_slices()
  assert n == 13507
  assert [s[1] for s in sl] == [(12, 64), (64, 64), (64, 64), (64, 64),
                                (64, 3)]
  # FLOPs per pixel of the forward pass.
  assert 2 * sum(k * n for (_, (k, n), _) in sl) == 26496


def test_bf16_rounding():
  x = np.array([1.0, 1.00390625, 1.005859375, -3.0e-3, 65504.0], np.float32)
  q = mo.bf16(x)
  assert q[0] == 1.0
  assert q[1] == 1.0            # tie rounds to even (mantissa ...0)
  assert q[2] == 1.0078125      # above the tie rounds up
  assert np.all(np.abs(q - x) <= np.abs(x) * 2.0 ** -8)


def test_float64_gradient_matches_finite_differences():
  rng = np.random.default_rng(0)
  p = mo.init_params(1).astype(np.float64)
  x = rng.normal(size=(40, 12))
  t = rng.normal(size=(40, 3))
  mask = (rng.uniform(size=40) > 0.2).astype(np.float64)
  loss, g = mo.loss_and_grad_f64(p, x, t, mask, 0.01)
  idx = rng.choice(p.size, 64, replace=False)
  eps = 1e-6
  for i in idx:
    pp, pm = p.copy(), p.copy()
    pp[i] += eps
    pm[i] -= eps
    fd = (mo.loss_and_grad_f64(pp, x, t, mask, 0.01)[0] -
          mo.loss_and_grad_f64(pm, x, t, mask, 0.01)[0]) / (2 * eps)
    assert abs(fd - g[i]) <= 1e-6 + 1e-5 * abs(fd)


def test_quantised_gradient_close_to_float64():
  rng = np.random.default_rng(1)
  p = mo.init_params(2)
  x = rng.normal(size=(256, 12)).astype(np.float32)
  t = rng.normal(size=(256, 3)).astype(np.float32)
  mask = np.ones(256, np.float32)
  _, gq = mo.loss_and_grad(p, x, t, mask, 1.0 / 768)
  _, g64 = mo.loss_and_grad_f64(p, x, t, mask, 1.0 / 768)
  # bfloat16 operands: ~3 significant digits.
  assert np.linalg.norm(gq - g64) <= 2e-2 * np.linalg.norm(g64)


def test_phong_single_light_matches_point_light():
  rng = np.random.default_rng(3)
  n = 100
  pos = rng.normal(size=(n, 3))
  nrm = rng.normal(size=(n, 3))
  nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
  eye = np.array([0.3, -0.2, 4.0])
  vd = eye - pos
  vd /= np.linalg.norm(vd, axis=1, keepdims=True)
  alb = rng.uniform(size=(n, 3))
  light = np.array([2.0, 3.0, 1.0])
  g = np.concatenate([pos, nrm, vd, alb], 1)
  ours = mo.phong(g, np.array([[*light, 1.0]]), ka=0.3, kd=1.4, ks=0.5,
                  ns=16.0)
  # shaders.point_light multiplies the ambient term by the decay too; the
  # many-light form counts ambient once, undecayed.
  full = mo.point_light_single(alb, nrm, pos, light, eye, 0.3, 1.4, 0.5, 16.0)
  amb = mo.point_light_single(alb, nrm, pos, light, eye, 0.3, 0.0, 0.0, 16.0)
  np.testing.assert_allclose(ours - 0.3 * alb, full - amb, rtol=1e-12,
                             atol=1e-14)


def test_adam_first_step_is_sign_times_lr():
  p = np.zeros(5, np.float32)
  g = np.array([1.0, -2.0, 0.5, 3.0, -0.1], np.float32)
  out, _, _ = mo.adam(p, g, np.zeros(5, np.float32), np.zeros(5, np.float32),
                      1, lr=0.01)
  np.testing.assert_allclose(out, -0.01 * np.sign(g), rtol=1e-5)
