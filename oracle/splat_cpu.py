"""TEST INFRASTRUCTURE ONLY: ctypes loader for liboracle_splat.so.

`render` runs splat_cpu.c: the float32 render-splat of one view (pass-through
shader, rasterization given), its L2 loss against a target and the gradients
with respect to vertex positions and colours, on `threads` host threads.  It
restates diffren/jax/render.py:107-132 with splat.py:26-229,
interpolate.py:24-63, composite.py:20-37 and transforms.py:40-78; the bench's
CPU baseline times it, and tests compare it with splat_oracle.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liboracle_splat.so")
_fn = None


def _load():
  global _fn
  if _fn is None:
    if not os.path.exists(_LIB):
      raise FileNotFoundError(_LIB)
    fn = ctypes.CDLL(_LIB).splat_cpu_render
    fn.restype = ctypes.c_double
    p, i = ctypes.c_void_p, ctypes.c_int32
    fn.argtypes = [p, p, p, i, i, p, i, p, p, i, i, i, p, ctypes.c_float, p,
                   p, p, i]
    _fn = fn
  return _fn


def render(positions, colors, camera, triangles, ids, barys, target=None,
           loss_scale=1.0, threads=0, gradients=True):
  """Returns (image [H,W,C], loss, d_positions [V,3], d_colors [V,C]); the
  loss and gradients are None without a target."""
  pos = np.ascontiguousarray(positions, np.float32)
  col = np.ascontiguousarray(colors, np.float32)
  cam = np.ascontiguousarray(camera, np.float32)
  tri = np.ascontiguousarray(triangles, np.int32)
  ids = np.ascontiguousarray(ids, np.int32)
  b = np.ascontiguousarray(barys, np.float32)
  L, H, W = ids.shape
  V, C = col.shape
  assert pos.shape == (V, 3) and cam.shape == (4, 4) and b.shape == (L, H, W, 3)
  img = np.empty((H, W, C), np.float32)
  tgt = None if target is None else np.ascontiguousarray(target, np.float32)
  dp = np.empty((V, 3), np.float32) if tgt is not None and gradients else None
  dc = np.empty((V, C), np.float32) if tgt is not None and gradients else None
  addr = lambda a: None if a is None else a.ctypes.data
  loss = _load()(pos.ctypes.data, col.ctypes.data, cam.ctypes.data, V, C,
                 tri.ctypes.data, tri.shape[0], ids.ctypes.data, b.ctypes.data,
                 L, H, W, addr(tgt), float(loss_scale), img.ctypes.data,
                 addr(dp), addr(dc), int(threads))
  if loss < 0:
    raise ValueError("splat_cpu_render: unsupported sizes")
  return img, (None if tgt is None else loss), dp, dc
