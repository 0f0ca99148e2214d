"""TEST INFRASTRUCTURE ONLY: ctypes loaders for the rasterizer oracles.

`rasterize`      -> liboracle_raster.so, the C restatement in raster_oracle.c
                    of diffren/common/kernels/rasterize_triangles_impl.cc:27-118.
`ref_rasterize`  -> _ref/libref_raster.so, the reference source itself compiled
                    from its own files (only where it was built).

Both return (ids [L,H,W] i32, z [L,H,W] f32, barycentrics [L,H,W,3] f32) with
the reference CPU handler's background values (ids -1, z 1, barycentrics 0,
rasterize_triangles_cpu.cc:55-60).
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_ORACLE = os.path.join(_HERE, "liboracle_raster.so")
_REF = os.path.join(_HERE, "_ref", "libref_raster.so")
_libs = {}


def _load(path, sym):
  if path not in _libs:
    if not os.path.exists(path):
      raise FileNotFoundError(path)
    lib = ctypes.CDLL(path)
    fn = getattr(lib, sym)
    fn.restype = None
    fn.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int32,
                   ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                   ctypes.c_int32, ctypes.c_void_p, ctypes.c_void_p,
                   ctypes.c_void_p]
    _libs[path] = fn
  return _libs[path]


def ref_available():
  return os.path.exists(_REF)


def _run(fn, vertices, triangles, width, height, layers, culling_mode,
         with_barys=True):
  v = np.ascontiguousarray(vertices, dtype=np.float32)
  t = np.ascontiguousarray(triangles, dtype=np.int32)
  assert v.ndim == 2 and v.shape[1] == 4 and t.ndim == 2 and t.shape[1] == 3
  ids = np.empty((layers, height, width), np.int32)
  z = np.empty((layers, height, width), np.float32)
  b = np.empty((layers, height, width, 3), np.float32) if with_barys else None
  fn(v.ctypes.data, t.ctypes.data, t.shape[0], width, height, layers,
     int(culling_mode), ids.ctypes.data, z.ctypes.data,
     b.ctypes.data if b is not None else None)
  return ids, z, b


def rasterize(vertices, triangles, width, height, layers, culling_mode,
              with_barys=True):
  return _run(_load(_ORACLE, "oracle_rasterize"), vertices, triangles, width,
              height, layers, culling_mode, with_barys)


def ref_rasterize(vertices, triangles, width, height, layers, culling_mode,
                  with_barys=True):
  return _run(_load(_REF, "ref_rasterize"), vertices, triangles, width,
              height, layers, culling_mode, with_barys)
