"""ctypes binding of libdiffren_amd.so, the gfx950 C ABI (include/diffren_amd.h).

There is no fallback: if the library is missing or a call fails, this raises.
Device buffers are torch tensors on the current HIP device; work is enqueued on
torch's current stream.
"""
import ctypes
import os

import torch

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libdiffren_amd.so")

_c_int = ctypes.c_int32
_c_i64 = ctypes.c_int64
_c_ptr = ctypes.c_void_p
_c_size = ctypes.c_size_t


RASTER_SPARSE = 1
RASTER_BINS_ONLY = 2
RASTER_TILES_ONLY = 4
RASTER_PREFIX = 8
RENDER_ACCUM_BUFFERS = 4
RENDER_SHADED = 8
MLP_ALBEDO_MODULATE = 1
MLP_GRAD_OVERWRITE = 2
PAD_CONSTANT = 0
PAD_EDGE = 1


class RasterConfig(ctypes.Structure):
  _fields_ = [("tile_w", ctypes.c_int32), ("tile_h", ctypes.c_int32),
              ("bin_capacity", ctypes.c_int64), ("flags", ctypes.c_int32),
              ("tile_threads", ctypes.c_int32)]


class RasterJob(ctypes.Structure):
  _fields_ = [("vertices", _c_ptr), ("triangles", _c_ptr),
              ("triangle_ids", _c_ptr), ("z_buffer", _c_ptr),
              ("barycentrics", _c_ptr), ("camera", _c_ptr),
              ("num_vertices", _c_int),
              ("num_triangles", _c_int), ("width", _c_int),
              ("height", _c_int), ("triangle_begin", _c_i64),
              ("tile_begin", _c_i64)]


class PhongParams(ctypes.Structure):
  _fields_ = [("ka", ctypes.c_float), ("kd", ctypes.c_float),
              ("ks", ctypes.c_float), ("ns", ctypes.c_float),
              ("attenuation_constant", ctypes.c_float),
              ("attenuation_linear", ctypes.c_float),
              ("attenuation_quadratic", ctypes.c_float)]


class SplatJob(ctypes.Structure):
  _fields_ = [("clip_vertices", _c_ptr), ("positions", _c_ptr),
              ("triangles", _c_ptr),
              ("colors", _c_ptr), ("camera", _c_ptr),
              ("triangle_ids", _c_ptr), ("z_buffer", _c_ptr),
              ("barycentrics", _c_ptr), ("image", _c_ptr),
              ("target", _c_ptr), ("d_image", _c_ptr),
              ("grad_scratch", _c_ptr), ("d_records", _c_ptr),
              ("accum", _c_ptr), ("norm", _c_ptr), ("d_camera", _c_ptr),
              ("accum_weights", _c_ptr),
              ("shaded", _c_ptr), ("d_shaded", _c_ptr),
              ("num_vertices", _c_int),
              ("num_triangles", _c_int), ("width", _c_int),
              ("height", _c_int), ("tile_begin", _c_i64),
              ("fixed_shift", _c_int), ("reserved", _c_int)]


# name -> (restype, argtypes).  Keep in the order of include/diffren_amd.h.
SIGNATURES = {
    "dr_last_error": (ctypes.c_char_p, []),
    "dr_version": (ctypes.c_char_p, []),
    "dr_raster_jobs_prepare": (_c_int, [_c_ptr, _c_int, _c_ptr, _c_ptr,
                                        _c_ptr]),
    "dr_raster_workspace_size": (_c_size, [_c_i64, _c_i64, _c_ptr]),
    "dr_rasterize_triangles": (_c_int, [_c_ptr, _c_int, _c_ptr, _c_int,
                                        _c_int, _c_int, _c_int, _c_int,
                                        _c_ptr, _c_ptr, _c_ptr, _c_ptr,
                                        _c_ptr, _c_size, _c_ptr]),
    "dr_rasterize_triangles_noalloc": (_c_int, [_c_ptr, _c_int, _c_ptr,
                                                _c_int, _c_int, _c_int,
                                                _c_int, _c_int, _c_ptr,
                                                _c_ptr, _c_ptr, _c_ptr,
                                                _c_ptr]),
    "dr_rasterize_triangles_batched": (_c_int, [_c_ptr, _c_int, _c_i64,
                                                _c_i64, _c_int, _c_int,
                                                _c_ptr, _c_ptr, _c_size,
                                                _c_ptr]),
    "dr_raster_bin_stats": (_c_int, [_c_ptr, _c_ptr, _c_ptr, _c_ptr]),
    "dr_transform_homogeneous": (_c_int, [_c_ptr, _c_ptr, _c_int, _c_i64,
                                          _c_ptr, _c_ptr]),
    "dr_transform_homogeneous_multi": (_c_int, [_c_ptr, _c_int, _c_ptr,
                                                _c_int, _c_i64, _c_ptr,
                                                _c_ptr]),
    "dr_transform_homogeneous_bwd": (_c_int, [_c_ptr, _c_ptr, _c_int, _c_i64,
                                              _c_ptr, _c_ptr]),
    "dr_diff_barycentrics_fwd": (_c_int, [_c_ptr, _c_i64, _c_ptr, _c_int,
                                          _c_ptr, _c_int, _c_int, _c_int,
                                          _c_ptr, _c_ptr]),
    "dr_diff_barycentrics_bwd": (_c_int, [_c_ptr, _c_i64, _c_ptr, _c_int,
                                          _c_ptr, _c_int, _c_int, _c_int,
                                          _c_ptr, _c_ptr, _c_ptr]),
    "dr_interpolate_fwd": (_c_int, [_c_ptr, _c_ptr, _c_ptr, _c_int, _c_ptr,
                                    _c_int, _c_int, _c_ptr, _c_i64, _c_ptr,
                                    _c_ptr]),
    "dr_interpolate_bwd": (_c_int, [_c_ptr, _c_ptr, _c_ptr, _c_int, _c_ptr,
                                    _c_int, _c_int, _c_i64, _c_ptr, _c_ptr]),
    "dr_interpolate_bwd_barycentrics": (_c_int, [_c_ptr, _c_ptr, _c_int,
                                                 _c_ptr, _c_int, _c_int,
                                                 _c_ptr, _c_i64, _c_ptr,
                                                 _c_ptr]),
    "dr_splat_coefficient": (ctypes.c_float, []),
    "dr_splat_fwd": (_c_int, [_c_ptr, _c_ptr, _c_ptr, _c_int, _c_int, _c_int,
                              _c_int, _c_ptr, _c_ptr, _c_ptr, _c_ptr, _c_ptr]),
    "dr_splat_bwd_workspace_size": (_c_size, [_c_int, _c_int, _c_int]),
    "dr_splat_bwd": (_c_int, [_c_ptr, _c_ptr, _c_ptr, _c_int, _c_int, _c_int,
                              _c_int, _c_ptr, _c_ptr, _c_ptr, _c_ptr, _c_size,
                              _c_ptr]),
    "dr_bilinear_resample_fwd": (_c_int, [_c_ptr, _c_int, _c_int, _c_int,
                                          _c_ptr, _c_i64, _c_int, _c_ptr,
                                          _c_ptr]),
    "dr_bilinear_resample_bwd": (_c_int, [_c_ptr, _c_int, _c_int, _c_int,
                                          _c_ptr, _c_i64, _c_int, _c_ptr,
                                          _c_ptr, _c_ptr, _c_ptr]),
    "dr_gbuffer": (_c_int, [_c_ptr, _c_ptr, _c_ptr, _c_int, _c_ptr, _c_ptr,
                            _c_ptr, _c_int, _c_ptr, _c_i64, _c_ptr, _c_ptr]),
    "dr_phong_shade": (_c_int, [_c_ptr, _c_i64, _c_ptr, _c_int, _c_ptr,
                                _c_ptr, _c_ptr]),
    "dr_mlp_num_params": (_c_int, []),
    "dr_mlp_packed_bytes": (_c_size, []),
    "dr_mlp_pack": (_c_int, [_c_ptr, _c_ptr, _c_ptr]),
    "dr_mlp_forward": (_c_int, [_c_ptr, _c_ptr, _c_ptr, _c_i64, _c_ptr,
                                _c_ptr, _c_int, _c_int, _c_ptr]),
    "dr_mlp_train_workspace_size": (_c_size, [_c_int]),
    "dr_mlp_train_step": (_c_int, [_c_ptr, _c_ptr, _c_ptr, _c_ptr, _c_ptr,
                                   _c_i64, ctypes.c_float, _c_ptr, _c_int,
                                   _c_ptr, _c_ptr, _c_int, _c_ptr, _c_size,
                                   _c_ptr]),
    "dr_mlp_adam": (_c_int, [_c_ptr, _c_ptr, _c_ptr, _c_ptr, _c_int,
                             ctypes.c_float, ctypes.c_float, ctypes.c_float,
                             ctypes.c_float, _c_ptr]),
    "dr_splat_jobs_prepare": (_c_int, [_c_ptr, _c_int, _c_int, _c_ptr]),
    "dr_render_scratch_floats": (_c_int, [_c_int]),
    "dr_render_record_size": (_c_int, [_c_int]),
    "dr_render_splat_workspace_size": (_c_size, [_c_i64]),
    "dr_render_splat_forward": (_c_int, [_c_ptr, _c_int, _c_i64, _c_int,
                                         _c_int, _c_int, ctypes.c_float,
                                         _c_ptr, _c_ptr, _c_size, _c_ptr]),
    "dr_fixed_to_float": (_c_int, [_c_ptr, _c_i64, _c_int, _c_ptr, _c_ptr]),
    "dr_render_splat_backward": (_c_int, [_c_ptr, _c_int, _c_i64, _c_int,
                                          _c_int, _c_int, _c_ptr]),
}


def table_to_device(table, device):
  """Copies a ctypes array of job structs into a device byte tensor."""
  host = torch.frombuffer(bytearray(bytes(table)), dtype=torch.uint8)
  return host.to(device)

_lib = None


class NativeError(RuntimeError):
  pass


def lib():
  """Loads the library once; raises if it has not been built."""
  global _lib
  if _lib is None:
    if not os.path.exists(LIB_PATH):
      raise NativeError(
          f"{LIB_PATH} is missing: build it with "
          "`python -m diffren_amd.build_native` (there is no CPU fallback)")
    l = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
      fn = getattr(l, name)
      fn.restype = res
      fn.argtypes = args
    _lib = l
  return _lib


def check(rc, what):
  if rc != 0:
    msg = lib().dr_last_error().decode()
    raise NativeError(f"{what} failed (status {rc}): {msg}")


def ptr(t):
  """Device pointer of a tensor (None -> NULL)."""
  return None if t is None else ctypes.c_void_p(t.data_ptr())


def stream_handle(stream=None):
  s = torch.cuda.current_stream() if stream is None else stream
  return ctypes.c_void_p(s.cuda_stream)


def require_device(*tensors):
  for t in tensors:
    if t is not None and not t.is_cuda:
      raise ValueError("diffren_amd kernels take device tensors; got a "
                       f"tensor on {t.device}")


def workspace(nbytes, device):
  """Caller-owned scratch from torch's caching allocator (256-byte aligned)."""
  n = max(int(nbytes), 256)
  return torch.empty(n, dtype=torch.uint8, device=device)
