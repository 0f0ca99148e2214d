"""Enums of diffren's public API (diffren/jax/constants.py:20-30).

The integer values are part of the C ABI: FaceCullingMode must equal
diffren::FaceCullingMode (diffren/common/kernels/rasterize_triangles_types.h:24).
"""
import enum


class FaceCullingMode(enum.IntEnum):
  NONE = 0
  BACK = 1
  FRONT = 2


class CompositingMode(enum.IntEnum):
  NONE = 0
  OVER = 1
  SPLAT_OVER = 2
