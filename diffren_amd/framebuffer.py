"""Framebuffer: the rasterizer's per-pixel outputs for deferred shading.

Mirrors diffren.jax.internal.framebuffer.Framebuffer
(diffren/jax/internal/framebuffer.py:23-129): barycentrics, triangle_id,
vertex_ids and foreground_mask of shape [..., L, H, W, C] plus shaded
`attributes`.  Here the buffers are torch tensors in device memory.

`vertex_ids` and `foreground_mask` are derived on first access instead of being
written by the rasterizer: the interpolation and splat kernels read the triangle
id and gather vertex ids themselves (rasterize.py:109-113 writes both as extra
full-frame passes).
"""
import dataclasses
from typing import Dict, Optional

import torch


@dataclasses.dataclass
class Framebuffer:
  barycentrics: torch.Tensor           # [L, H, W, 3] f32
  triangle_id: torch.Tensor            # [L, H, W, 1] i32, -1 on background
  triangles: Optional[torch.Tensor] = None  # [T, 3] i32, for vertex_ids
  z_buffer: Optional[torch.Tensor] = None   # [L, H, W] f32 NDC depth
  attributes: Dict[str, torch.Tensor] = dataclasses.field(default_factory=dict)
  _vertex_ids: Optional[torch.Tensor] = None
  _foreground_mask: Optional[torch.Tensor] = None

  def __post_init__(self):
    shapes = [tuple(self.barycentrics.shape), tuple(self.triangle_id.shape)]
    shapes += [tuple(v.shape) for v in self.attributes.values()]
    for s in shapes[1:]:
      if s[:-1] != shapes[0][:-1]:
        raise ValueError("Expected all input shapes to match (up to "
                         f"channels), but found {shapes}")

  @property
  def vertex_ids(self):
    """triangles[triangle_id] ([..., 3] i32); id -1 wraps to the last
    triangle like the reference's gather (barycentrics are 0 there)."""
    if self._vertex_ids is None:
      if self.triangles is None:
        raise ValueError("vertex_ids needs the triangles tensor")
      ids = self.triangle_id[..., 0].long()
      ids = torch.where(ids < 0, ids + self.triangles.shape[0], ids)
      self._vertex_ids = self.triangles[ids]
    return self._vertex_ids

  @property
  def foreground_mask(self):
    """1.0 where a triangle covers the pixel (rasterize.py:109: z != 1)."""
    if self._foreground_mask is None:
      self._foreground_mask = (self.triangle_id >= 0).float()
    return self._foreground_mask

  @property
  def is_batched(self):
    return self.barycentrics.dim() == 5

  @property
  def is_multi_layer(self):
    return self.barycentrics.dim() in (4, 5)

  @property
  def batch_size(self):
    return self.barycentrics.shape[0] if self.is_batched else 1

  @property
  def num_layers(self):
    if self.is_batched:
      return self.barycentrics.shape[1]
    if self.is_multi_layer:
      return self.barycentrics.shape[0]
    return 1

  @property
  def height(self):
    return self.barycentrics.shape[-3]

  @property
  def width(self):
    return self.barycentrics.shape[-2]

  @property
  def pixel_count(self):
    return self.num_layers * self.height * self.width

  @property
  def background_mask(self):
    return 1 - self.foreground_mask
