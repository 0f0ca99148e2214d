"""Mesh utilities used around the render path.

compute_vertex_normals: area-weighted vertex normals
(diffren/jax/utils/mesh.py:22-52): each face's unnormalised normal
(v1 - v0) x (v2 - v0) is added to its three vertices, then normalised as
transforms.l2_normalize does.
"""
import torch

from diffren_amd import shading


def compute_vertex_normals(vertices, triangles):
  """vertices [V, 3] float, triangles [T, 3] int -> [V, 3] unit normals."""
  v = torch.as_tensor(vertices)
  t = torch.as_tensor(triangles, device=v.device).long()
  corners = v[t]
  n = torch.linalg.cross(corners[:, 1] - corners[:, 0],
                         corners[:, 2] - corners[:, 0])
  acc = [torch.zeros_like(v).index_add_(0, t[:, i], n) for i in range(3)]
  return shading.l2_normalize(acc[0] + acc[1] + acc[2], dim=1)
