"""transform_homogeneous on the GPU (diffren/jax/utils/transforms.py:40-78).

clip = M @ [x, y, z, 1] per vertex, computed by the dr_transform_homogeneous
kernel in one fixed summation order.  Differentiable with respect to the
vertices (kernel) and the matrix (a 4x4 torch product, tiny).
"""
import torch

from diffren_amd import _native


class _Transform(torch.autograd.Function):

  @staticmethod
  def forward(ctx, matrix, vertices):
    m = matrix.detach().to(torch.float32).contiguous()
    v = vertices.detach().to(torch.float32).contiguous()
    out = torch.empty((v.shape[0], 4), dtype=torch.float32, device=v.device)
    _native.check(
        _native.lib().dr_transform_homogeneous(
            _native.ptr(m), _native.ptr(v), v.shape[1], v.shape[0],
            _native.ptr(out), _native.stream_handle()),
        "dr_transform_homogeneous")
    ctx.save_for_backward(m, v)
    return out

  @staticmethod
  def backward(ctx, d_clip):
    m, v = ctx.saved_tensors
    dm = dv = None
    d_clip = d_clip.contiguous()
    if ctx.needs_input_grad[1]:
      dv = torch.empty_like(v)
      _native.check(
          _native.lib().dr_transform_homogeneous_bwd(
              _native.ptr(m), _native.ptr(d_clip), v.shape[1], v.shape[0],
              _native.ptr(dv), _native.stream_handle()),
          "dr_transform_homogeneous_bwd")
    if ctx.needs_input_grad[0]:
      vh = v if v.shape[1] == 4 else torch.nn.functional.pad(v, (0, 1),
                                                             value=1.0)
      dm = d_clip.t() @ vh
    return dm, dv


def transform_homogeneous(matrices, vertices):
  """[4,4] matrix, [N,3|4] vertices -> [N,4] (transforms.py:40-78)."""
  if matrices.dim() != vertices.dim():
    raise ValueError(
        "length of matrices and vertices shapes must match but are "
        f"{tuple(matrices.shape)} and {tuple(vertices.shape)}")
  if vertices.shape[-1] not in (3, 4):
    raise ValueError("vertices must be either 3-D (xyz) or 4-D (xyzw) but is "
                     f"{vertices.shape[-1]}")
  if matrices.dim() != 2:
    return torch.stack([transform_homogeneous(m, v)
                        for m, v in zip(matrices, vertices)])
  _native.require_device(matrices, vertices)
  return _Transform.apply(matrices, vertices)
