"""TEST INFRASTRUCTURE ONLY: CPU oracle for interpolate -> splat -> over.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may import
this module; nothing in diffren_amd/ does.

A restatement, in torch CPU tensor ops (float32 to mirror the reference's FP32
arithmetic, float64 for the gradient oracle), of:
  interpolate   diffren/jax/internal/interpolate.py:24-63, 118-145
                (+ the vertex_ids gather of rasterize.py:111-113)
  splat         diffren/jax/internal/splat.py:26-175  (splat_at_pixel_centers)
                diffren/jax/internal/splat.py:178-229 (splat_shaded_pixels)
  over          diffren/jax/composite.py:20-37
  transform     diffren/jax/utils/transforms.py:40-78
It follows the reference's own formulation -- 3 padded (H+2)x(W+2)
accumulators, one "push" of every source layer per 3x3 tap -- which is a
different formulation from the GPU's per-target pull, so agreement is not a
re-typed formula.  Gradients are torch autograd through this restatement
("autodiff of the restatement": jax is not installed, so the reference's own
jax.grad cannot run).

Pins (tests/test_oracle_splat.py): the reference PNG fixtures
Two_Triangles_Splat_{Composite,Layer_0..2}.png and Unlit_Cube_*.png at the
reference's tolerance (compare_images.py:26-70), closed forms (the splat
coefficient, a flat fully-covered surface normalising back to its colour,
`over` of one layer being the identity, single-layer splats leaving the fg/bg
accumulators empty) and central finite differences of the float64 path.
"""
import math

import numpy as np
import torch

SPLAT_VARIANCE = 0.25
EXTRA_ACCUMULATION_EPSILON = 0.05


def splat_coefficient():
  """splat.py:63-78: (1 + 0.05) / sum of the 3x3 Gaussian taps."""
  scale = -1.0 / (2.0 * SPLAT_VARIANCE)
  total = sum(math.exp(scale * (u * u + v * v)) for u in (-1, 0, 1)
              for v in (-1, 0, 1))
  return (1.0 + EXTRA_ACCUMULATION_EPSILON) / total


def _t(x, dtype):
  return torch.as_tensor(np.asarray(x) if not torch.is_tensor(x) else x,
                         dtype=dtype)


def transform(matrix, vertices, dtype=torch.float64):
  """clip = M @ [v, 1] with the GPU kernel's summation order."""
  m = _t(matrix, dtype)
  v = _t(vertices, dtype) if torch.is_tensor(vertices) else _t(vertices, dtype)
  if v.shape[1] == 3:
    v = torch.cat([v, torch.ones_like(v[:, :1])], 1)
  rows = []
  for i in range(4):
    rows.append(((m[i, 0] * v[:, 0] + m[i, 1] * v[:, 1]) + m[i, 2] * v[:, 2])
                + m[i, 3] * v[:, 3])
  return torch.stack(rows, 1)


def interpolate(attribute, ids, barys, triangles, background=None):
  """[V,C] -> [L,H,W,C]: m * sum_k b_k A[vid_k] + (1 - m) * bg."""
  a = attribute
  ids = torch.as_tensor(np.asarray(ids)).long()
  tris = torch.as_tensor(np.asarray(triangles)).long()
  # Tensors (possibly differentiable barycentrics) are used as they are.
  b = (barys if isinstance(barys, torch.Tensor) else
       torch.as_tensor(np.asarray(barys))).to(a.dtype)
  vid = tris[torch.where(ids < 0, ids + tris.shape[0], ids)]  # wraps like jnp
  corners = a[vid]                                            # [...,3,C]
  summed = (corners[..., 0, :] * b[..., 0:1] +
            corners[..., 1, :] * b[..., 1:2]) + corners[..., 2, :] * b[..., 2:3]
  m = (ids >= 0).to(a.dtype)[..., None]
  if background is None:
    bg = torch.zeros(a.shape[1], dtype=a.dtype)
  else:
    bg = torch.as_tensor(np.asarray(background), dtype=a.dtype).reshape(-1)
  return m * summed + (1 - m) * bg


def over(layers):
  out = layers[-1]
  for i in range(2, layers.shape[0] + 1):
    out = layers[-i] + (1.0 - layers[-i, ..., -1:]) * out
  return out


def splat_at_pixel_centers(xyz, color, mask):
  """xyz [L,H,W,3], color [L,H,W,C], mask [L,H,W,1] -> (out, accum, norm)."""
  dtype = color.dtype
  L, H, W, C = color.shape
  coef = splat_coefficient()
  acc_c = torch.zeros((3, H + 2, W + 2, C), dtype=dtype)
  acc_w = torch.zeros((3, H + 2, W + 2, 1), dtype=dtype)
  # z of every layer, padded with 1.0 (background depth).
  zpad = torch.ones((L, H + 2, W + 2, 1), dtype=dtype)
  zpad[:, 1:-1, 1:-1] = xyz[..., 2:3]
  for l in range(L):
    # Offset from the splat position to the centre of its own pixel.
    to_center = (torch.floor(xyz[l, ..., :2]) + 0.5) - xyz[l, ..., :2]
    for u in (-1, 0, 1):
      for v in (-1, 0, 1):
        d = to_center + torch.tensor([u, v], dtype=dtype)
        d2 = d[..., 0:1] ** 2 + d[..., 1:2] ** 2
        w = mask[l] * coef * torch.exp(-2.0 * d2)           # [H,W,1]
        # The tap lands on pixel (x+u, y+v): rows y+v+1, cols x+u+1 padded.
        r0, c0 = v + 1, u + 1
        tap_w = torch.zeros((H + 2, W + 2, 1), dtype=dtype)
        tap_c = torch.zeros((H + 2, W + 2, C), dtype=dtype)
        tap_w[r0:r0 + H, c0:c0 + W] = w
        tap_c[r0:r0 + H, c0:c0 + W] = w * color[l]
        # z of the source layers as seen from each target pixel.
        src_z = torch.ones((L, H + 2, W + 2, 1), dtype=dtype)
        src_z[:, r0:r0 + H, c0:c0 + W] = xyz[..., 2:3]
        d_center = torch.abs(src_z - zpad[0])
        c_idx = torch.argmin(d_center, dim=0)
        c_min = torch.amin(d_center, dim=0)
        d_tap = torch.abs(zpad - src_z[0])
        t_idx = torch.argmin(d_tap, dim=0)
        t_min = torch.amin(d_tap, dim=0)
        surface = torch.where(t_min < c_min, -t_idx, c_idx)
        masks = torch.stack([(surface > l), (surface == l),
                             (surface < l)]).to(dtype)      # [3,H+2,W+2,1]
        acc_c = acc_c + tap_c[None] * masks
        acc_w = acc_w + tap_w[None] * masks
  acc_c = acc_c[:, 1:-1, 1:-1]
  acc_w = acc_w[:, 1:-1, 1:-1]
  norm = acc_c * (1.0 / (torch.clamp(acc_w - 1.0, min=0.0) + 1.0))
  return over(norm), acc_c, norm


def viewport_positions(clip_buffer, width, height):
  """splat.py:218-221: (xyz / w + 1) * [W, H, 1] * 0.5."""
  ndc = clip_buffer[..., :3] / clip_buffer[..., 3:4]
  scale = torch.tensor([width, height, 1], dtype=clip_buffer.dtype)
  return (ndc + 1.0) * scale * 0.5


def splat_shaded_pixels(clip_vertices, ids, barys, triangles, shaded):
  """splat.py:178-229 given clip-space vertices [V,4]."""
  L, H, W, _ = shaded.shape
  clip = interpolate(clip_vertices, ids, barys, triangles, (0, 0, 1, 1))
  xyz = viewport_positions(clip, W, H)
  mask = (torch.as_tensor(np.asarray(ids)) >= 0).to(shaded.dtype)[..., None]
  return splat_at_pixel_centers(xyz, shaded, mask)


def render_splat(vertices, colors, camera, ids, barys, triangles):
  """render.py:107-132 for a pass-through shader, rasterization given.

  vertices [V,3], colors [V,C] (tensors, may require grad), camera [4,4].
  Returns the [H,W,C] image.
  """
  clip = transform(camera, vertices, vertices.dtype)
  shaded = interpolate(colors, ids, barys, triangles, (0.0,))
  out, _, _ = splat_shaded_pixels(clip, ids, barys, triangles, shaded)
  return out
