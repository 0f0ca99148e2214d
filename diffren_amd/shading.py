"""Deferred shading on the GPU: G-buffers, many-light Phong and a learned
MLP shader with its training loop.

gbuffer(...)          -> [L,H,W,12] position, normal, view direction, albedo
phong(...)            -> [...,3] many-light form of shaders.point_light
                         (diffren/jax/utils/shaders.py:100-173), ambient once
texture_map(...)      -> [...,H,W,C] shaders.texture_map
                         (diffren/jax/utils/shaders.py:47-72) on the
                         bilinear resampling kernel, edge clamped
load_texture(...)     -> shaders.load_texture (:25-34), lower-left origin
diffuse_light(...)    -> shaders.diffuse_light (:75-97)
point_light(...)      -> shaders.point_light (:100-173), one light, as the
                         reference writes it (elementwise torch ops)
MLPShader             12 -> 64x4 -> 3 MLP, bfloat16 MFMA, float32 accumulation,
                      trained by the library's own forward/backward/Adam
                      kernels; `shade_with_fallback` keeps the Phong result
                      wherever the network's per-pixel error exceeds a bound;
                      save / load keep weights and Adam state in a
                      safetensors file.
"""
import ctypes

import numpy as np
import safetensors
import safetensors.torch as safetensors_torch
import torch

from diffren_amd import _native
from diffren_amd import image

PHONG_DEFAULTS = dict(ka=0.3, kd=1.4, ks=0.5, ns=16.0,
                      attenuation_constant=1.0, attenuation_linear=0.7,
                      attenuation_quadratic=1.8)


def gbuffer(framebuffer, positions, normals, albedo, eye):
  """[L,H,W,12] G-buffer from a rasterized Framebuffer and [V,3] attributes."""
  ids = framebuffer.triangle_id[..., 0].contiguous()
  b = framebuffer.barycentrics.contiguous()
  tris = framebuffer.triangles
  pos = positions.to(torch.float32).contiguous()
  nrm = normals.to(torch.float32).contiguous()
  alb = albedo.to(torch.float32).contiguous()
  eye_t = torch.as_tensor(eye, dtype=torch.float32,
                          device=pos.device).contiguous()
  out = torch.empty(tuple(ids.shape) + (12,), dtype=torch.float32,
                    device=pos.device)
  _native.check(
      _native.lib().dr_gbuffer(
          _native.ptr(ids), _native.ptr(b), _native.ptr(tris), tris.shape[0],
          _native.ptr(pos), _native.ptr(nrm), _native.ptr(alb), pos.shape[0],
          _native.ptr(eye_t), ids.numel(), _native.ptr(out),
          _native.stream_handle()), "dr_gbuffer")
  return out


def load_texture(texture_filename, lower_left_origin=True):
  """shaders.py:25-34: an image file as float32 [Ht, Wt, C] in [0, 1], rows
  flipped so that v = 0 is the bottom row (OpenGL convention)."""
  from PIL import Image
  with Image.open(texture_filename) as im:
    tex = np.asarray(im).astype(np.float32) / 255.0
  if lower_left_origin:
    tex = np.ascontiguousarray(np.flipud(tex))
  return tex


def texture_map(uv_image, texture_image_or_filename):
  """Bilinear texture lookup with CLAMP_TO_EDGE (shaders.py:47-72).

  uv_image: [..., H, W, 2] UV coordinates in [0, 1]; the texture is a
  [Ht, Wt, C] float tensor or array, or a file name for load_texture.
  Returns [..., H, W, C], differentiable with respect to both inputs.
  """
  texture_image = texture_image_or_filename
  if isinstance(texture_image, str):
    texture_image = load_texture(texture_image)
  tex = torch.as_tensor(texture_image, dtype=torch.float32,
                        device=uv_image.device)
  scale = torch.tensor([tex.shape[1], tex.shape[0]], dtype=torch.float32,
                       device=uv_image.device)
  query = uv_image * scale
  return image.bilinear_resample(tex, query, pad_mode="edge")


def l2_normalize(x, dim=-1, eps=1e-12):
  """transforms.py:26-37: x * rsqrt(sum(x^2) + eps) along dim."""
  return x * torch.rsqrt((x * x).sum(dim=dim, keepdim=True) + eps)


def diffuse_light(diffuse_image, normals_image,
                  light_direction=(0.5, 0.5, 1.0), ka=0.5, kd=0.5):
  """shaders.py:75-97: ambient plus clamped n.l directional diffuse."""
  d = l2_normalize(torch.as_tensor(light_direction, dtype=torch.float32,
                                   device=normals_image.device))
  n_dot_l = torch.clamp((normals_image * d).sum(-1, keepdim=True), 0.0, 1.0)
  return diffuse_image * ka + diffuse_image * kd * n_dot_l


def point_light(diffuse_image, normals_image, pixel_positions,
                point_locations, eye_position=(0.0, 0.0, 0.0), ka=0.3,
                kd=1.4, ks=0.0, ns=0.0, attenuation_constant=1.0,
                attenuation_linear=0.7, attenuation_quadratic=1.8):
  """shaders.py:100-173: one point light with ambient, diffuse and
  Blinn-Phong specular terms and quadratic attenuation (the many-light
  kernel is `phong`)."""
  dev = pixel_positions.device
  light = torch.as_tensor(point_locations, dtype=torch.float32, device=dev)
  eye = torch.as_tensor(eye_position, dtype=torch.float32, device=dev)
  to_light = light - pixel_positions
  light_dir = l2_normalize(to_light)
  view_dir = l2_normalize(eye - pixel_positions)
  half = l2_normalize(light_dir + view_dir)
  nol = torch.clamp((normals_image * light_dir).sum(-1, keepdim=True), 0.0,
                    1.0)
  noh = torch.clamp((normals_image * half).sum(-1, keepdim=True), 0.0, 1.0)
  dist = torch.linalg.vector_norm(to_light, dim=-1, keepdim=True)
  decay = 1.0 / (attenuation_constant + attenuation_linear * dist +
                 attenuation_quadratic * dist * dist)
  luma = ka + kd * nol + ks * noh ** ns
  return diffuse_image * luma * decay


def phong(gbuf, lights, **params):
  """gbuf [...,12], lights [K,4] (xyz, intensity) -> [...,3]."""
  prm = dict(PHONG_DEFAULTS)
  prm.update(params)
  g = gbuf.contiguous()
  lt = lights.to(torch.float32).contiguous()
  out = torch.empty(tuple(g.shape[:-1]) + (3,), dtype=torch.float32,
                    device=g.device)
  cp = _native.PhongParams(**{k: float(v) for k, v in prm.items()})
  _native.check(
      _native.lib().dr_phong_shade(_native.ptr(g), g.numel() // 12,
                                   _native.ptr(lt), lt.shape[0],
                                   ctypes.byref(cp), _native.ptr(out),
                                   _native.stream_handle()), "dr_phong_shade")
  return out


def random_lights(num_lights, seed=0, r_min=4.0, r_max=5.0, center=(0, 0, 0)):
  """Seeded lights uniform in a spherical shell, intensities U[0, 1]."""
  rng = np.random.default_rng(seed)
  d = rng.normal(size=(num_lights, 3))
  d /= np.linalg.norm(d, axis=1, keepdims=True)
  r = np.cbrt(rng.uniform(r_min ** 3, r_max ** 3, size=(num_lights, 1)))
  pos = d * r + np.asarray(center, np.float64)
  inten = rng.uniform(0, 1, size=(num_lights, 1))
  return np.concatenate([pos, inten], 1).astype(np.float32)


MLP_SIZES = (12, 64, 64, 64, 64, 3)


def init_params(seed=0):
  """Fan-in uniform weights and biases, U(-1/sqrt(k), 1/sqrt(k)) for a layer
  with k inputs, flat in the kernels' order:
  W1[12,64] b1 W2[64,64] b2 W3 b3 W4 b4 W5[64,3] b5.  (He-normal weights with
  zero biases train this shader to about three times the per-pixel error.)"""
  rng = np.random.default_rng(seed)
  parts = []
  for k, n in zip(MLP_SIZES[:-1], MLP_SIZES[1:]):
    bound = 1.0 / np.sqrt(k)
    parts.append(rng.uniform(-bound, bound, size=k * n))
    parts.append(rng.uniform(-bound, bound, size=n))
  return np.concatenate(parts).astype(np.float32)


class MLPShader:
  """The learned deferred shader and its on-GPU training state.

  input_norm: optional [24] (shift 12 | scale 12): the network sees
  (x - shift) * scale.  albedo_modulate: the image is albedo * network output
  (the network learns the light response; G-buffer albedo is x[..., 9:12]).
  """

  def __init__(self, params, device, input_norm=None, albedo_modulate=False,
               train_blocks=256):
    lib = _native.lib()
    n = lib.dr_mlp_num_params()
    p = torch.as_tensor(np.asarray(params, np.float32), device=device)
    if p.numel() != n:
      raise ValueError(f"expected {n} parameters, got {p.numel()}")
    self.params = p.contiguous()
    self.packed = torch.empty(lib.dr_mlp_packed_bytes(), dtype=torch.uint8,
                              device=device)
    self.grad = torch.zeros_like(self.params)
    self.m = torch.zeros_like(self.params)
    self.v = torch.zeros_like(self.params)
    self.loss = torch.zeros(1, dtype=torch.float32, device=device)
    self.norm = None if input_norm is None else torch.as_tensor(
        np.asarray(input_norm, np.float32), device=device).contiguous()
    self.flags = _native.MLP_ALBEDO_MODULATE if albedo_modulate else 0
    self.train_blocks = int(train_blocks)
    self.train_ws = _native.workspace(
        lib.dr_mlp_train_workspace_size(self.train_blocks), device)
    self.step_count = 0
    self.pack()

  @staticmethod
  def normalization(x, eps=1e-6):
    """[24] shift/scale from features x [N, 12] (mean and 1/std)."""
    x = x.reshape(-1, 12).to(torch.float64)
    mean = x.mean(0)
    std = x.std(0)
    return torch.cat([mean, 1.0 / (std + eps)]).to(torch.float32).cpu(
        ).numpy()

  def pack(self):
    _native.check(_native.lib().dr_mlp_pack(_native.ptr(self.params),
                                            _native.ptr(self.packed),
                                            _native.stream_handle()),
                  "dr_mlp_pack")

  def forward(self, x, num_blocks=0):
    """x [...,12] -> [...,3]."""
    x = x.contiguous()
    y = torch.empty(tuple(x.shape[:-1]) + (3,), dtype=torch.float32,
                    device=x.device)
    _native.check(
        _native.lib().dr_mlp_forward(_native.ptr(self.packed),
                                     _native.ptr(self.params), _native.ptr(x),
                                     x.numel() // 12, _native.ptr(y),
                                     _native.ptr(self.norm), self.flags,
                                     num_blocks, _native.stream_handle()),
        "dr_mlp_forward")
    return y

  def loss_and_grad(self, x, target, mask=None, scale=None):
    """Clears and fills self.loss / self.grad for one pass over x."""
    x = x.contiguous()
    n = x.numel() // 12
    t = target.contiguous()
    mk = None if mask is None else mask.to(torch.float32).contiguous()
    if scale is None:
      cnt = n if mask is None else max(float(mk.sum().item()), 1.0)
      scale = 1.0 / (3.0 * cnt)
    _native.check(
        _native.lib().dr_mlp_train_step(
            _native.ptr(self.packed), _native.ptr(self.params),
            _native.ptr(x), _native.ptr(t), _native.ptr(mk), n, float(scale),
            _native.ptr(self.norm), self.flags | _native.MLP_GRAD_OVERWRITE,
            _native.ptr(self.grad),
            _native.ptr(self.loss), self.train_blocks,
            _native.ptr(self.train_ws), self.train_ws.numel(),
            _native.stream_handle()), "dr_mlp_train_step")
    return self.loss

  def adam_step(self, lr=1e-3, beta1=0.9, beta2=0.999, eps=1e-8):
    self.step_count += 1
    _native.check(
        _native.lib().dr_mlp_adam(_native.ptr(self.params),
                                  _native.ptr(self.grad), _native.ptr(self.m),
                                  _native.ptr(self.v), self.step_count, lr,
                                  beta1, beta2, eps, _native.stream_handle()),
        "dr_mlp_adam")
    self.pack()

  def train_step(self, x, target, mask=None, scale=None, lr=1e-3):
    loss = self.loss_and_grad(x, target, mask, scale)
    self.adam_step(lr)
    return loss

  def save(self, path):
    """Writes the network and its training state to a safetensors file:
    float32 master weights, Adam moments, the input normalisation and, as
    metadata, the Adam step count and the albedo-modulation flag.  The
    bfloat16 operand images are not stored (load re-packs them)."""
    tensors = {"params": self.params, "adam_m": self.m, "adam_v": self.v}
    if self.norm is not None:
      tensors["input_norm"] = self.norm
    tensors = {k: v.detach().cpu().contiguous() for k, v in tensors.items()}
    meta = {
        "format": "diffren_amd.MLPShader",
        "layers": ",".join(str(s) for s in MLP_SIZES),
        "step_count": str(self.step_count),
        "albedo_modulate": str(int(bool(self.flags &
                                        _native.MLP_ALBEDO_MODULATE))),
    }
    safetensors_torch.save_file(tensors, str(path), metadata=meta)

  @classmethod
  def load(cls, path, device, train_blocks=256):
    """The shader saved by `save`, on `device`, ready to run or to train on
    from the saved Adam state."""
    with safetensors.safe_open(str(path), framework="pt") as f:
      meta = f.metadata() or {}
      if meta.get("format") != "diffren_amd.MLPShader":
        raise ValueError(f"{path}: not a saved MLPShader")
      layers = ",".join(str(s) for s in MLP_SIZES)
      if meta.get("layers") != layers:
        raise ValueError(f"{path}: layer sizes {meta.get('layers')} differ "
                         f"from this library's {layers}")
      t = {k: f.get_tensor(k) for k in f.keys()}
    norm = t.get("input_norm")
    mlp = cls(t["params"].numpy(), device,
              input_norm=None if norm is None else norm.numpy(),
              albedo_modulate=meta.get("albedo_modulate") == "1",
              train_blocks=train_blocks)
    mlp.m.copy_(t["adam_m"].to(device))
    mlp.v.copy_(t["adam_v"].to(device))
    mlp.step_count = int(meta.get("step_count", "0"))
    return mlp


def shade_with_fallback(mlp, gbuf, reference, tolerance=1e-2):
  """Network shading with the plain shader kept where the per-pixel error
  (max over channels) exceeds `tolerance`; returns (image, fallback mask)."""
  y = mlp.forward(gbuf)
  err = (y - reference).abs().amax(-1)
  fb = err > tolerance
  return torch.where(fb[..., None], reference, y), fb
