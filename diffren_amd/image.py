"""Bilinear image resampling on the GPU (diffren/jax/utils/image.py:21-118).

`bilinear_resample(image, locations, image_order, locations_order, pad_mode)`
has the reference's arguments and semantics: texel (i, j) is centred at
(j + 0.5, i + 0.5), locations are (x, y), the image is padded by one texel
with zeros ('constant') or its edge ('edge').  The gather runs in the
dr_bilinear_resample_fwd kernel and is differentiable with respect to the
image and the locations through a hand-written backward kernel.
"""
import torch

from diffren_amd import _native

_PAD = {"constant": _native.PAD_CONSTANT, "edge": _native.PAD_EDGE}


class _Resample(torch.autograd.Function):

  @staticmethod
  def forward(ctx, image, locations, pad):
    img = image.detach().to(torch.float32).contiguous()
    loc = locations.detach().to(torch.float32).contiguous()
    H, W, C = img.shape
    n = loc.shape[0]
    out = torch.empty((n, C), dtype=torch.float32, device=img.device)
    _native.check(
        _native.lib().dr_bilinear_resample_fwd(
            _native.ptr(img), H, W, C, _native.ptr(loc), n, pad,
            _native.ptr(out), _native.stream_handle()),
        "dr_bilinear_resample_fwd")
    ctx.save_for_backward(img, loc)
    ctx.pad = pad
    return out

  @staticmethod
  def backward(ctx, d_out):
    img, loc = ctx.saved_tensors
    H, W, C = img.shape
    d_out = d_out.to(torch.float32).contiguous()
    d_img = torch.zeros_like(img) if ctx.needs_input_grad[0] else None
    d_loc = torch.empty_like(loc) if ctx.needs_input_grad[1] else None
    _native.check(
        _native.lib().dr_bilinear_resample_bwd(
            _native.ptr(img), H, W, C, _native.ptr(loc), loc.shape[0],
            ctx.pad, _native.ptr(d_out), _native.ptr(d_img),
            _native.ptr(d_loc), _native.stream_handle()),
        "dr_bilinear_resample_bwd")
    return d_img, d_loc, None


def bilinear_resample(image, locations, image_order="HWC",
                      locations_order="HWC", pad_mode="constant"):
  """[H1,W1,C] (or [C,H1,W1]) image sampled at [H2,W2,2] (or [2,H2,W2])
  (x, y) locations -> [H2,W2,C] (or [C,H2,W2]) (image.py:21-118).
  Leading dimensions of an HWC `locations` beyond [H2, W2] are kept."""
  if image_order not in ("HWC", "CHW") or locations_order not in ("HWC",
                                                                  "CHW"):
    raise ValueError(f"Bad ordering constant: Got image order {image_order}, "
                     f"locations order {locations_order}")
  loc_axis = -1 if locations_order == "HWC" else 0
  if (image.dim() != 3 or locations.dim() < 3 or
      locations.shape[loc_axis] != 2):
    raise ValueError("Bad shapes for bilinear resample. Got image shape "
                     f"{tuple(image.shape)}, locations shape "
                     f"{tuple(locations.shape)}")
  if pad_mode not in _PAD:
    raise ValueError(f"Unsupported padding mode: {pad_mode}")
  _native.require_device(image, locations)
  img = image if image_order == "HWC" else image.permute(1, 2, 0)
  loc = locations if locations_order == "HWC" else locations.permute(1, 2, 0)
  out = _Resample.apply(img, loc.reshape(-1, 2), _PAD[pad_mode])
  out = out.reshape(tuple(loc.shape[:-1]) + (img.shape[-1],))
  return out if image_order == "HWC" else out.permute(2, 0, 1)
