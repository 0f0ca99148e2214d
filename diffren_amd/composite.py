"""Premultiplied back-to-front `over` (diffren/jax/composite.py:20-37)."""


def over(color_layers):
  """[n, H, W, C] front-to-back layers -> [H, W, C]; last channel is alpha."""
  out = color_layers[-1]
  for i in range(2, color_layers.shape[0] + 1):
    alpha = color_layers[-i, ..., -1:]
    out = color_layers[-i] + (1.0 - alpha) * out
  return out
