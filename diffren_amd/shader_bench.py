"""Learned deferred shader benchmark: `python -m diffren_amd.shader_bench`.

Workload: G-buffers (position, normal, view direction, albedo) of 16 seeded
icosphere scenes (642 vertices, 1280 triangles, radial noise 5%) rendered at
1024x1024 (16.8M pixels); ground truth is the many-light Phong shader with 64
seeded point lights uniform in a shell of radius 4-5 around the object and
intensities U[0, 1].  The MLP 12 -> 64 x 4 -> 3 (bfloat16 MFMA, float32
accumulation) is trained on foreground pixels by the library's own training
and Adam kernels.  Prints one JSON line: the loss curve, per-pixel error
against Phong, the fraction of pixels that fall back to Phong (error >
tolerance), inference, training-step and Phong times, and the MLP kernels'
MFMA utilisation against the dense bf16 peak (2.5 PFLOP/s).
"""
import argparse
import json

import numpy as np
import torch

from diffren_amd import rasterize
from diffren_amd import scenes
from diffren_amd import shading

BF16_PEAK_TFLOPS = 2500.0
FWD_FLOP_PER_PX = 2 * (12 * 64 + 3 * 64 * 64 + 64 * 3)   # 26,496
TRAIN_FLOP_PER_PX = 77952   # forward + dW for all layers + dX for 2..5


def gbuffers(num_scenes, res, dev):
  out, masks = [], []
  for seed in range(num_scenes):
    s = scenes.icosphere_scene(seed=seed)
    v = torch.as_tensor(s["vertices"], device=dev)
    tris = torch.as_tensor(s["triangles"], device=dev)
    cam = torch.as_tensor(s["camera"], device=dev)
    fb = rasterize.rasterize_triangles(v, tris, cam, res, res, 1)
    nrm = scenes.vertex_normals(s["vertices"].astype(np.float64),
                                s["triangles"]).astype(np.float32)
    g = shading.gbuffer(fb, v, torch.as_tensor(nrm, device=dev),
                        torch.as_tensor(s["colors"][:, :3].copy(),
                                        device=dev), (0.0, 0.0, 0.0))
    out.append(g[0])
    masks.append(fb.triangle_id[0, ..., 0] >= 0)
  return torch.stack(out), torch.stack(masks)


def timed(fn, reps=10):
  start = torch.cuda.Event(enable_timing=True)
  end = torch.cuda.Event(enable_timing=True)
  fn()
  start.record()
  for _ in range(reps):
    fn()
  end.record()
  torch.cuda.synchronize()
  return start.elapsed_time(end) / reps


def run(args):
  dev = torch.device("cuda:0")
  g, mask = gbuffers(args.scenes, args.resolution, dev)
  lights = torch.as_tensor(
      shading.random_lights(args.lights, seed=args.seed, r_min=4.0, r_max=5.0,
                            center=(0.0, 0.0, -3.0)), device=dev)
  ref = shading.phong(g, lights)
  x_all = g.reshape(-1, 12)
  t_all = ref.reshape(-1, 3)
  m_all = mask.reshape(-1)
  fg = torch.nonzero(m_all).squeeze(1)
  n_px = x_all.shape[0]

  norm = shading.MLPShader.normalization(x_all[fg])
  net = shading.MLPShader(shading.init_params(args.seed), dev,
                          input_norm=norm,
                          albedo_modulate=not args.no_modulate)
  gen = torch.Generator(device=dev).manual_seed(args.seed)
  curve = []
  for step in range(args.steps):
    idx = fg[torch.randint(0, fg.numel(), (args.batch,), device=dev,
                           generator=gen)]
    lr = args.lr * (0.1 ** (step / max(args.steps - 1, 1)))
    loss = net.train_step(x_all[idx].contiguous(), t_all[idx].contiguous(),
                          None, 1.0 / (3 * args.batch), lr=lr)
    if step % max(args.steps // 10, 1) == 0 or step == args.steps - 1:
      curve.append((step, float(loss.item())))

  y = net.forward(x_all)
  err = (y - t_all).abs().amax(-1)[m_all]
  img, fb_mask = shading.shade_with_fallback(net, x_all, t_all,
                                             args.tolerance)
  final_err = (img - t_all).abs().amax(-1)[m_all]

  phong_ms = timed(lambda: shading.phong(g, lights))
  mlp_ms = timed(lambda: net.forward(x_all))
  # The training pass is timed on a fixed 1M-pixel batch.
  train_px = 1 << 20
  xb = x_all[fg[:train_px]].contiguous()
  tb = t_all[fg[:train_px]].contiguous()
  train_ms = timed(lambda: net.loss_and_grad(xb, tb, None, 1.0))
  fwd_tflops = FWD_FLOP_PER_PX * n_px / (mlp_ms * 1e-3) / 1e12
  train_tflops = TRAIN_FLOP_PER_PX * train_px / (train_ms * 1e-3) / 1e12
  return {
      "workload": (f"{args.scenes} icosphere scenes at {args.resolution}^2 "
                   f"({n_px} px, {fg.numel()} foreground), {args.lights} "
                   "point lights"),
      "train_steps": args.steps, "batch_px": args.batch,
      "albedo_modulate": not args.no_modulate,
      "loss_curve": curve,
      "mlp_error_fg": {"mean": float(err.mean()), "p50": float(
          err.median()), "p99": float(torch.quantile(err[:1 << 24].float(),
                                                     0.99)),
                       "max": float(err.max())},
      "tolerance": args.tolerance,
      "fallback_fraction_fg": float(fb_mask[m_all].float().mean()),
      "composited_error_max": float(final_err.max()),
      "phong_ms": round(phong_ms, 4),
      "mlp_forward_ms": round(mlp_ms, 4),
      "mlp_forward_tflops": round(fwd_tflops, 1),
      "mlp_forward_mfma_util": round(fwd_tflops / BF16_PEAK_TFLOPS, 4),
      "mlp_train_pass_px": train_px,
      "mlp_train_pass_ms": round(train_ms, 4),
      "mlp_train_tflops": round(train_tflops, 1),
      "mlp_train_mfma_util": round(train_tflops / BF16_PEAK_TFLOPS, 4),
      "speed_ratio_phong_over_mlp": round(phong_ms / mlp_ms, 3),
  }


def main():
  p = argparse.ArgumentParser()
  p.add_argument("--scenes", type=int, default=16)
  p.add_argument("--resolution", type=int, default=1024)
  p.add_argument("--lights", type=int, default=64)
  p.add_argument("--steps", type=int, default=8000)
  p.add_argument("--batch", type=int, default=1 << 18)
  p.add_argument("--lr", type=float, default=3e-3)
  p.add_argument("--tolerance", type=float, default=1e-2)
  p.add_argument("--seed", type=int, default=0)
  p.add_argument("--no-modulate", action="store_true",
                 help="network predicts RGB directly instead of the light "
                 "response multiplied by the G-buffer albedo")
  print(json.dumps(run(p.parse_args())), flush=True)


if __name__ == "__main__":
  main()
