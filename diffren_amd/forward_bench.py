"""Forward-only benchmark: `python -m diffren_amd.forward_bench`.

Workload: the 1M-triangle UV sphere (500 rows x 1000 columns, 4-octave
seeded Perlin displacement of amplitude 0.1) with float32 vertex normals
plus 8 per-vertex features ~ N(0, 1) and alpha 1 as a pass-through shader
(C = Cs = 12), rendered at 1024x1024 with 3 layers from 8 seeded poses on a
radius-3 orbit.  One step = rasterize + interpolate + splat + over of all 8
views.  The batched fused path (BatchedSplatRenderer.render captured in a
HIP graph) is timed next to the per-view paths (render_triangles per view,
no autograd: unfused passes -- the plain path -- and with the splat on the
fused kernels), and the fused target kernel's launches are timed one by one for
its algorithmic rate: L*P*16 (ids + barycentrics) + T*12 + V*4*C read and
P*4*Cs written per view, against the 8 TB/s HBM peak.  Prints one JSON line.
"""
import argparse
import json

import numpy as np
import torch

from diffren_amd import batch
from diffren_amd import mesh as mesh_lib
from diffren_amd import render
from diffren_amd import scenes

HBM_PEAK_GBS = 8000.0


def timed(fn, reps):
  fn()
  torch.cuda.synchronize()
  start = torch.cuda.Event(enable_timing=True)
  end = torch.cuda.Event(enable_timing=True)
  start.record()
  for _ in range(reps):
    fn()
  end.record()
  torch.cuda.synchronize()
  return start.elapsed_time(end) / reps


def run(args):
  dev = torch.device("cuda:0")
  v, t = scenes.uv_sphere_mesh(args.rows, args.cols, seed=0)
  vt = torch.as_tensor(v, device=dev)
  tt = torch.as_tensor(t, device=dev)
  normals = mesh_lib.compute_vertex_normals(vt, tt)
  rng = np.random.default_rng(2)
  feats = torch.as_tensor(rng.normal(size=(v.shape[0], 8)).astype(np.float32),
                          device=dev)
  attrs = torch.cat([normals, feats, torch.ones_like(feats[:, :1])],
                    1).contiguous()
  C = attrs.shape[1]
  R, L = args.resolution, args.layers
  cams = [torch.as_tensor(c, device=dev)
          for c in scenes.orbit_cameras(args.views, seed=0)]
  jobs = [(0, c, R, R) for c in cams]
  r = batch.BatchedSplatRenderer(
      [dict(vertices=vt, triangles=tt, colors=attrs)], jobs, L, tile="auto",
      streams=args.streams)
  r.capture(forward_only=True)
  fused_ms = timed(r.replay, args.steps)

  # One launch of the fused forward (target) kernel per group, each timed
  # with its own event pair.
  pairs = []
  for _ in range(args.steps):
    for g in r.groups:
      a = torch.cuda.Event(enable_timing=True)
      b = torch.cuda.Event(enable_timing=True)
      a.record()
      r._splat_forward(g, 0)
      b.record()
      pairs.append((a, b))
  torch.cuda.synchronize()
  launch_ms = sum(a.elapsed_time(b) for a, b in pairs) / len(pairs)
  views_per_launch = len(jobs) / len(r.groups)
  P = R * R
  per_view = (L * P * 16 + t.shape[0] * 12 + v.shape[0] * 4 * C +
              P * 4 * C)
  achieved = per_view * views_per_launch / (launch_ms * 1e-3) / 1e9

  def run_per_view(fused):
    with torch.no_grad():
      for c in cams:
        render.render_triangles(vt, {"a": attrs}, tt, c, R, R,
                                lambda a: a["a"], num_layers=L,
                                fused_splat=fused)

  plain_ms = timed(lambda: run_per_view(False), args.plain_steps)
  per_view_fused_ms = timed(lambda: run_per_view(True), args.plain_steps)
  n = len(jobs)
  return {
      "workload": (f"1M-triangle Perlin sphere ({t.shape[0]} triangles), "
                   f"{R}x{R}, {L} layers, {n} views, normals + 8 features "
                   f"+ alpha (C = {C}), forward only"),
      "fused_ms_per_step": round(fused_ms, 4),
      "fused_renders_per_s": round(n / (fused_ms * 1e-3), 1),
      "fused_mtri_per_s": round(n * t.shape[0] / (fused_ms * 1e-3) / 1e6, 1),
      "plain_ms_per_step": round(plain_ms, 4),
      "plain_renders_per_s": round(n / (plain_ms * 1e-3), 1),
      "speedup_vs_plain": round(plain_ms / fused_ms, 2),
      "per_view_fused_ms_per_step": round(per_view_fused_ms, 4),
      "speedup_vs_per_view_fused": round(per_view_fused_ms / fused_ms, 2),
      "target_kernel": {
          "kernel": f"target_kernel<{C}>", "avg_launch_ms": round(launch_ms, 4),
          "views_per_launch": views_per_launch,
          "algorithmic_bytes_per_view": per_view,
          "achieved_gbs": round(achieved, 1),
          "frac_of_hbm_peak": round(achieved / HBM_PEAK_GBS, 4)},
      "streams": len(r.groups),
      "tile_threads": r.raster_config.tile_threads,
  }


def main():
  p = argparse.ArgumentParser()
  p.add_argument("--rows", type=int, default=500)
  p.add_argument("--cols", type=int, default=1000)
  p.add_argument("--views", type=int, default=8)
  p.add_argument("--resolution", type=int, default=1024)
  p.add_argument("--layers", type=int, default=3)
  p.add_argument("--steps", type=int, default=20)
  p.add_argument("--plain-steps", type=int, default=3)
  p.add_argument("--streams", type=lambda v: v if v == "auto" else int(v),
                 default="auto")
  print(json.dumps(run(p.parse_args())), flush=True)


if __name__ == "__main__":
  main()
