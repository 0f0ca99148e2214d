"""Batching benchmark: `python -m diffren_amd.batch_bench`.

Workload: 64 independent noise-displaced spheres whose triangle counts are
drawn log-uniformly from 1k to 500k (2.87M triangles in total for seed 0;
the expectation is about 5M), each with
its own seeded orbit camera and square resolution drawn from {128, 256, 512,
1024}, 3 depth layers, RGBA vertex colours.  One step = forward render of
all 64 jobs + L2 loss against targets rendered from perturbed copies of the
meshes + backward to every mesh's vertex positions and colours, run by
BatchedSplatRenderer (one launch per stage for all jobs, tile shape and bin
capacity from the autotuner, captured in a HIP graph).  Reported next to it:
the same work as one render_triangles call plus autograd backward per job,
with the splat as separate passes (the plain path) and on the fused kernels
(per-view fused).  Prints one JSON line.
"""
import argparse
import json

import numpy as np
import torch

from diffren_amd import batch
from diffren_amd import render
from diffren_amd import scenes


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
  work = scenes.batching_workload(args.meshes, seed=args.seed)
  rng = np.random.default_rng(args.seed + 1)
  meshes, targets_mesh, jobs = [], [], []
  for i, s in enumerate(work):
    m = dict(vertices=torch.as_tensor(s["vertices"], device=dev),
             triangles=torch.as_tensor(s["triangles"], device=dev),
             colors=torch.as_tensor(s["colors"], device=dev))
    pert = (s["vertices"] + rng.normal(0, 0.01, s["vertices"].shape)).astype(
        np.float32)
    meshes.append(m)
    targets_mesh.append(dict(m, vertices=torch.as_tensor(pert, device=dev)))
    cam = torch.as_tensor(s["camera"], device=dev)
    jobs.append((i, cam, s["resolution"], s["resolution"]))
  L = args.layers
  tris = sum(int(s["triangles"].shape[0]) for s in work)
  pixels = sum(s["resolution"] ** 2 for s in work)

  tr = batch.BatchedSplatRenderer(targets_mesh, jobs, L)
  tr.render()
  torch.cuda.synchronize()
  targets = [img.clone() for img in tr.images]
  del tr

  r = batch.BatchedSplatRenderer(meshes, jobs, L, tile="auto",
                                 streams=args.streams)
  for i, t in enumerate(targets):
    r.targets[i].copy_(t)
  r.capture()
  fused_ms = timed(r.replay, args.steps)

  def run_per_view(fused):
    for i, s in enumerate(work):
      m = meshes[i]
      vv = m["vertices"].clone().requires_grad_(True)
      cc = m["colors"].clone().requires_grad_(True)
      res = s["resolution"]
      img = render.render_triangles(vv, {"c": cc}, m["triangles"], jobs[i][1],
                                    res, res, lambda a: a["c"], num_layers=L,
                                    fused_splat=fused)
      loss = ((img - targets[i]) ** 2).sum() / (res * res * 4)
      loss.backward()

  plain_ms = timed(lambda: run_per_view(False), args.plain_steps)
  per_view_fused_ms = timed(lambda: run_per_view(True), args.plain_steps)
  n = len(jobs)
  return {
      "workload": (f"{n} meshes, {tris} triangles in total (log-uniform "
                   f"1k-500k each), resolutions from {{128,256,512,1024}} "
                   f"({pixels} px in total), {L} layers, RGBA, fwd + L2 loss "
                   "+ bwd"),
      "fused_ms_per_step": round(fused_ms, 4),
      "fused_renders_per_s": round(n / (fused_ms * 1e-3), 1),
      "fused_mtri_per_s": round(tris / (fused_ms * 1e-3) / 1e6, 1),
      "fused_mpix_per_s": round(pixels / (fused_ms * 1e-3) / 1e6, 1),
      "plain_ms_per_step": round(plain_ms, 4),
      "plain_renders_per_s": round(n / (plain_ms * 1e-3), 1),
      "speedup_vs_plain": round(plain_ms / fused_ms, 2),
      "per_view_fused_ms_per_step": round(per_view_fused_ms, 4),
      "speedup_vs_per_view_fused": round(per_view_fused_ms / fused_ms, 2),
      "raster_tile": [r.raster_config.tile_w, r.raster_config.tile_h],
      "bin_capacity": int(r.raster_config.bin_capacity),
      "streams": len(r.groups),
      "tile_threads": r.raster_config.tile_threads,
      "streams_timings": r.streams_timings,
  }


def main():
  p = argparse.ArgumentParser()
  p.add_argument("--meshes", type=int, default=64)
  p.add_argument("--layers", type=int, default=3)
  p.add_argument("--steps", type=int, default=20)
  p.add_argument("--plain-steps", type=int, default=2)
  p.add_argument("--streams", type=lambda v: v if v == "auto" else int(v),
                 default="auto")
  p.add_argument("--seed", type=int, default=0)
  print(json.dumps(run(p.parse_args())), flush=True)


if __name__ == "__main__":
  main()
