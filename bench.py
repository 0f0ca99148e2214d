"""Benchmark: batched fused forward+backward rasterize-then-splat.

Workload (the differentiation configuration): a 1,000,000-triangle UV sphere
(500 x 1000 + 2 poles, 500,002 vertices, 4-octave Perlin displacement of
amplitude 0.1), RGBA vertex colours, rendered at 1024x1024 with 3 depth layers
and back-face culling from 8 seeded camera poses on a radius-3 orbit (60 degree
FOV).  One step = forward render of all 8 views + photometric L2 loss against
targets rendered from a copy with vertices perturbed by N(0, 0.01) + backward
to vertex positions and colours.  Inputs are resident in HBM before timing.

Reported next to it, measured in the same run:
  * the plain path: one render_triangles call per view (one rasterizer launch
    per image, then transform, interpolation, viewport and splat as separate
    passes, autograd backward), the pattern of the reference's
    vmap_method='sequential_unrolled' and of its separate XLA passes;
  * the per-view fused path: the same calls with the splat on the fused
    render kernels (render_triangles' default), one image per launch;
  * per-stage kernel times from HIP events and the roofline of the dominant
    kernel;
  * the CPU baseline: the reference CPU rasterizer (compiled from the
    reference sources into oracle/_ref, one view per host thread) plus the
    float32 C restatement of interpolate/splat forward+backward on all host
    threads, for the workload's views;
  * the vertex-gradient relative error of one view against the float64 CPU
    oracle.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def _groups_arg(v):
  return v if v == "auto" else int(v)


def parse():
  p = argparse.ArgumentParser()
  p.add_argument("--gpus", type=int, default=1)
  p.add_argument("--steps", type=int, default=20)
  p.add_argument("--warmup", type=int, default=3)
  p.add_argument("--views", type=int, default=8)
  p.add_argument("--resolution", type=int, default=1024)
  p.add_argument("--layers", type=int, default=3)
  p.add_argument("--rows", type=int, default=500)
  p.add_argument("--cols", type=int, default=1000)
  p.add_argument("--plain-steps", type=int, default=3)
  p.add_argument("--no-cpu-baseline", action="store_true")
  p.add_argument("--no-grad-check", action="store_true")
  p.add_argument("--no-graph", action="store_true")
  p.add_argument("--streams", type=_groups_arg, default="auto",
                 help="concurrent job groups in the fused step, or 'auto' "
                 "(the autotuner times 1, 2 and 4 groups)")
  p.add_argument("--tile-threads",
                 type=lambda v: v if v == "auto" else int(v), default=None,
                 help="threads per raster tile block (128 or 256), or "
                 "'auto' (timed on the step; the default with --streams "
                 "auto, else 256)")
  p.add_argument("--deterministic", action="store_true",
                 help="fixed-point (order-independent) gradient sums")
  p.add_argument("--fused-transform", action="store_true",
                 help="transform vertices inside the raster and splat loads "
                 "instead of a transform launch into clip-space buffers")
  p.add_argument("--pmc-probe", action="store_true",
                 help="run only the fused step + a calibration copy, for "
                 "rocprofv3 --pmc passes")
  return p.parse_args()


def log(*a):
  print(*a, file=sys.stderr, flush=True)


def build_workload(args, dev):
  import numpy as np
  import torch
  from diffren_amd import scenes
  v, t = scenes.uv_sphere_mesh(args.rows, args.cols, seed=0)
  rng = np.random.default_rng(1)
  colors = np.concatenate([rng.uniform(0, 1, (v.shape[0], 3)),
                           np.ones((v.shape[0], 1))], 1).astype(np.float32)
  cams = scenes.orbit_cameras(args.views, seed=0)
  v_target = (v + rng.normal(0, 0.01, v.shape)).astype(np.float32)
  mesh = dict(vertices=torch.as_tensor(v, device=dev),
              triangles=torch.as_tensor(t, device=dev),
              colors=torch.as_tensor(colors, device=dev))
  target_mesh = dict(vertices=torch.as_tensor(v_target, device=dev),
                     triangles=mesh["triangles"], colors=mesh["colors"])
  cams_t = [torch.as_tensor(c, device=dev) for c in cams]
  return dict(v=v, t=t, colors=colors, cams=cams, mesh=mesh,
              target_mesh=target_mesh, cams_t=cams_t)


def time_launches(r, stage, reps):
  """Average duration (ms) of one launch of a step's stage, each group's
  launch bracketed by its own pair of HIP events on the stream it is
  launched on (the current stream):
    raster_bins   binning half of the rasterizer (clear, count, 2 scans,
                  fill) -- re-bins the current vertices;
    raster_tiles  the tile pass alone (raster_tile_kernel) over those bins;
    splat_target  fused target pass (target_kernel + the loss reduce);
    splat_source  fused source pass (source_kernel)."""
  import torch
  from diffren_amd import _native
  fn = {"raster_bins": lambda g: r._rasterize(g, _native.RASTER_BINS_ONLY),
        "raster_tiles": lambda g: r._rasterize(g, _native.RASTER_TILES_ONLY),
        "splat_target": lambda g: r._splat_forward(g, 1),
        "splat_source": r._splat_backward}[stage]
  pairs = []
  for _ in range(reps):
    for g in r.groups:
      a = torch.cuda.Event(enable_timing=True)
      b = torch.cuda.Event(enable_timing=True)
      a.record()
      fn(g)
      b.record()
      pairs.append((a, b))
  torch.cuda.synchronize()
  return sum(a.elapsed_time(b) for a, b in pairs) / len(pairs)  # ms


def copy_ceiling(dev, mib=256, reps=10):
  """Measured device-to-device copy rate (GB/s, read + write bytes) of a
  `mib` MiB buffer: the attainable HBM rate next to the datasheet peak."""
  import torch
  a = torch.ones(mib * 1024 * 1024 // 4, dtype=torch.float32, device=dev)
  b = torch.empty_like(a)
  b.copy_(a)
  torch.cuda.synchronize()
  start = torch.cuda.Event(enable_timing=True)
  end = torch.cuda.Event(enable_timing=True)
  start.record()
  for _ in range(reps):
    b.copy_(a)
  end.record()
  torch.cuda.synchronize()
  ms = start.elapsed_time(end) / reps
  return 2 * a.numel() * 4 / (ms * 1e-3) / 1e9


def time_stage(fn, reps):
  import torch
  start = torch.cuda.Event(enable_timing=True)
  end = torch.cuda.Event(enable_timing=True)
  fn()
  start.record()
  for _ in range(reps):
    fn()
  end.record()
  torch.cuda.synchronize()
  return start.elapsed_time(end) / reps  # ms


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "traffic.json")


def measured_traffic(stage, groups):
  """HBM bytes per launch of the stage's kernel from the committed
  rocprofv3 FETCH_SIZE / WRITE_SIZE passes (scaled by the copy-kernel
  calibration recorded with them), or None when not measured."""
  if not os.path.exists(TRAFFIC_FILE):
    return None
  with open(TRAFFIC_FILE) as f:
    t = json.load(f)
  e = t.get("kernels", {}).get(stage)
  if not e:
    return None
  per_step = (e["fetch_kib"] * t["fetch_scale"] +
              e["write_kib"] * t["write_scale"]) * 1024 * e["launches"]
  return int(per_step / groups)


def pmc_probe(args):
  """Program for rocprofv3 --pmc passes: the fused step (eager) three times
  with the default job groups, then a 256 MiB device copy whose FETCH_SIZE /
  WRITE_SIZE calibrate the counters' units."""
  import torch
  from diffren_amd import batch
  dev = torch.device("cuda:0")
  w = build_workload(args, dev)
  R, L = args.resolution, args.layers
  jobs = [(0, w["cams_t"][i], R, R) for i in range(args.views)]
  # Two fixed groups of four views when not given (the count the autotuner
  # picks for this workload): the counters' per-launch figures then cover
  # the bench's launches.
  r = batch.BatchedSplatRenderer([w["mesh"]], jobs, L, tile=(16, 16),
                                 streams=(2 if args.streams == "auto" else
                                          args.streams),
                                 fused_transform=args.fused_transform,
                                 tile_threads=args.tile_threads)
  for t in r.targets:
    t.uniform_()
  for _ in range(3):
    r.step()
  torch.cuda.synchronize()
  a = torch.empty(64 * 1024 * 1024, dtype=torch.float32, device=dev)
  a.fill_(1.0)
  b = torch.empty_like(a)
  torch.cuda.synchronize()
  b.copy_(a)
  torch.cuda.synchronize()
  print("pmc probe done", flush=True)


def plain_step(w, args, fused=False):
  import torch
  from diffren_amd import render
  vv = w["mesh"]["vertices"].clone().requires_grad_(True)
  cc = w["mesh"]["colors"].clone().requires_grad_(True)
  npix = args.views * args.resolution ** 2 * 4
  for i in range(args.views):
    img = render.render_triangles(vv, {"c": cc}, w["mesh"]["triangles"],
                                  w["cams_t"][i], args.resolution,
                                  args.resolution, lambda a: a["c"],
                                  num_layers=args.layers, fused_splat=fused)
    loss = ((img - w["targets"][i]) ** 2).sum() / npix
    loss.backward()
  return vv.grad, cc.grad


def cpu_baseline(w, args, reps=3):
  """The workload's fwd+bwd on the host's cores, timed `reps` times over
  all of its views: the reference CPU rasterizer (its own source compiled
  here, single-threaded per image as rasterize_triangles_impl.cc:37 is; the
  views run concurrently, one per thread), then for each view the float32 C
  restatement of interpolate -> splat -> over -> L2 -> backward
  (oracle/splat_cpu.c) on all threads.  renders/s = views / the median
  step time."""
  import numpy as np
  import torch
  from oracle import raster_oracle
  from oracle import splat_cpu
  from oracle import splat_oracle as so
  kind = "reference" if raster_oracle.ref_available() else "port"
  ras = raster_oracle.ref_rasterize if kind == "reference" else \
      raster_oracle.rasterize
  R, L, B = args.resolution, args.layers, len(w["cams"])
  threads = int(torch.get_num_threads())
  t0 = time.perf_counter()
  clip = so.transform(w["cams"][0], w["v"], torch.float32)
  ras(clip.numpy(), w["t"], R, R, L, 1)
  raster_ms = (time.perf_counter() - t0) * 1e3
  targets = [t.cpu().numpy() for t in w["targets"]]
  steps, raster_par, splat_ms = [], None, []
  for _ in range(reps):
    t1 = time.perf_counter()
    raster_par, frames = cpu_raster_parallel(w, args, ras)
    t2 = time.perf_counter()
    for k, (ids, _, b) in enumerate(frames):
      splat_cpu.render(w["v"], w["colors"], w["cams"][k], w["t"], ids, b,
                       targets[k], 1.0 / (R * R * 4),
                       threads=threads)
    t3 = time.perf_counter()
    steps.append((t3 - t1) * 1e3)
    splat_ms.append((t3 - t2) * 1e3 / B)
  step_ms = float(np.median(steps))
  splat_view_ms = float(np.median(splat_ms))
  return dict(value=B / (step_ms * 1e-3), unit="renders/s", cores=threads,
              kind="port",
              sample=(f"{reps} timed runs of the workload's {B} views, "
                      f"fwd+bwd (median {step_ms:.0f} ms): the {kind} CPU "
                      f"rasterizer one view per thread on "
                      f"{raster_par['threads']} threads "
                      f"({raster_par['ms']:.0f} ms for {B}; "
                      f"{raster_ms:.0f} ms for one view alone), then each "
                      f"view's C float32 interpolate/splat/over/L2 fwd+bwd "
                      f"restatement on {threads} threads "
                      f"({splat_view_ms:.0f} ms per view)"),
              raster_ms=raster_ms, splat_ms=splat_view_ms, step_ms=step_ms,
              raster_kind=kind, raster_parallel=raster_par)


def cpu_raster_parallel(w, args, ras):
  """The CPU rasterizer on all views at once, one single-threaded image per
  host thread (the reference kernel is single-threaded per image,
  rasterize_triangles_impl.cc:37): views/s and the thread count, and the
  views' (ids, z, barycentrics).  ctypes releases the interpreter lock
  during each call, so the images run concurrently."""
  from concurrent.futures import ThreadPoolExecutor
  import torch
  from oracle import splat_oracle as so
  R, L = args.resolution, args.layers
  clips = [so.transform(c, w["v"], torch.float32).numpy() for c in w["cams"]]
  n = len(clips)
  threads = max(1, min(n, int(torch.get_num_threads())))
  with ThreadPoolExecutor(threads) as pool:
    t0 = time.perf_counter()
    frames = list(pool.map(lambda c: ras(c, w["t"], R, R, L, 1), clips))
    dt = time.perf_counter() - t0
  return dict(views=n, threads=threads, ms=dt * 1e3, views_per_s=n / dt), \
      frames


def finite_differences(r, vertices, target, loss_scale, count=64, eps=1e-3,
                       seed=3):
  """Central differences (rasterization frozen) of the fused path's L2 loss
  for `count` seeded vertices with a non-zero gradient, x 3 coordinates,
  against r.d_vertices[0] from the last step.  Only the pixels a vertex
  touches change between the two renders, so their float64 difference is
  the exact local loss difference of the float32 images.

  An entry "crosses the kink" when, between its two renders and the
  unperturbed one, some pixel's accumulated splat weight of some layer class
  moves across 1, where the normalisation 1 / (max(w - 1, 0) + 1)
  (splat.py:168-170) changes branch: the quotient then mixes two branches of
  the loss.  Needs r built with return_accum_buffers=True.

  Returns the relative error ||fd - g|| / ||g|| over all entries and over
  the entries that cross no kink, and the number that do."""
  import numpy as np
  import torch
  g = r.d_vertices[0].clone()
  base_side = r.accum_weights[0] > 1.0
  cand = torch.nonzero(g.abs().sum(1) > 0).squeeze(1).cpu().numpy()
  sel = np.random.default_rng(seed).choice(cand, min(count, cand.size),
                                           replace=False)
  base = vertices.clone()
  tgt = target.double()
  fd, an, kink = [], [], []
  for v in sel:
    for c in range(3):
      losses, pos = [], []
      crossed = False
      for sgn in (1.0, -1.0):
        vertices.copy_(base)
        vertices[v, c] += sgn * eps
        pos.append(float(vertices[v, c]))
        r.transform()
        r.splat_forward(0)
        losses.append(((r.images[0].double() - tgt) ** 2))
        crossed |= bool(((r.accum_weights[0] > 1.0) != base_side).any())
      d = (losses[0] - losses[1]).sum().item() * loss_scale
      fd.append(d / (pos[0] - pos[1]))
      an.append(float(g[v, c]))
      kink.append(crossed)
  vertices.copy_(base)
  r.transform()
  r.splat_forward(0)
  fd, an, kink = np.array(fd), np.array(an), np.array(kink)
  free = ~kink
  return dict(
      rel_err=float(np.linalg.norm(fd - an) / np.linalg.norm(an)),
      rel_err_kink_free=(float(np.linalg.norm((fd - an)[free]) /
                               np.linalg.norm(an[free]))
                         if free.any() else None),
      entries=int(fd.size), crossing_kink=int(kink.sum()))


def grad_check(w, args, dev):
  """Vertex/colour gradient of view 0: fused GPU path vs float64 oracle."""
  import numpy as np
  import torch
  from diffren_amd import batch
  from oracle import raster_oracle
  from oracle import splat_oracle as so
  R, L = args.resolution, args.layers
  r = batch.BatchedSplatRenderer([w["mesh"]], [(0, w["cams_t"][0], R, R)], L,
                                 loss_scale=1.0 / (R * R * 4),
                                 return_accum_buffers=True,
                                 deterministic=args.deterministic)
  r.targets[0].copy_(w["targets"][0])
  r.step()
  torch.cuda.synchronize()
  ids = r.ids[0].cpu().numpy()
  b = r.barys[0].cpu().numpy()
  clip32 = so.transform(w["cams"][0], w["v"], torch.float32).numpy()
  ids_o, _, b_o = raster_oracle.rasterize(clip32, w["t"], R, R, L, 1)
  fg = ids_o >= 0
  # Sparse raster outputs: barycentrics are defined under triangles only.
  ids_exact = bool(np.array_equal(ids, ids_o) and np.array_equal(
      b.view(np.uint32)[fg], b_o.view(np.uint32)[fg]))
  v64 = torch.as_tensor(w["v"], dtype=torch.float64).requires_grad_(True)
  c64 = torch.as_tensor(w["colors"], dtype=torch.float64).requires_grad_(True)
  out = so.render_splat(v64, c64, w["cams"][0], ids_o, b_o, w["t"])
  loss = ((out - w["targets"][0].cpu().double()) ** 2).sum() / (R * R * 4)
  loss.backward()

  def rel(a, b):
    a = a.cpu().double()
    return float((a - b).abs().max() / b.abs().max())

  # Image of view 0 (before the difference quotients below re-render it with
  # moved vertices) against the float32 oracle restatement, which forms the
  # viewport coordinates and splat offsets in the reference's float32 order
  # (splat.py:103-104, 218-221): the fused path and the plain per-view path.
  # Against float64 both carry the float32 rounding of x ~ 1000 px.
  from diffren_amd import render
  out64 = out.detach()
  with torch.no_grad():
    out32 = so.render_splat(torch.as_tensor(w["v"]),
                            torch.as_tensor(w["colors"]), w["cams"][0], ids_o,
                            b_o, w["t"])
    plain = render.render_triangles(
        w["mesh"]["vertices"], {"c": w["mesh"]["colors"]},
        w["mesh"]["triangles"], w["cams_t"][0], R, R, lambda a: a["c"],
        num_layers=L, fused_splat=False)
  image_err = rel(r.images[0], out32.double())
  plain_image_err = rel(plain, out32.double())
  image_err64 = rel(r.images[0], out64)
  plain_image_err64 = rel(plain, out64)
  fused_equals_plain = bool(torch.equal(r.images[0], plain))

  # Central differences at eps 1e-3 (0.3-0.44 px of fragment motion here,
  # enough to move neighbouring pixels' accumulated weights across the
  # normalisation kink at 1, 5 % below the full-coverage weight 1.05), and
  # at 3e-4 and 1e-4, below which no entry crosses it.
  fd = {str(eps): finite_differences(r, w["mesh"]["vertices"],
                                     w["targets"][0], 1.0 / (R * R * 4),
                                     eps=eps) for eps in (1e-3, 3e-4, 1e-4)}
  return dict(vertex_grad_rel_err=rel(r.d_vertices[0], v64.grad),
              vertex_grad_fd_rel_err=fd,
              color_grad_rel_err=rel(r.d_colors[0], c64.grad),
              loss_rel_err=abs(r.loss.item() - loss.item()) / abs(loss.item()),
              image_rel_err=image_err,
              plain_image_rel_err=plain_image_err,
              image_rel_err_vs_float64=image_err64,
              plain_image_rel_err_vs_float64=plain_image_err64,
              fused_image_equals_plain=fused_equals_plain,
              triangle_ids_bit_exact=ids_exact)


def main():
  args = parse()
  if args.gpus != 1:
    print(f"bench.py runs on one GPU only; --gpus {args.gpus} is not "
          "supported (views and meshes would shard across GPUs with one "
          "gradient all-reduce, which this build does not implement)",
          file=sys.stderr)
    sys.exit(2)
  if args.pmc_probe:
    pmc_probe(args)
    return
  import numpy as np
  import torch
  from diffren_amd import batch

  dev = torch.device("cuda:0")
  log("building workload")
  w = build_workload(args, dev)
  R, L, B = args.resolution, args.layers, args.views
  T = w["t"].shape[0]
  V = w["v"].shape[0]
  C = 4
  jobs = [(0, w["cams_t"][i], R, R) for i in range(B)]

  # Targets: forward renders of the perturbed mesh.
  tr = batch.BatchedSplatRenderer([w["target_mesh"]], jobs, L)
  tr.render()
  torch.cuda.synchronize()
  w["targets"] = [img.clone() for img in tr.images]
  del tr

  r = batch.BatchedSplatRenderer([w["mesh"]], jobs, L, tile="auto",
                                 streams=args.streams,
                                 fused_transform=args.fused_transform,
                                 deterministic=args.deterministic,
                                 tile_threads=args.tile_threads)
  log(f"autotuned raster tile {r.raster_config.tile_w}x"
      f"{r.raster_config.tile_h}, bin capacity {r.raster_config.bin_capacity}"
      f", timings {r.autotune_timings}; {len(r.groups)} job groups"
      f" (step ms by group count: {r.streams_timings}); tile pass in "
      f"{r.raster_config.tile_threads}-thread blocks (step ms: "
      f"{r.tile_threads_timings})")
  for i in range(B):
    r.targets[i].copy_(w["targets"][i])
  r.step()
  entries, overflow = r.bin_stats()
  log(f"bin entries {entries} overflow {overflow}")

  use_graph = not args.no_graph
  if use_graph:
    r.capture()
    run = r.replay
  else:
    run = r.step
  for _ in range(args.warmup):
    run()
  torch.cuda.synchronize()
  t0 = time.perf_counter()
  for _ in range(args.steps):
    run()
  torch.cuda.synchronize()
  dt = (time.perf_counter() - t0) / args.steps
  renders_per_s = B / dt
  log(f"batched fused: {dt * 1e3:.3f} ms/step, {renders_per_s:.1f} renders/s")

  # Per-stage HIP-event times (eager, same stream as the kernels).
  reps = 10
  stage_ms = dict(
      transform=time_stage(r.transform, reps),
      rasterize=time_stage(r.rasterize, reps),
      splat_target=time_stage(lambda: r.splat_forward(1), reps),
      splat_source=time_stage(r.splat_backward, reps))
  log("stage ms (all views): " + json.dumps(stage_ms))

  # Algorithmic bytes per view of each stage, fp32/i32, API-level inputs
  # and outputs counted once, no intermediate scratch:
  #   rasterize  T*12 + V*16 read (triangles, clip vertices), L*P*20
  #              written (ids, z, barycentrics);
  #   target     (interpolate -> splat -> over -> L2 gradient) L*P*20 (ids,
  #              z, barycentrics) + T*12 + V*16 + V*4C (clip vertices,
  #              colours) + P*4C (target image) read, P*4C written (image);
  #   source     (the VJP) L*P*16 (ids, barycentrics) + T*12 + V*16 + V*4C
  #              + P*4C (d image) read, V*4*(3 + C) written (d vertices,
  #              d colours; once per mesh per launch).
  # The gradient scratch between the two splat passes is an intermediate and
  # is not counted.
  P = R * R
  lp = L * P
  ng = len(r.groups)
  vpl = B / ng  # views per launch
  per_launch = {
      "raster_tiles": vpl * (T * 12 + V * 16 + lp * 20),
      "splat_target": vpl * (lp * 20 + T * 12 + V * 16 + V * 4 * C +
                             P * 4 * C + P * 4 * C),
      "splat_source": vpl * (lp * 16 + T * 12 + V * 16 + V * 4 * C +
                             P * 4 * C) + V * 4 * (3 + C),
  }
  kernel_names = {
      "raster_tiles": "raster_tile_kernel<%d, %d, %d>" % (
          r.raster_config.tile_w, r.raster_config.tile_h,
          r.raster_config.tile_threads or 256),
      "splat_target": "target_kernel<4, false>",
      "splat_source": "source_kernel<4, false>",
  }
  launch_ms = {k: time_launches(r, k, reps)
               for k in ("raster_bins", "raster_tiles", "splat_target",
                         "splat_source")}
  ceiling = copy_ceiling(dev)
  kernels = {}
  for k, nbytes in per_launch.items():
    gbs = nbytes / (launch_ms[k] * 1e-3) / 1e9
    kernels[k] = dict(kernel=kernel_names[k],
                      avg_launch_ms=round(launch_ms[k], 4),
                      algorithmic_bytes_per_launch=int(nbytes),
                      achieved_gbs=round(gbs, 1),
                      frac=round(gbs / HBM_PEAK_GBS, 4),
                      frac_of_copy_ceiling=round(gbs / ceiling, 4),
                      traffic=measured_traffic(k, ng))
  log("launch ms: " + json.dumps(launch_ms))
  dom = max(per_launch, key=lambda k: launch_ms[k])
  d = kernels[dom]
  roofline = dict(bound="hbm", kernel=d["kernel"], achieved=d["achieved_gbs"],
                  peak=HBM_PEAK_GBS, unit="GB/s", frac=d["frac"],
                  traffic=d["traffic"],
                  algorithmic_bytes_per_launch=d["algorithmic_bytes_per_launch"],
                  avg_launch_ms=d["avg_launch_ms"], launches_per_step=ng,
                  copy_ceiling_gbs=round(ceiling, 1),
                  binning_half_ms=round(launch_ms["raster_bins"], 4),
                  kernels=kernels)

  # Per-view paths: one render_triangles per view + autograd (one warm-up
  # step each): unfused passes (plain), and the splat on the fused kernels.
  per_view_rps = {}
  for fused in (False, True):
    plain_step(w, args, fused)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.plain_steps):
      plain_step(w, args, fused)
    torch.cuda.synchronize()
    pdt = (time.perf_counter() - t0) / args.plain_steps
    per_view_rps[fused] = B / pdt
    log(f"per view ({'fused splat' if fused else 'plain'}): "
        f"{pdt * 1e3:.3f} ms/step, {B / pdt:.1f} renders/s")
  plain_rps = per_view_rps[False]

  extra = {}
  if not args.no_grad_check:
    log("float64 gradient check (view 0)")
    extra.update(grad_check(w, args, dev))
    log(json.dumps(extra))
  cpu = None
  if not args.no_cpu_baseline:
    log("CPU baseline (view 0)")
    cpu = cpu_baseline(w, args)
    log(json.dumps(cpu))

  out = {
      "metric": ("fwd+bwd renders/s @1M tris 1024² (batched vs plain "
                 "GPU); Mtri/s; grad rel err"),
      "value": round(renders_per_s, 2),
      "unit": "renders/s",
      "n_gpus": 1,
      "steps": args.steps,
      "warmup": args.warmup,
      "ms_per_step": round(dt * 1e3, 4),
      "higher_is_better": True,
      "scaling": "weak",
      "vs_baseline": None,
      "dtype": "f32",
      "data": "synthetic (seeded Perlin-displaced UV sphere, seeded cameras)",
      "config": {"workload": "1M-triangle sphere, 1024x1024, 3 layers, 8 "
                             "views, RGBA, fwd + L2 loss + bwd, batched "
                             f"fused path, {len(r.groups)} job groups on "
                             f"as many streams" +
                             (" (autotuned)" if args.streams == "auto"
                              else "") +
                             (", transform on load" if args.fused_transform
                              else "") +
                             (", deterministic gradients" if
                              args.deterministic else "") +
                             (" in a HIP graph" if use_graph else ""),
                 "triangles": T, "vertices": V, "resolution": R, "layers": L,
                 "views": B, "channels": C, "job_groups": len(r.groups),
                 "tile_threads": r.raster_config.tile_threads,
                 "parallelism": "single GPU"},
      "mtri_per_s": round(renders_per_s * T / 1e6, 1),
      "plain_renders_per_s": round(plain_rps, 2),
      "speedup_vs_plain": round(renders_per_s / plain_rps, 2),
      "per_view_fused_renders_per_s": round(per_view_rps[True], 2),
      "speedup_vs_per_view_fused": round(renders_per_s / per_view_rps[True],
                                         2),
      "stage_ms": {k: round(v, 4) for k, v in stage_ms.items()},
      "roofline": roofline,
      "cpu_baseline": cpu,
  }
  out.update(extra)
  print(json.dumps(out), flush=True)


if __name__ == "__main__":
  main()
