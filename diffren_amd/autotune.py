"""Autotuner for the batched render: screen-tile shape, bin capacity, the
number of concurrent job groups and the tile pass's block size.

For a batch of raster jobs it times every candidate tile shape with HIP
events on the current stream (after one warm-up launch), keeps the fastest,
then sizes the bin capacity from the entry count that configuration actually
needed (dr_raster_bin_stats) plus headroom, so the default over-allocation
(3 entries per triangle) shrinks to what the workload uses.  tune_groups
times the whole captured step for each candidate number of job groups (each
group's kernels run on their own stream, so groups overlap) and keeps the
fastest; tune_tile_threads then times that step with the tile pass in 256-
and 128-thread blocks (128 interleaves better with the other groups' kernels
on sparse frames, 256 is faster on dense ones).  Results are cached in memory per workload signature (image sizes,
triangle counts, layers, culling) and optionally in a JSON file.
"""
import ctypes
import json
import os

import torch

from diffren_amd import _native

CANDIDATE_TILES = [(8, 8), (16, 16), (32, 8), (32, 16), (64, 8), (32, 32)]
CANDIDATE_GROUPS = (1, 2, 4)
CANDIDATE_TILE_THREADS = (256, 128)
HEADROOM = 1.25

_cache = {}


def signature(jobs, num_layers, culling):
  """jobs: list of (width, height, num_triangles)."""
  return json.dumps([sorted(map(list, jobs)), int(num_layers), int(culling)])


def _load(path):
  if path and os.path.exists(path):
    with open(path) as f:
      _cache.update(json.load(f))


def _save(path):
  if path:
    os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
    with open(path, "w") as f:
      json.dump(_cache, f)


def time_launch(fn, reps):
  start = torch.cuda.Event(enable_timing=True)
  end = torch.cuda.Event(enable_timing=True)
  fn()
  start.record()
  for _ in range(reps):
    fn()
  end.record()
  torch.cuda.synchronize()
  return start.elapsed_time(end) / reps


def tune_raster(table_host, num_jobs, num_layers, culling, device,
                job_sizes, reps=5, cache_path=None, candidates=None, flags=0):
  """Returns (tile_w, tile_h, bin_capacity, timings_ms).

  table_host: ctypes array of dr_raster_job with the device pointers filled;
  job_sizes: list of (width, height, num_triangles) for the signature.
  """
  _load(cache_path)
  key = signature(job_sizes, num_layers, culling) + f"/{flags}"
  if key in _cache:
    e = _cache[key]
    return e["tile"][0], e["tile"][1], e["capacity"], e["timings"]
  lib = _native.lib()
  timings = {}
  best = None
  for tw, th in (candidates or CANDIDATE_TILES):
    if num_layers * tw * th * 8 > 65536:
      continue
    cfg = _native.RasterConfig(tw, th, 0, flags, 0)
    tt, tiles = ctypes.c_int64(), ctypes.c_int64()
    _native.check(lib.dr_raster_jobs_prepare(table_host, num_jobs,
                                             ctypes.byref(cfg),
                                             ctypes.byref(tt),
                                             ctypes.byref(tiles)),
                  "dr_raster_jobs_prepare")
    table = _native.table_to_device(table_host, device)
    ws = _native.workspace(lib.dr_raster_workspace_size(
        tt.value, tiles.value, ctypes.byref(cfg)), device)

    def launch():
      _native.check(lib.dr_rasterize_triangles_batched(
          _native.ptr(table), num_jobs, tt.value, tiles.value, num_layers,
          culling, ctypes.byref(cfg), _native.ptr(ws), ws.numel(),
          _native.stream_handle()), "dr_rasterize_triangles_batched")

    ms = time_launch(launch, reps)
    need, over = ctypes.c_int64(), ctypes.c_int32()
    _native.check(lib.dr_raster_bin_stats(_native.ptr(ws), ctypes.byref(need),
                                          ctypes.byref(over),
                                          _native.stream_handle()),
                  "dr_raster_bin_stats")
    timings[f"{tw}x{th}"] = ms
    if best is None or ms < best[0]:
      best = (ms, tw, th, need.value)
  _, tw, th, need = best
  capacity = int(need * HEADROOM) + 4096
  _cache[key] = dict(tile=[tw, th], capacity=capacity, timings=timings)
  _save(cache_path)
  return tw, th, capacity, timings


def _tune_step(key, field, candidates, run_with, reps, warmup, cache_path):
  """(choice, timings_ms): the candidate whose step runs fastest.
  run_with(c) prepares the renderer for candidate c and returns a function
  that enqueues one step (a graph replay); cached under key/field."""
  _load(cache_path)
  key = key + "/" + field
  if key in _cache:
    e = _cache[key]
    return e[field], e["timings"]
  timings = {}
  best = None
  for c in candidates:
    fn = run_with(c)
    for _ in range(warmup):
      fn()
    ms = time_launch(fn, reps)
    timings[str(c)] = ms
    if best is None or ms < best[0]:
      best = (ms, c)
  _cache[key] = {field: best[1], "timings": timings}
  _save(cache_path)
  return best[1], timings


def tune_groups(key, run_with, num_jobs, reps=40, warmup=5, cache_path=None,
                candidates=None):
  """Returns (groups, timings_ms): the number of job groups whose step runs
  fastest.  run_with(n) prepares the renderer for n groups and returns a
  function that enqueues one step (a graph replay); key identifies the
  workload (signature() plus whatever else fixes the step)."""
  cands = [n for n in (candidates or CANDIDATE_GROUPS) if n <= num_jobs]
  return _tune_step(key, "groups", cands, run_with, reps, warmup, cache_path)


def tune_tile_threads(key, run_with, reps=40, warmup=5, cache_path=None,
                      candidates=None):
  """Returns (threads, timings_ms): the tile-pass block size
  (dr_raster_config.tile_threads) whose step runs fastest; run_with(t) as
  for tune_groups."""
  return _tune_step(key, "tile_threads",
                    list(candidates or CANDIDATE_TILE_THREADS), run_with,
                    reps, warmup, cache_path)
