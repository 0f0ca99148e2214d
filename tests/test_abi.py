"""The C-ABI library loads and exports every entry point include/*.h declares.

No compute calls: these run without a GPU.
"""
import ctypes
import glob
import os
import re

from diffren_amd import _native

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
  names = set()
  for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
    text = open(h).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names |= set(re.findall(r"\b(dr_[a-z0-9_]+)\s*\(", text))
  return names


def test_library_exports_every_declared_symbol():
  lib = ctypes.CDLL(_native.LIB_PATH)
  names = declared_symbols()
  assert len(names) >= 10
  missing = [n for n in sorted(names) if not hasattr(lib, n)]
  assert not missing, missing


def test_python_binding_covers_every_declared_symbol():
  assert declared_symbols() <= set(_native.SIGNATURES)


def test_host_only_entry_points():
  lib = _native.lib()
  assert lib.dr_version().startswith(b"diffren_amd")
  assert abs(lib.dr_splat_coefficient() - 0.6503143820850361) < 1e-7
  cfg = _native.RasterConfig(32, 8, 0)
  n = lib.dr_raster_workspace_size(1000, 100, ctypes.byref(cfg))
  assert n >= 4 * 1000
  # Bad tile size is rejected by the host-side preparation.
  job = (_native.RasterJob * 1)()
  job[0].width = 16
  job[0].height = 16
  job[0].triangle_ids = 256
  job[0].z_buffer = 256
  bad = _native.RasterConfig(7, 7, 0)
  rc = lib.dr_raster_jobs_prepare(job, 1, ctypes.byref(bad), None, None)
  assert rc != 0 and b"tile" in lib.dr_last_error()
  ok = lib.dr_raster_jobs_prepare(job, 1, ctypes.byref(cfg), None, None)
  assert ok == 0
  # The tile pass's block size: 0 (256), 128 or 256.
  for t, good in ((0, True), (128, True), (256, True), (64, False),
                  (512, False)):
    c = _native.RasterConfig(32, 8, 0, 0, t)
    rc = lib.dr_raster_jobs_prepare(job, 1, ctypes.byref(c), None, None)
    assert (rc == 0) == good, t
    if not good:
      assert b"tile_threads" in lib.dr_last_error()
  assert lib.dr_splat_bwd_workspace_size(4, 5, 3) == 4 * 5 * 3 * 4 * 4


def test_header_enums_match_python():
  from diffren_amd import constants
  text = open(os.path.join(ROOT, "include", "diffren_amd.h")).read()
  assert "DR_CULL_NONE = 0, DR_CULL_BACK = 1, DR_CULL_FRONT = 2" in text
  assert [int(m) for m in constants.FaceCullingMode] == [0, 1, 2]
  # Raster flag bits: header values against the Python mirrors, and no two
  # flags sharing a bit.
  import re
  flags = {n: int(v) for n, v in
           re.findall(r"DR_RASTER_([A-Z_]+) = (\d+)", text)}
  assert flags == {"SPARSE": _native.RASTER_SPARSE,
                   "BINS_ONLY": _native.RASTER_BINS_ONLY,
                   "TILES_ONLY": _native.RASTER_TILES_ONLY,
                   "PREFIX": _native.RASTER_PREFIX}
  assert sum(flags.values()) == (1 | 2 | 4 | 8)


def test_struct_layouts_match_header(tmp_path):
  """Every ctypes mirror has the header struct's size and field offsets
  (a C program built from include/diffren_amd.h with gcc prints them)."""
  import subprocess
  mirrors = {"dr_raster_job": _native.RasterJob,
             "dr_raster_config": _native.RasterConfig,
             "dr_splat_job": _native.SplatJob,
             "dr_phong_params": _native.PhongParams}
  lines = ['#include <stddef.h>', '#include <stdio.h>',
           '#include "diffren_amd.h"', 'int main(void) {']
  for name, cls in mirrors.items():
    lines.append(f'  printf("{name} size %zu\\n", sizeof({name}));')
    for field, _ in cls._fields_:
      lines.append(f'  printf("{name} {field} %zu\\n", '
                   f'offsetof({name}, {field}));')
  lines += ['  return 0;', '}']
  src = tmp_path / "layout.c"
  src.write_text("\n".join(lines) + "\n")
  exe = tmp_path / "layout"
  subprocess.run(["gcc", "-std=c11", "-I", os.path.join(ROOT, "include"),
                  str(src), "-o", str(exe)], check=True)
  got = subprocess.run([str(exe)], check=True, capture_output=True,
                       text=True).stdout.split("\n")
  want = []
  for name, cls in mirrors.items():
    want.append(f"{name} size {ctypes.sizeof(cls)}")
    for field, _ in cls._fields_:
      want.append(f"{name} {field} {getattr(cls, field).offset}")
  assert got[:-1] == want


def test_render_channel_counts_and_job_checks():
  lib = _native.lib()
  for c in (1, 4, 8, 9, 12, 13, 16, 20, 24, 28):
    assert lib.dr_render_record_size(c) > 0
    assert lib.dr_render_scratch_floats(c) > 0
  for c in (0, 17, 21, 29, 32):
    assert lib.dr_render_record_size(c) == 0
    assert lib.dr_render_scratch_floats(c) == 0
  job = (_native.SplatJob * 1)()
  job[0].width = 40
  job[0].height = 20
  job[0].triangle_ids = 256
  job[0].z_buffer = 256
  job[0].barycentrics = 256
  tiles = ctypes.c_int64()
  assert lib.dr_splat_jobs_prepare(job, 1, 4, ctypes.byref(tiles)) == 0
  assert tiles.value == 2 * 3  # 32x8 tiles: two columns, three rows
  assert lib.dr_splat_jobs_prepare(job, 1, 17, ctypes.byref(tiles)) != 0
  assert b"channel" in lib.dr_last_error()
  job[0].grad_scratch = 260
  assert lib.dr_splat_jobs_prepare(job, 1, 4, ctypes.byref(tiles)) != 0
  assert b"grad_scratch" in lib.dr_last_error()
  job[0].grad_scratch = 0
  # Frames of C % 4 == 0 channels move as 16-byte rows: a misaligned image
  # is refused there, and taken for other channel counts.
  job[0].image = 264
  assert lib.dr_splat_jobs_prepare(job, 1, 4, ctypes.byref(tiles)) != 0
  assert b"16-byte" in lib.dr_last_error()
  assert lib.dr_splat_jobs_prepare(job, 1, 3, ctypes.byref(tiles)) == 0
  job[0].image = 272
  assert lib.dr_splat_jobs_prepare(job, 1, 4, ctypes.byref(tiles)) == 0
  job[0].image = 0
  job[0].num_triangles = 3
  assert lib.dr_splat_jobs_prepare(job, 1, 4, ctypes.byref(tiles)) != 0
  assert b"null mesh" in lib.dr_last_error()
