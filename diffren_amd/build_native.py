"""Builds the gfx950 HIP library libdiffren_amd.so in-tree with hipcc.

Every .hip file under diffren_amd/csrc is compiled for gfx950 only (one
hipcc process per file, in parallel, objects under diffren_amd/_build) and
linked into one shared library exporting the C ABI declared in
include/diffren_amd.h.  Floating-point contraction is off library-wide: the
rasterizer's bit-exact parity depends on it, and interpolation/splat follow
the reference's unfused FP32 expression order.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libdiffren_amd.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    "-ffp-contract=off", "-fno-fast-math",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # Matrix-core results in ordinary VGPRs: the MLP kernels read every
    # accumulator back for ReLU / bfloat16 conversion.
    "-mllvm", "-amdgpu-mfma-vgpr-form=1",
    "-I" + os.path.join(ROOT, "include"), "-I" + CSRC,
    "-Wno-unused-command-line-argument",
]


def sources():
  return sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC)
                if f.endswith(".hip"))


def _compile(src, verbose):
  obj = os.path.join(OBJ, os.path.basename(src)[:-4] + ".o")
  cmd = [HIPCC] + FLAGS + ["-c", src, "-o", obj]
  if verbose:
    print(" ".join(cmd), flush=True)
  res = subprocess.run(cmd, capture_output=True, text=True)
  return obj, res


def build(verbose=False, jobs=None):
  os.makedirs(OBJ, exist_ok=True)
  srcs = sources()
  jobs = jobs or min(len(srcs), max(1, (os.cpu_count() or 1)), 16)
  with ThreadPoolExecutor(jobs) as pool:
    results = list(pool.map(lambda s: _compile(s, verbose), srcs))
  failed = [r for _, r in results if r.returncode != 0]
  for r in failed:
    sys.stderr.write(r.stdout + r.stderr)
  if failed:
    raise RuntimeError("hipcc failed building libdiffren_amd.so")
  tmp = LIB + ".tmp"
  cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC"] + \
      [o for o, _ in results] + ["-o", tmp]
  if verbose:
    print(" ".join(cmd), flush=True)
  res = subprocess.run(cmd, capture_output=True, text=True)
  if res.returncode != 0:
    sys.stderr.write(res.stdout + res.stderr)
    raise RuntimeError("hipcc failed linking libdiffren_amd.so")
  os.replace(tmp, LIB)
  return LIB


if __name__ == "__main__":
  print(build(verbose=True))
