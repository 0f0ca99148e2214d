"""Builds the gfx950 HIP library libdiffren_amd.so in-tree with hipcc.

Every .hip file under diffren_amd/csrc is compiled for gfx950 only and linked
into one shared library exporting the C ABI declared in include/diffren_amd.h.
Floating-point contraction is off library-wide: the rasterizer's bit-exact
parity depends on it, and interpolation/splat follow the reference's unfused
FP32 expression order.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libdiffren_amd.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
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


def build(verbose=False):
  tmp = LIB + ".tmp"
  cmd = [HIPCC] + FLAGS + sources() + ["-o", tmp]
  if verbose:
    print(" ".join(cmd), flush=True)
  res = subprocess.run(cmd, capture_output=True, text=True)
  if res.returncode != 0:
    sys.stderr.write(res.stdout + res.stderr)
    raise RuntimeError("hipcc failed building libdiffren_amd.so")
  os.replace(tmp, LIB)
  return LIB


if __name__ == "__main__":
  print(build(verbose=True))
