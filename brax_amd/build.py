"""Build recipe for libbrax_gen.so (gfx950), in-tree.

Regenerates the topology header from the registered systems, then compiles
brax_amd/csrc/bgx.hip with hipcc into brax_amd/libbrax_gen.so.
"""

from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libbrax_gen.so")
ARCH = os.environ.get("BGX_OFFLOAD_ARCH", "gfx950")

HIPCC_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", "-shared",
    f"--offload-arch={ARCH}",
    # packed FP32 (v_pk_*_f32) stays off: plain FP32 VALU only
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def hipcc() -> str:
  for c in (os.environ.get("HIPCC"), shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
    if c and os.path.exists(c):
      return c
  raise RuntimeError("hipcc not found")


def build(verbose: bool = False, resource_usage: bool = False) -> str:
  sys.path.insert(0, os.path.dirname(HERE))
  from brax_amd import codegen
  codegen.write_topology_header(os.path.join(CSRC, "topologies.inc"))
  tmp = LIB + ".tmp"
  cmd = [hipcc()] + HIPCC_FLAGS + [os.path.join(CSRC, "bgx.hip"), "-o", tmp]
  if resource_usage:
    cmd.append("-Rpass-analysis=kernel-resource-usage")
  if verbose:
    print(" ".join(cmd), flush=True)
  res = subprocess.run(cmd, capture_output=True, text=True)
  if res.returncode != 0:
    raise RuntimeError(f"hipcc failed ({res.returncode}):\n{res.stderr[-8000:]}")
  if resource_usage or verbose:
    print(res.stderr, flush=True)
  os.replace(tmp, LIB)
  return LIB


if __name__ == "__main__":
  build(verbose=True, resource_usage="--usage" in sys.argv)
