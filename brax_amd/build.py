"""Build recipe for libbrax_gen.so (gfx950), in-tree.

Regenerates the topology header from the registered systems, compiles
brax_amd/csrc/bgx_topo.hip once per topology (-DBGX_TOPO=<struct>) and
brax_amd/csrc/bgx.hip (the C ABI) in parallel, and links them into
brax_amd/libbrax_gen.so.

  python brax_amd/build.py [--usage] [--jobs N]
"""

from __future__ import annotations

import concurrent.futures as cf
import os
import re
import shutil
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libbrax_gen.so")
ARCH = os.environ.get("BGX_OFFLOAD_ARCH", "gfx950")

HIPCC_FLAGS = [
    "-O3", "-std=c++17", "-fPIC",
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


def _run(cmd, verbose):
  if verbose:
    print(" ".join(cmd), flush=True)
  res = subprocess.run(cmd, capture_output=True, text=True)
  if res.returncode != 0:
    raise RuntimeError(f"hipcc failed ({res.returncode}): {' '.join(cmd)}\n"
                       f"{res.stderr[-8000:]}")
  return res.stderr


def topology_structs(inc_path: str):
  text = open(inc_path).read()
  return re.findall(r"X\(\d+, (Topo_\w+)\)", text)


def build(verbose: bool = False, resource_usage: bool = False, jobs: int = 0) -> str:
  sys.path.insert(0, os.path.dirname(HERE))
  from brax_amd import codegen
  inc = os.path.join(CSRC, "topologies.inc")
  codegen.write_topology_header(inc)
  structs = topology_structs(inc)
  extra = ["-Rpass-analysis=kernel-resource-usage"] if resource_usage else []
  jobs = jobs or min(8, (os.cpu_count() or 4), len(structs) + 1)
  with tempfile.TemporaryDirectory(prefix="bgx_build_") as tmp:
    units = [(os.path.join(CSRC, "bgx.hip"), os.path.join(tmp, "bgx.o"), [])]
    for s in structs:
      units.append((os.path.join(CSRC, "bgx_topo.hip"), os.path.join(tmp, f"{s}.o"),
                    [f"-DBGX_TOPO={s}"]))
    cmds = [[hipcc()] + HIPCC_FLAGS + d + extra + ["-c", src, "-o", obj]
            for src, obj, d in units]
    with cf.ThreadPoolExecutor(jobs) as ex:
      logs = list(ex.map(lambda c: _run(c, verbose), cmds))
    out = LIB + ".tmp"
    _run([hipcc(), "-shared", "-fPIC"] +
         [o for _, o, _ in units] + ["-o", out], verbose)
  if resource_usage:
    for log in logs:
      print(log, flush=True)
  os.replace(out, LIB)
  return LIB


if __name__ == "__main__":
  jobs = 0
  if "--jobs" in sys.argv:
    jobs = int(sys.argv[sys.argv.index("--jobs") + 1])
  build(verbose=True, resource_usage="--usage" in sys.argv, jobs=jobs)
