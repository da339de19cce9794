"""Build recipe for libbrax_gen.so (gfx950), in-tree.

Regenerates the topology header from the registered systems, compiles
brax_amd/csrc/bgx_topo.hip once per tree topology (-DBGX_TOPO=<struct>),
brax_amd/csrc/bgx_chain.hip (the kernels of every hinge-chain topology) and
brax_amd/csrc/bgx.hip (the C ABI) in parallel, and links them into
brax_amd/libbrax_gen.so.

  python brax_amd/build.py [--usage] [--jobs N] [--cached]

--cached reuses object files under brax_amd/_build/ whose inputs (every
source under csrc/, the public header, the flags) are unchanged.
"""

from __future__ import annotations

import concurrent.futures as cf
import hashlib
import os
import re
import shutil
import subprocess
import sys

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


def topology_structs(inc_path: str, macro: str = "BGX_FOR_EACH_TREE_TOPOLOGY"):
  """The topology structs listed by one X-macro of topologies.inc (by
  default the tree topologies: one object file each)."""
  text = open(inc_path).read()
  start = text.index(f"#define {macro}(X)")
  end = text.find("\n\n", start)
  return re.findall(r"X\(\d+, (Topo_\w+)\)", text[start:end if end >= 0 else None])


CACHE = os.path.join(HERE, "_build")


def _inputs_digest(unit_src: str) -> str:
  """Digest of what one object file is built from: its own .hip file, the
  shared headers under csrc/ (every .h and the generated topologies.inc),
  the public header and the flags.  A change to one .hip file rebuilds only
  its objects."""
  h = hashlib.sha1()
  for name in sorted(os.listdir(CSRC)):
    if name.endswith(".hip"):
      continue
    h.update(name.encode())
    with open(os.path.join(CSRC, name), "rb") as f:
      h.update(f.read())
  with open(os.path.join(os.path.dirname(HERE), "include", "brax_gen.h"), "rb") as f:
    h.update(f.read())
  h.update(" ".join(HIPCC_FLAGS).encode())
  h.update(os.path.basename(unit_src).encode())
  with open(unit_src, "rb") as f:
    h.update(f.read())
  return h.hexdigest()[:16]


def build(verbose: bool = False, resource_usage: bool = False, jobs: int = 0,
          cached: bool = False) -> str:
  sys.path.insert(0, os.path.dirname(HERE))
  from brax_amd import codegen
  inc = os.path.join(CSRC, "topologies.inc")
  codegen.write_topology_header(inc)
  structs = topology_structs(inc)
  extra = ["-Rpass-analysis=kernel-resource-usage"] if resource_usage else []
  jobs = jobs or min(8, (os.cpu_count() or 4), len(structs) + 1)
  os.makedirs(CACHE, exist_ok=True)
  units = [(os.path.join(CSRC, "bgx.hip"), "bgx", []),
           (os.path.join(CSRC, "bgx_chain.hip"), "bgx_chain", [])]
  for s in structs:
    units.append((os.path.join(CSRC, "bgx_topo.hip"), s, [f"-DBGX_TOPO={s}"]))
  objs, cmds = [], []
  for src, name, d in units:
    obj = os.path.join(CACHE, f"{name}-{_inputs_digest(src)}.o")
    objs.append(obj)
    if cached and os.path.exists(obj) and not resource_usage:
      continue
    cmds.append([hipcc()] + HIPCC_FLAGS + d + extra + ["-c", src, "-o", obj + ".tmp"])
  with cf.ThreadPoolExecutor(jobs) as ex:
    logs = list(ex.map(lambda c: _run(c, verbose), cmds))
  for c in cmds:
    os.replace(c[-1], c[-1][:-4])
  out = LIB + ".tmp"
  _run([hipcc(), "-shared", "-fPIC"] + objs + ["-o", out], verbose)
  if resource_usage:
    for log in logs:
      print(log, flush=True)
  os.replace(out, LIB)
  # drop objects of other input versions
  keep = {os.path.basename(o) for o in objs}
  for f in os.listdir(CACHE):
    if f.endswith(".o") and f not in keep:
      os.remove(os.path.join(CACHE, f))
  return LIB


if __name__ == "__main__":
  jobs = 0
  if "--jobs" in sys.argv:
    jobs = int(sys.argv[sys.argv.index("--jobs") + 1])
  build(verbose=True, resource_usage="--usage" in sys.argv, jobs=jobs,
        cached="--cached" in sys.argv)
