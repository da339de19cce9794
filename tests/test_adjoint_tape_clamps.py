"""Memory safety of the adjoint's tape readers on the CPU.

The reverse sweep (stage A2) and the tangent pass (stage B) take loop counts
and indices from the tape stage A1 writes: the active-contact count, the
solver iterations, each slot's link, each candidate's slot and frame flag.
tests/host/adj_tape_check.cpp compiles the same device code for the host
(g++, address and undefined-behaviour sanitizers, every addressed array on
the heap at its exact size), tapes a contact-rich substep and runs A2 and B
on that tape and on copies whose integer fields hold NaN, infinities, huge,
negative and just-out-of-range values.  The kernels clamp every such field
(bgx_step.h tape_int), so no case may read or write out of bounds.
"""

import os
import shutil
import subprocess

import numpy as np
import pytest

from brax_amd import codegen
from brax_amd import systems as S

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "host", "adj_tape_check.cpp")


def _build(topo, out_dir):
  exe = os.path.join(out_dir, f"adj_check_{topo}")
  cmd = ["g++", "-std=c++17", "-O1", "-g", "-fsanitize=address,undefined",
         "-fno-sanitize-recover=all", "-DBGX_HOST_CHECK", f"-DBGX_TOPO=Topo_{topo}", SRC,
         "-o", exe]
  subprocess.run(cmd, check=True, capture_output=True, text=True)
  return exe


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
@pytest.mark.parametrize("name", ["quadruped", "ant"])
def test_corrupted_tapes_stay_in_bounds(name, tmp_path):
  sys_ = S.quadruped(0) if name == "quadruped" else S.ant()
  topo = codegen.Topology(sys_)
  exe = _build(topo.name, str(tmp_path))
  rng = np.random.default_rng(0)
  q = sys_.init_q.copy()
  q[2] -= 0.15  # torso lowered: several feet and the torso in contact
  qd = 0.1 * rng.standard_normal(sys_.qd_size())
  act = rng.uniform(-1, 1, sys_.act_size())
  inp = os.path.join(str(tmp_path), "inputs.bin")
  np.concatenate([codegen.pack_model(sys_, topo), q, qd, act]).astype(np.float32).tofile(inp)
  env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0")
  res = subprocess.run([exe, inp], capture_output=True, text=True, env=env, timeout=600)
  print(res.stdout)
  assert res.returncode == 0, res.stdout + res.stderr[-4000:]
  assert "valid tape:" in res.stdout and "no out-of-bounds access" in res.stdout
  n_c = int(res.stdout.split("valid tape: ")[1].split()[0])
  assert n_c >= 2  # the valid tape exercises the slot paths
