"""Exposure of the oracle's two unpinned details on config-0 inputs.

No fixture available here pins (1) the exact form of jaxopt's line-search
acceptance test, in particular its ``+ eps`` slack, or (2) the tangent frame
of MJX's plane-capsule contacts (only the normal, frame[0], is pinned, by
brax/contact_test.py:40,74).  This test runs the oracle's C++ restatement
(oracle/cstep.cpp, pinned to the torch oracle to 1e-9) on 256 seeded
quadrupeds for 20 env steps under the alternative readings and prints how far
q and qd move, in units of the parity band |d| <= 1e-5 + 1e-4 |ref|:

* ls_eps = 0 instead of float32 eps: bitwise the same trajectories (the
  slack is far below every objective change the solver compares);
* capsule tangent = orthogonals(n) (the sphere rule) instead of the capsule
  axis projected on the plane: 44% of envs leave the band after one env
  step.  The frame is material, and it stays parity unpinned.
"""

import numpy as np

import bench
from oracle.cstep import CStep

N, STEPS = 256, 20


def _traj(cs, q0, qd0, acts, dtype):
  q, qd, out = q0.astype(dtype), qd0.astype(dtype), []
  for t in range(STEPS):
    q, qd = cs.rollout(q, qd, acts[t:t + 1].astype(dtype), 5, dtype=dtype)
    out.append((q.astype(np.float64), qd.astype(np.float64)))
  return out


def _band(a, b):
  r = lambda x, y: np.abs(x - y) / (1e-5 + 1e-4 * np.abs(y))  # noqa: E731
  return np.maximum(r(a[0], b[0]).max(1), r(a[1], b[1]).max(1))


def test_unpinned_details_exposure():
  sys_, q0, qd0, _ = bench.workload("quadruped", N)
  acts = np.random.default_rng(5).uniform(-1, 1, (STEPS, N, 8))
  base = _traj(CStep(sys_), q0, qd0, acts, np.float64)
  runs = {
      "ls_eps=0, float64": _traj(CStep(sys_, ls_eps=0.0), q0, qd0, acts, np.float64),
      "ls_eps=0, float32": _traj(CStep(sys_, ls_eps=0.0), q0, qd0, acts, np.float32),
      "float32 (for scale)": _traj(CStep(sys_), q0, qd0, acts, np.float32),
      "capsule tangent = orthogonals(n), float64":
          _traj(CStep(sys_, capsule_frame="orthogonals"), q0, qd0, acts, np.float64),
  }
  out = {}
  for name, tr in runs.items():
    for t in (0, 19):
      r = _band(tr[t], base[t])
      out[name, t + 1] = (float(np.mean(r > 1)), float(np.median(r)), float(r.max()))
      print(f"{name:42s} after {t + 1:2d} env steps: {100 * np.mean(r > 1):5.1f}% of envs "
            f"outside the band, median {np.median(r):.3g}, max {r.max():.3g}")
  # the slack changes no decision: float64 trajectories are bitwise equal,
  # float32 ones equal the float32 run with the slack
  assert out["ls_eps=0, float64", 1][2] == 0 and out["ls_eps=0, float64", 20][2] == 0
  assert out["ls_eps=0, float32", 20] == out["float32 (for scale)", 20]
  # the capsule frame is material within one env step
  assert out["capsule tangent = orthogonals(n), float64", 1][0] > 0.1
