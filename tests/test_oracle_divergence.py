"""The humanoid workload diverges in the oracle too.

The humanoid bench line (16384 synthetic humanoids, q noise U(-0.01, 0.01),
qd U(-0.01, 0.01), actions U(-0.4, 0.4), 1000 env steps of 5 substeps)
reports about 2% of envs non-finite at the end.  This test runs the float64
oracle (oracle/cstep.cpp, pinned to the torch oracle) on 1024 envs of the
same input distribution (the bench's seeded states, numpy-seeded actions)
and finds a like fraction diverging: the humanoids sink through the floor
under random torques (8 projected-gradient iterations cannot hold the deep
contacts) and then blow up.  The divergence belongs to the reference
algorithm on these inputs, not to float32 arithmetic or the kernel.
"""

import numpy as np

import bench
from oracle.cstep import CStep


def test_humanoid_workload_diverges_in_the_float64_oracle():
  sys_, q0, qd0, _ = bench.workload("humanoid", 16384)
  n, steps = 1024, 1000
  acts = np.random.default_rng(1).uniform(-0.4, 0.4, (steps, n, sys_.act_size()))
  acts = acts.astype(np.float32).astype(np.float64)
  q, qd = CStep(sys_).rollout(q0[:n].astype(np.float64), qd0[:n].astype(np.float64), acts, 5,
                              dtype=np.float64)
  bad = ~(np.isfinite(q).all(1) & np.isfinite(qd).all(1))
  sunk = np.isfinite(q[:, 2]) & (q[:, 2] < 0.0)
  print(f"\nfloat64 oracle, {n} humanoids x {steps} env steps: {int(bad.sum())} non-finite "
        f"({100 * bad.mean():.1f}%), {int(sunk.sum())} more with the pelvis below the floor",
        flush=True)
  assert bad.mean() >= 0.01
