"""The oracle's C++ restatement (oracle/cstep.cpp, the CPU baseline) against
the torch oracle: in float64 both follow the same algorithm and agree to
rounding on every compiled topology kind (free root with contacts and
limits, stacked hinge triples, contact-free chains and pendulums)."""

import numpy as np
import pytest
import torch

from brax_amd import systems as S
from oracle.cstep import CStep
from oracle.generalized import Oracle

CASES = [("quadruped", lambda: S.quadruped(0)), ("ant", S.ant),
         ("humanoid", lambda: S.humanoid(0)), ("chain7", lambda: S.chain(7, 1)),
         ("pendulum3", lambda: S.pendulum(3))]


@pytest.mark.parametrize("name,make", CASES)
def test_cstep_float64_matches_torch_oracle(name, make):
  sys = make()
  rng = np.random.default_rng(0)
  n, T = 8, 4
  q0 = sys.init_q + rng.uniform(-0.1, 0.1, (n, sys.q_size()))
  qd0 = 0.1 * rng.standard_normal((n, sys.qd_size()))
  acts = rng.uniform(-1, 1, (T, n, sys.act_size()))
  o = Oracle(sys, torch.float64, inverse="exact")
  st = o.init(q0, qd0)
  for t in range(T):
    st = o.env_step(st, acts[t], 5)
  q, qd = CStep(sys).rollout(q0, qd0, acts, 5, dtype=np.float64, threads=2)
  np.testing.assert_allclose(q, st.q.numpy(), rtol=1e-9, atol=1e-10)
  np.testing.assert_allclose(qd, st.qd.numpy(), rtol=1e-9, atol=1e-9)


def test_cstep_float32_close_to_float64_and_counts_rows():
  sys = S.quadruped(0)
  rng = np.random.default_rng(1)
  n = 16
  q0 = sys.init_q + rng.uniform(-0.1, 0.1, (n, sys.q_size()))
  qd0 = 0.1 * rng.standard_normal((n, sys.qd_size()))
  acts = rng.uniform(-1, 1, (1, n, sys.act_size()))
  cs = CStep(sys)
  q64, qd64, rows = cs.rollout(q0, qd0, acts, 5, dtype=np.float64, rows=True)
  q32, qd32 = cs.rollout(q0, qd0, acts, 5, dtype=np.float32)
  assert np.abs(q32 - q64).max() < 1e-3
  # the starting pose has the feet on the floor: contact rows are active
  assert 0 < rows < 4 * 17 + 8
