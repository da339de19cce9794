"""bench.py's bookkeeping on the CPU: the measured-traffic lookup reads the
committed rocprofv3 summaries of the exact launch it reports, the op-count
model keeps the per-substep figures the roofline fractions were quoted
with, and the heterogeneous CPU baseline steps every env with its own
System (compared with the C++ oracle run per System)."""

import numpy as np
import pytest
import torch

import bench
from brax_amd import systems as S


def test_measured_traffic_reads_the_committed_profiles():
  assert bench.measured_traffic("k_rollout", "quadruped", 65536, 1000) == 12004260019
  assert bench.measured_traffic("k_rollout", "quadruped", 65536, 20) == 286779560
  assert bench.measured_traffic("k_mtree", "humanoid", 16384, 1000) == 2143824051
  assert bench.measured_traffic("k_rollout", "hetero", 32768, 500,
                                rel="profiles/hetero_32768.md") == 1765421468
  # no profile of that launch: no figure
  assert bench.measured_traffic("k_rollout", "quadruped", 65536, 7) is None
  assert bench.measured_traffic("k_rollout", "quadruped", 1024, 1000) is None
  assert bench.measured_traffic("k_mtree", "quadruped", 65536, 1000) is None


def test_flop_model_is_frozen():
  q = S.quadruped(0)
  args = (q.num_links(), q.qd_size(), q.num_contacts(), q.num_limit_rows(),
          len(q.geom_type), q.solver_iterations)
  assert args == (9, 14, 17, 8, 10, 4)
  # the figures the quadruped's roofline fractions are quoted with
  assert bench.flop_per_substep(*args, 8.42) == pytest.approx(43771.0, abs=60)
  assert bench.flop_per_substep(*args, 9.49) == pytest.approx(45327.4, abs=60)


class _Rows:
  """The row bookkeeping of a HeteroBatch (env e's q, qd and action slices
  in the packed buffers), without a device."""

  def __init__(self, envs):
    self.off = []
    o = np.zeros(3, int)
    for s in envs:
      w = np.array([s.q_size(), s.qd_size(), s.act_size()])
      self.off.append((o.copy(), w))
      o += w
    self.sizes = o

  def env_rows(self, e):
    o, w = self.off[e]
    return tuple(slice(int(o[k]), int(o[k] + w[k])) for k in range(3))


def test_hetero_cpu_baseline_steps_each_env_with_its_own_system():
  rng = np.random.default_rng(3)
  quad = S.quadruped(0)
  chains = [S.chain(L, 100 + L) for L in (2, 5, 9)]
  envs = [quad] * 6 + [chains[i % 3] for i in range(9)]
  envs = [envs[i] for i in rng.permutation(len(envs))]
  hb = _Rows(envs)
  q0 = np.zeros(hb.sizes[0], np.float32)
  qd0 = np.zeros(hb.sizes[1], np.float32)
  for e, s in enumerate(envs):
    r = hb.env_rows(e)
    q0[r[0]] = s.init_q + rng.uniform(-0.1, 0.1, s.q_size())
    qd0[r[1]] = 0.1 * rng.standard_normal(s.qd_size())
  K = 3
  acts = rng.uniform(-1, 1, (K + 2, hb.sizes[2])).astype(np.float32)
  cpu, rows = bench.hetero_cpu_baseline(envs, hb, torch.tensor(q0), torch.tensor(qd0),
                                        torch.tensor(acts), K, chunk=4)
  assert cpu["kind"] == "port" and cpu["value"] > 0 and cpu["cores"] >= 1
  assert f"first {len(envs)} envs (6 quadrupeds, 9 chains) x {K} env steps" in cpu["sample"]
  # the quadrupeds' mean active rows: the oracle on those envs alone
  from oracle.cstep import CStep
  qi = [e for e, s in enumerate(envs) if s is quad]
  q = np.stack([q0[hb.env_rows(e)[0]] for e in qi])
  qd = np.stack([qd0[hb.env_rows(e)[1]] for e in qi])
  a = np.stack([acts[:K, hb.env_rows(e)[2]] for e in qi], 1)
  _, _, ref_rows = CStep(quad).rollout(q, qd, a, bench.N_FRAMES, threads=1, rows=True)
  assert rows == pytest.approx(ref_rows, rel=1e-12)
