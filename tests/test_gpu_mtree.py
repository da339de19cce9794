"""The multi-lane tree kernel (bgx_mtree.h: a lane per dof, the default
for the humanoid topologies) against the float64 oracle and against the
one-thread kernel (BGX_KERNELS_LANES = 1) on the same inputs."""

import os

import numpy as np
import pytest
import torch

from brax_amd import systems as S
from brax_amd.generalized import autotune, pipeline
from test_gpu_parity import _one_step_ok, _ratio

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
DEV = "cuda"


def _t(x):
  return torch.tensor(np.asarray(x), dtype=torch.float32, device=DEV)


def _lanes(sys, k):
  m = pipeline.device_model(sys, torch.device(DEV))
  autotune.set_launch(m, autotune.LANES, k)
  return autotune.get_launch(m, autotune.LANES)


@pytest.fixture
def humanoid():
  sys = S.humanoid(0)
  yield sys
  _lanes(sys, 0)


def test_humanoid_uses_the_multilane_kernel(humanoid):
  assert _lanes(humanoid, 0) == 32
  assert _lanes(humanoid, 1) == 1


def test_humanoid_one_env_step_both_kernels(humanoid):
  """One env step (5 substeps) of the golden humanoid inputs, each kernel
  against the float64 golden with the float32 oracle's exemptions."""
  sys = humanoid
  g = np.load(os.path.join(GOLDEN, "humanoid.npz"))
  out = {}
  for k in (1, 32):
    _lanes(sys, k)
    st = pipeline.init(sys, _t(g["q0"]), _t(g["qd0"]))
    st = pipeline.pipeline_step(sys, st, _t(g["acts"][0]), n_frames=5)
    torch.cuda.synchronize()
    print(f"\nlanes {k}:", end=" ")
    assert _one_step_ok(st.q.cpu(), st.qd.cpu(), g) <= 1.0
    out[k] = (st.q.cpu().numpy(), st.qd.cpu().numpy())
  r = np.maximum(_ratio(out[32][0], out[1][0]).max(1), _ratio(out[32][1], out[1][1]).max(1))
  print(f"multi-lane vs one-thread: band ratio median {np.median(r):.3g} max {r.max():.3g}")


def test_humanoid_first_substep_forces_vs_oracle(humanoid):
  """qf_smooth and qdd of the first substep of the golden humanoid inputs,
  each kernel against the float64 oracle (exact inverse).  qf_smooth: in
  the band on every env.  qdd comes out of the first contact solve, where
  float32 rounding flips solver decisions on most of these envs (the
  float32 oracle leaves the band on 10 of 16): qdd in the band on every env
  whose float32 result stays in it under twelve one-ulp input changes, and
  within K_EXEMPT of that float32 spread on the others."""
  import oracle.generalized as G
  from test_gpu_parity import K_EXEMPT
  sys = humanoid
  g = np.load(os.path.join(GOLDEN, "humanoid.npz"))
  q0, qd0, a0 = g["q0"], g["qd0"], g["acts"][0]
  o64 = G.Oracle(sys, torch.float64, inverse="exact")
  o32 = G.Oracle(sys, torch.float32, inverse="exact")
  r64 = o64.step(o64.init(q0.astype(np.float64), qd0.astype(np.float64)), a0.astype(np.float64))
  rng = np.random.default_rng(0)
  spread = np.zeros(q0.shape[0])
  for i in range(13):
    ulp = (lambda x: x) if i == 0 else \
        (lambda x: x * (1 + rng.choice([-1.0, 1.0], x.shape) * 2.0 ** -23))
    r32 = o32.step(o32.init(ulp(q0).astype(np.float32), ulp(qd0).astype(np.float32)),
                   a0.astype(np.float32))
    spread = np.maximum(spread, _ratio(r32.qdd.double(), r64.qdd).max(1))
  for k in (1, 32):
    _lanes(sys, k)
    st = pipeline.init(sys, _t(q0), _t(qd0))
    pr = pipeline.probe(sys, st, _t(a0), 1)
    torch.cuda.synchronize()
    rq = _ratio(pr["qf_smooth"].cpu(), r64.qf_smooth).max(1)
    rd = _ratio(pr["qdd"].cpu(), r64.qdd).max(1)
    stable = spread <= 1
    print(f"\nlanes {k}: qf_smooth max {rq.max():.3g}; qdd max {rd.max():.3g}; float32 "
          f"spread max {spread.max():.3g}; {int(stable.sum())} envs stable in float32, gpu "
          f"max there {rd[stable].max() if stable.any() else 0:.3g}", flush=True)
    assert rq.max() <= 1.0
    assert (not stable.any()) or rd[stable].max() <= 1.0
    assert np.all(rd <= K_EXEMPT * np.maximum(spread, 1.0))


def test_ragged_batch_and_rollout(humanoid):
  """An env's results do not depend on the batch around it (37 envs: the
  last wavefront half empty) and a rollout equals repeated steps."""
  sys = humanoid
  _lanes(sys, 32)
  rng = np.random.default_rng(5)
  n = 64
  q0 = sys.init_q + rng.uniform(-0.05, 0.05, (n, sys.q_size()))
  q0[:, 2] -= 0.05
  qd0 = 0.1 * rng.standard_normal((n, sys.qd_size()))
  acts = _t(rng.uniform(-0.4, 0.4, (3, n, sys.act_size())))
  st = pipeline.init(sys, _t(q0), _t(qd0))
  full = pipeline.rollout(sys, st, acts, 5)[0]
  part = pipeline.rollout(sys, pipeline.init(sys, _t(q0[:37]), _t(qd0[:37])), acts[:, :37], 5)[0]
  torch.cuda.synchronize()
  assert torch.equal(full.q[:37], part.q) and torch.equal(full.qd[:37], part.qd)
  stp = st
  for k in range(3):
    stp = pipeline.pipeline_step(sys, stp, acts[k], n_frames=5)
  assert torch.equal(stp.q, full.q) and torch.equal(stp.qd, full.qd)
  assert torch.isfinite(full.q).all()
