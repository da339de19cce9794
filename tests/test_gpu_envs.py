"""The Ant env on the GPU pipeline: observation layout, reward terms and
termination as brax/envs/ant.py:205-279 computes them from the pipeline
states, and the wrappers around it."""

import numpy as np
import pytest
import torch

from brax_amd.envs import Ant, AutoResetWrapper, EpisodeWrapper

pytestmark = pytest.mark.gpu


def test_ant_reset_step_and_reward_terms():
  env = Ant()
  rng = torch.Generator().manual_seed(0)
  st = env.reset(rng, 64)
  assert st.obs.shape == (64, env.sys.q_size() - 2 + env.sys.qd_size())
  act = torch.rand(64, env.action_size, generator=rng).cuda() * 2 - 1
  ps0 = st.pipeline_state
  st1 = env.step(st, act)
  torch.cuda.synchronize()
  ps1 = st1.pipeline_state
  vel = (ps1.x.pos[:, 0] - ps0.x.pos[:, 0]) / env.dt
  expect = vel[:, 0] + 1.0 - 0.5 * (act * act).sum(-1)
  np.testing.assert_allclose(st1.reward.cpu().numpy(), expect.cpu().numpy(), rtol=1e-6,
                             atol=1e-6)
  z = ps1.x.pos[:, 0, 2]
  np.testing.assert_array_equal(st1.done.cpu().numpy(),
                                ((z < 0.2) | (z > 1.0)).float().cpu().numpy())
  np.testing.assert_array_equal(st1.obs.cpu().numpy(),
                                torch.cat([ps1.q[:, 2:], ps1.qd], -1).cpu().numpy())


def test_ant_wrapped_episodes_restart():
  env = AutoResetWrapper(EpisodeWrapper(Ant(), episode_length=4, action_repeat=1))
  rng = torch.Generator().manual_seed(1)
  st = env.reset(rng, 32)
  first_q = st.pipeline_state.q.clone()
  act = torch.zeros(32, env.action_size, device="cuda")
  for t in range(4):
    st = env.step(st, act)
  torch.cuda.synchronize()
  # every env ends its episode at step 4 and is back at its first state
  assert bool((st.done > 0).all())
  np.testing.assert_array_equal(st.pipeline_state.q.cpu().numpy(), first_q.cpu().numpy())
  assert torch.isfinite(st.info["episode_metrics"]["sum_reward"]).all()
