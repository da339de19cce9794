"""Analytic policy gradients through the adjoint: a few updates of a small
policy on the Ant env (wrapped for episodes and resets) give finite,
nonzero gradients and move the parameters."""

import math

import pytest
import torch

from brax_amd.envs import Ant, AutoResetWrapper, EpisodeWrapper
from brax_amd.training import apg

pytestmark = pytest.mark.gpu


def test_apg_updates_the_policy():
  env = AutoResetWrapper(EpisodeWrapper(Ant(), episode_length=100, action_repeat=1))
  cfg = apg.Config(num_envs=64, horizon_length=8, num_updates=3, learning_rate=1e-3)
  policy = apg.Policy(env.observation_size, env.action_size).cuda()
  before = [p.detach().clone() for p in policy.parameters()]
  policy, hist = apg.train(env, cfg, policy)
  torch.cuda.synchronize()
  assert len(hist) == 3
  for m in hist:
    assert math.isfinite(m["loss"]) and math.isfinite(m["grad_norm"])
    assert m["grad_norm"] > 0
  moved = sum(float((p.detach() - b).abs().max()) for p, b in zip(policy.parameters(), before))
  assert moved > 0
