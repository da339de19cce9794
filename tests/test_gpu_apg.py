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


def test_apg_policy_gradient_matches_oracle_autograd():
  """The trainer's loss (-mean reward over a horizon, apg.horizon_loss) for a
  fixed small policy on 8 Ant envs over 3 env steps: its gradient with
  respect to every policy parameter through the GPU env (the adjoint of each
  env step and of the frames the reward reads) against float64 autograd
  through the same env driven by the oracle (the reference's jax.grad of its
  loss, brax/training/agents/apg/train.py:188-198).  Relative error of the
  whole gradient below 1e-3, and of each layer's below 1e-2."""
  import copy

  import numpy as np

  from oracle_envs import OracleAnt

  env, ref = Ant(), OracleAnt()
  torch.manual_seed(3)
  policy = apg.Policy(env.observation_size, env.action_size, hidden=(16, 16),
                      normalize=False).cuda()
  policy64 = copy.deepcopy(policy).double().cpu()
  st = env.reset(torch.Generator().manual_seed(4), 8)
  so = ref.reset(torch.Generator().manual_seed(4), 8)
  loss, _, _ = apg.horizon_loss(env, policy, st, 3)
  loss.backward()
  loss64, _, _ = apg.horizon_loss(ref, policy64, so, 3)
  loss64.backward()
  torch.cuda.synchronize()
  g = [p.grad.detach().double().cpu().numpy().ravel() for p in policy.parameters()]
  r = [p.grad.detach().numpy().ravel() for p in policy64.parameters()]
  per_layer = [float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-12)) for a, b in zip(g, r)]
  gg, rr = np.concatenate(g), np.concatenate(r)
  rel = float(np.linalg.norm(gg - rr) / np.linalg.norm(rr))
  print(f"\nAPG loss gpu {float(loss):.6f} float64 oracle {float(loss64):.6f}; policy gradient "
        f"rel err {rel:.2e}, per tensor {np.array2string(np.array(per_layer), precision=2)}",
        flush=True)
  assert abs(float(loss) - float(loss64)) <= 1e-4 * abs(float(loss64)) + 1e-5
  assert rel < 1e-3
  assert max(per_layer) < 1e-2
