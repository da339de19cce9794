"""Episode and auto-reset wrappers (brax/envs/wrappers/training.py:74-158)
on a stand-in env of CPU tensors: step counts, truncation, episode metrics
and restart from the first state follow the reference's rules."""

import torch

from brax_amd.envs import AutoResetWrapper, EpisodeWrapper
from brax_amd.envs.base import Env, State
from brax_amd.generalized.base import Motion, State as PState, Transform


class Counter(Env):
  """q counts up by the action; done once q reaches 3; reward = action."""

  def reset(self, rng, batch_size):
    n = batch_size
    q = torch.zeros(n, 1)
    ps = PState(q=q, qd=torch.zeros(n, 1), x=Transform(torch.zeros(n, 1, 3), torch.zeros(n, 1, 4)),
                xd=Motion(torch.zeros(n, 1, 3), torch.zeros(n, 1, 3)))
    z = torch.zeros(n)
    return State(ps, q.clone(), z.clone(), z.clone(), {"m": z.clone()}, {})

  def step(self, state, action):
    ps = state.pipeline_state
    q = ps.q + action
    nps = ps.replace(q=q)
    done = (q[:, 0] >= 3).float()
    return state.replace(pipeline_state=nps, obs=q.clone(), reward=action[:, 0],
                         done=done, metrics={"m": q[:, 0]})


def test_episode_wrapper_truncates_and_accumulates():
  env = EpisodeWrapper(Counter(), episode_length=4, action_repeat=2)
  st = env.reset(None, 2)
  act = torch.tensor([[0.5], [0.25]])
  st = env.step(st, act)  # two substeps: q = 1.0, 0.5
  assert torch.equal(st.info["steps"], torch.tensor([2.0, 2.0]))
  assert torch.equal(st.reward, torch.tensor([1.0, 0.5]))
  assert torch.equal(st.done, torch.tensor([0.0, 0.0]))
  st = env.step(st, act)  # q = 2.0, 1.0; 4 steps reached: truncated
  assert torch.equal(st.done, torch.tensor([1.0, 1.0]))
  assert torch.equal(st.info["truncation"], torch.tensor([1.0, 1.0]))
  assert torch.equal(st.info["episode_metrics"]["sum_reward"], torch.tensor([2.0, 1.0]))
  assert torch.equal(st.info["episode_metrics"]["length"], torch.tensor([4.0, 4.0]))


def test_auto_reset_restarts_done_envs_from_their_first_state():
  env = AutoResetWrapper(EpisodeWrapper(Counter(), episode_length=100, action_repeat=1))
  st = env.reset(None, 2)
  act = torch.tensor([[2.0], [1.0]])
  st = env.step(st, act)  # q = 2, 1
  st = env.step(st, act)  # env 0 reaches 4 -> done, restarts at q = 0; env 1 at 2
  assert torch.equal(st.done, torch.tensor([1.0, 0.0]))
  assert torch.equal(st.pipeline_state.q[:, 0], torch.tensor([0.0, 2.0]))
  assert torch.equal(st.obs[:, 0], torch.tensor([0.0, 2.0]))
  # a terminated (not truncated) episode
  assert torch.equal(st.info["truncation"], torch.tensor([0.0, 0.0]))
  st = env.step(st, act)  # the step counter of env 0 restarted
  assert torch.equal(st.info["steps"], torch.tensor([1.0, 3.0]))
