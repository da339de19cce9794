"""Training wrappers (brax/envs/wrappers/training.py:59-158), batched.

The env axis is present from reset, so the reference's VmapWrapper is the
identity here."""

from __future__ import annotations

import dataclasses

import torch

from brax_amd.envs.base import Env, State


class Wrapper(Env):

  def __init__(self, env: Env):
    self.env = env

  def reset(self, rng: torch.Generator, batch_size: int) -> State:
    return self.env.reset(rng, batch_size)

  def step(self, state: State, action: torch.Tensor) -> State:
    return self.env.step(state, action)

  @property
  def observation_size(self) -> int:
    return self.env.observation_size

  @property
  def action_size(self) -> int:
    return self.env.action_size

  @property
  def unwrapped(self) -> Env:
    return self.env.unwrapped

  def __getattr__(self, name):
    if name == "env":
      raise AttributeError(name)
    return getattr(self.env, name)


class EpisodeWrapper(Wrapper):
  """Steps count and done at episode end (training.py:74-125): each step
  repeats the action action_repeat times and sums the rewards."""

  def __init__(self, env: Env, episode_length: int, action_repeat: int):
    super().__init__(env)
    self.episode_length = int(episode_length)
    self.action_repeat = int(action_repeat)

  def reset(self, rng: torch.Generator, batch_size: int) -> State:
    state = self.env.reset(rng, batch_size)
    z = torch.zeros_like(state.done)
    info = dict(state.info)
    info["steps"] = z.clone()
    info["truncation"] = z.clone()
    info["episode_done"] = z.clone()
    em = {"sum_reward": z.clone(), "length": z.clone()}
    for k in state.metrics:
      em[k] = z.clone()
    info["episode_metrics"] = em
    return state.replace(info=info)

  def step(self, state: State, action: torch.Tensor) -> State:
    rewards = torch.zeros_like(state.reward)
    for _ in range(self.action_repeat):
      state = self.env.step(state, action)
      rewards = rewards + state.reward
    state = state.replace(reward=rewards)
    info = dict(state.info)
    steps = info["steps"] + self.action_repeat
    ended = steps >= self.episode_length
    done = torch.where(ended, torch.ones_like(state.done), state.done)
    info["truncation"] = torch.where(ended, 1 - state.done, torch.zeros_like(state.done))
    info["steps"] = steps
    prev_done = info["episode_done"]
    em = dict(info["episode_metrics"])
    em["sum_reward"] = em["sum_reward"] * (1 - prev_done) + rewards
    em["length"] = em["length"] * (1 - prev_done) + self.action_repeat
    for k, v in state.metrics.items():
      if k != "reward":
        em[k] = em[k] * (1 - prev_done) + v
    info["episode_metrics"] = em
    info["episode_done"] = done
    return state.replace(done=done, info=info)


class AutoResetWrapper(Wrapper):
  """Envs that are done restart from their first state (training.py:
  128-158)."""

  def reset(self, rng: torch.Generator, batch_size: int) -> State:
    state = self.env.reset(rng, batch_size)
    info = dict(state.info)
    info["first_pipeline_state"] = state.pipeline_state
    info["first_obs"] = state.obs
    return state.replace(info=info)

  def step(self, state: State, action: torch.Tensor) -> State:
    info = dict(state.info)
    if "steps" in info:
      info["steps"] = torch.where(state.done > 0, torch.zeros_like(info["steps"]),
                                  info["steps"])
    state = state.replace(done=torch.zeros_like(state.done), info=info)
    state = self.env.step(state, action)
    done = state.done > 0

    def where_done(x, y):
      if not torch.is_tensor(x) or x.dim() == 0 or x.shape[0] != done.shape[0]:
        return y
      return torch.where(done.reshape((-1,) + (1,) * (x.dim() - 1)), x, y)

    first, ps = state.info["first_pipeline_state"], state.pipeline_state
    fields = {}
    for f in dataclasses.fields(ps):
      a, b = getattr(first, f.name), getattr(ps, f.name)
      if dataclasses.is_dataclass(b):
        fields[f.name] = type(b)(**{g.name: where_done(getattr(a, g.name), getattr(b, g.name))
                                    for g in dataclasses.fields(b)})
      else:
        fields[f.name] = where_done(a, b)
    obs = where_done(state.info["first_obs"], state.obs)
    return state.replace(pipeline_state=type(ps)(**fields), obs=obs)
