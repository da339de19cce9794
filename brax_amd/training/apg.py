"""Analytic policy gradients through the hand-written adjoint
(brax/training/agents/apg/train.py:176-236), in torch.

The loss is minus the mean reward of a horizon of env steps taken by a
deterministic policy; its gradient reaches the policy parameters through
the env's differentiable pipeline steps (each a one-step adjoint of the
GPU step).  Gradients are clipped to a global norm, then elementwise to
[-1, 1], then applied with Adam (the reference's optax chain); the env
state carries over between updates without gradient, as the reference's
scan does from one loss call to the next.
"""

from __future__ import annotations

import dataclasses
from typing import Callable, Dict, List, Optional, Sequence

import torch

from brax_amd.envs.base import Env


class Policy(torch.nn.Module):
  """MLP with swish hidden layers and a tanh output in [-1, 1]; optional
  running observation normalisation (running_statistics in the
  reference)."""

  def __init__(self, obs_size: int, action_size: int,
               hidden: Sequence[int] = (32, 32, 32, 32), normalize: bool = True):
    super().__init__()
    sizes = [obs_size] + list(hidden) + [action_size]
    self.layers = torch.nn.ModuleList(
        torch.nn.Linear(a, b) for a, b in zip(sizes[:-1], sizes[1:]))
    self.normalize = normalize
    self.register_buffer("count", torch.zeros(()))
    self.register_buffer("mean", torch.zeros(obs_size))
    self.register_buffer("var", torch.ones(obs_size))

  @torch.no_grad()
  def update_normalizer(self, obs: torch.Tensor):
    obs = obs.reshape(-1, obs.shape[-1]).float()
    n = obs.shape[0]
    tot = self.count + n
    delta = obs.mean(0) - self.mean
    self.mean += delta * (n / tot)
    self.var = (self.var * self.count + obs.var(0, unbiased=False) * n +
                delta * delta * self.count * n / tot) / tot
    self.count = tot

  def forward(self, obs: torch.Tensor) -> torch.Tensor:
    x = obs
    if self.normalize and float(self.count) > 0:
      x = (x - self.mean) / torch.sqrt(self.var + 1e-6)
    for i, layer in enumerate(self.layers):
      x = layer(x)
      if i < len(self.layers) - 1:
        x = torch.nn.functional.silu(x)
    return torch.tanh(x)


@dataclasses.dataclass
class Config:
  num_envs: int = 256
  horizon_length: int = 32
  num_updates: int = 10
  learning_rate: float = 1e-4
  max_gradient_norm: float = 1e9
  adam_b: Sequence[float] = (0.7, 0.95)
  seed: int = 0


def horizon_loss(env: Env, policy: torch.nn.Module, state, horizon: int):
  """Minus the mean reward of `horizon` env steps of the policy from state
  (the loss of brax/training/agents/apg/train.py:176-198); returns (loss,
  final state, observations)."""
  rewards, observations = [], []
  st = state
  for _ in range(horizon):
    observations.append(st.obs.detach())
    st = env.step(st, policy(st.obs))
    rewards.append(st.reward)
  return -torch.stack(rewards).mean(), st, observations


def train(env: Env, cfg: Config = Config(), policy: Optional[Policy] = None,
          progress: Optional[Callable[[int, Dict[str, float]], None]] = None):
  """Returns (policy, per-update metrics)."""
  gen = torch.Generator().manual_seed(cfg.seed)
  torch.manual_seed(cfg.seed)
  state = env.reset(gen, cfg.num_envs)
  if policy is None:
    policy = Policy(state.obs.shape[-1], env.action_size)
  policy = policy.to(state.obs.device)
  opt = torch.optim.Adam(policy.parameters(), lr=cfg.learning_rate,
                         betas=tuple(cfg.adam_b))
  history: List[Dict[str, float]] = []
  for it in range(cfg.num_updates):
    opt.zero_grad(set_to_none=True)
    loss, st, observations = horizon_loss(env, policy, state, cfg.horizon_length)
    loss.backward()
    params = [p for p in policy.parameters() if p.grad is not None]
    g_norm = torch.sqrt(sum((p.grad * p.grad).sum() for p in params))
    if float(g_norm) >= cfg.max_gradient_norm:
      for p in params:
        p.grad *= cfg.max_gradient_norm / g_norm
    for p in params:
      p.grad.clamp_(-1.0, 1.0)
    opt.step()
    policy.update_normalizer(torch.stack(observations))
    state = _detach(st)
    m = {"loss": float(loss.detach()), "grad_norm": float(g_norm),
         "params_norm": float(torch.sqrt(sum((p * p).sum() for p in policy.parameters())))}
    history.append(m)
    if progress:
      progress(it, m)
  return policy, history


def _detach(x):
  if torch.is_tensor(x):
    return x.detach()
  if dataclasses.is_dataclass(x) and not isinstance(x, type):
    return type(x)(**{f.name: _detach(getattr(x, f.name)) for f in dataclasses.fields(x)})
  if isinstance(x, dict):
    return {k: _detach(v) for k, v in x.items()}
  return x
