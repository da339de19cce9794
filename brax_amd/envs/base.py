"""Env state and the pipeline env (brax/envs/base.py:26-169), batched.

The reference vmaps an unbatched env; here every field carries the env axis
N from the start (what brax.envs.wrappers.training.VmapWrapper produces),
and the physics runs through brax_amd.generalized.pipeline on the GPU.
"""

from __future__ import annotations

import dataclasses
from typing import Any, Dict, Optional

import torch

from brax_amd import systems as S
from brax_amd.generalized import base as gbase
from brax_amd.generalized import pipeline


@dataclasses.dataclass
class State:
  """Environment state (brax/envs/base.py:26-36), batched over envs."""

  pipeline_state: Optional[gbase.State]
  obs: torch.Tensor
  reward: torch.Tensor
  done: torch.Tensor
  metrics: Dict[str, torch.Tensor] = dataclasses.field(default_factory=dict)
  info: Dict[str, Any] = dataclasses.field(default_factory=dict)

  def replace(self, **kw) -> "State":
    return dataclasses.replace(self, **kw)


class Env:
  """Interface of an environment (brax/envs/base.py:39-75)."""

  def reset(self, rng: torch.Generator, batch_size: int) -> State:
    raise NotImplementedError

  def step(self, state: State, action: torch.Tensor) -> State:
    raise NotImplementedError

  @property
  def observation_size(self) -> int:
    raise NotImplementedError

  @property
  def action_size(self) -> int:
    raise NotImplementedError

  @property
  def unwrapped(self) -> "Env":
    return self


class PipelineEnv(Env):
  """An env over the generalized pipeline (brax/envs/base.py:78-169): one
  env step is n_frames substeps with one action, run as one kernel launch."""

  def __init__(self, sys: S.System, backend: str = "generalized", n_frames: int = 1,
               device: Optional[torch.device] = None, debug: bool = False):
    if backend != "generalized":
      raise ValueError(f"backend {backend!r} is not available: brax_amd runs the "
                       "generalized pipeline")
    self.sys = sys
    self.backend = backend
    self._n_frames = int(n_frames)
    self._debug = debug
    self.device = torch.device(device if device is not None else "cuda")

  def pipeline_init(self, q: torch.Tensor, qd: torch.Tensor) -> gbase.State:
    return pipeline.init(self.sys, q, qd, debug=self._debug)

  def pipeline_step(self, pipeline_state: gbase.State, action: torch.Tensor) -> gbase.State:
    return pipeline.pipeline_step(self.sys, pipeline_state, action, self._n_frames,
                                  self._debug)

  @property
  def dt(self) -> float:
    """Time per env step (brax/envs/base.py:140-143)."""
    return float(self.sys.opt.timestep) * self._n_frames

  @property
  def observation_size(self) -> int:
    rng = torch.Generator(device="cpu").manual_seed(0)
    return int(self.reset(rng, 1).obs.shape[-1])

  @property
  def action_size(self) -> int:
    return self.sys.act_size()
