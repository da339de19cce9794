"""State of the generalized pipeline (brax/base.py:396-412).

Batched over a leading env axis N, as the reference's State is under
``jax.vmap``: q (N, q_size), qd (N, qd_size), x = Transform(pos (N, L, 3),
rot (N, L, 4)), xd = Motion(ang (N, L, 3), vel (N, L, 3)), contact.  All
tensors are float32 on the GPU.  The generalized extras of the reference's
State (brax/generalized/base.py:25-60: cinr, cdof, mass_mx, con_jac, ...) are
recomputed inside each substep from (q, qd) and are not carried.
"""

from __future__ import annotations

import dataclasses
from typing import Optional

import torch


@dataclasses.dataclass
class Transform:
  pos: torch.Tensor
  rot: torch.Tensor


@dataclasses.dataclass
class Motion:
  ang: torch.Tensor
  vel: torch.Tensor


@dataclasses.dataclass
class State:
  q: torch.Tensor
  qd: torch.Tensor
  x: Transform
  xd: Motion
  contact: Optional[object] = None
  status: Optional[torch.Tensor] = None  # per-env BGX_STATUS_* bits

  def replace(self, **kw) -> "State":
    return dataclasses.replace(self, **kw)
