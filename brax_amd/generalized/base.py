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
class Contact:
  """Contacts of a state (brax/base.py Contact; the generalized pipeline's
  debug output): every candidate of the topology, active (dist < 0) or not.
  pos (N, C, 3), frame (N, C, 3, 3) rows normal, t1, t2, dist (N, C);
  friction (C,) the tangential coefficient, solref (C, 2), solimp (C, 5),
  elasticity (C,) the mean of the two geoms'; link_idx = (link of geom 1,
  link of geom 2), each (C,) (-1 = world), the frame's normal pointing from
  geom 1 to geom 2."""
  pos: torch.Tensor
  frame: torch.Tensor
  dist: torch.Tensor
  friction: torch.Tensor
  solref: torch.Tensor
  solimp: torch.Tensor
  link_idx: tuple
  elasticity: Optional[torch.Tensor] = None


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
