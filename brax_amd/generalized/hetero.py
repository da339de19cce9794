"""Environments of different structure stepped together, in one launch.

The reference vmaps one System over a batch (brax/envs/wrappers/training.py:
59-72) and per-env System arrays with DomainRandomizationVmapWrapper
(training.py:223-260); envs of different structure need one compiled program
each.  ``HeteroBatch`` puts envs of several models in one batch
(``bgx_batch_create``): hinge chains of any length up to 32 links (per-env
model tables) and envs of one tree topology that has a batch kernel (the
quadruped).  ``launch_rollout`` runs every env through the rollout in one
persistent kernel launch (``bgx_batch_rollout``): its workgroups take work
items (64 tree envs, or one wavefront of chains) from a device work table,
longest first, until the table is exhausted.

Envs are grouped by model: a chain length (one SystemBatch of those envs'
own tables) or a tree System's numbers (one shared table per distinct
model table).  The batch's q, qd and action rows are packed group after
group, env after env; ``pack`` / ``unpack`` convert per-env rows, and
``env_rows`` gives an env's slices.
"""

from __future__ import annotations

import ctypes
import hashlib
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np
import torch

from brax_amd import _lib
from brax_amd import codegen
from brax_amd import systems as S
from brax_amd.generalized import pipeline


class _Group:

  def __init__(self, envs: List[int], sys_obj, device):
    self.envs = envs
    self.sys = sys_obj
    self.model = pipeline.device_model(sys_obj, device)
    self.n = len(envs)


def _table_digest(sys: S.System, topo: codegen.Topology) -> str:
  return hashlib.sha1(codegen.pack_model(sys, topo).tobytes()).hexdigest()


class HeteroBatch:
  """One System per env, stepped by one persistent kernel launch."""

  def __init__(self, systems: Sequence[S.System], device=None):
    self.device = torch.device(device if device is not None else "cuda")
    if self.device.index is None:
      self.device = torch.device("cuda", torch.cuda.current_device())
    compiled = _lib.topologies()
    keyed: Dict[str, List[int]] = {}
    digests: Dict[int, str] = {}
    for e, s in enumerate(systems):
      per_env = codegen.Topology(s, per_env=True)
      if per_env.signature() in compiled and per_env.chain:
        key = "chain:" + per_env.signature()
      else:
        shared = codegen.Topology(s)
        if shared.signature() not in compiled:
          raise NotImplementedError(f"topology of {s.name!r} is not compiled")
        # envs with equal model tables share one upload, whatever object
        # holds them
        if id(s) not in digests:
          digests[id(s)] = _table_digest(s, shared)
        key = f"tree:{shared.signature()}:{digests[id(s)]}"
      keyed.setdefault(key, []).append(e)
    self.groups: List[_Group] = []
    for key, envs in keyed.items():  # first appearance order
      obj = (pipeline.SystemBatch([systems[e] for e in envs]) if key.startswith("chain:")
             else systems[envs[0]])
      self.groups.append(_Group(envs, obj, self.device))
    self.n = len(systems)
    lib = _lib.load()
    models = (ctypes.c_void_p * max(1, len(self.groups)))(
        *[g.model.handle.value for g in self.groups])
    counts = (ctypes.c_int32 * max(1, len(self.groups)))(*[g.n for g in self.groups])
    handle = ctypes.c_void_p()
    with torch.cuda.device(self.device):
      _lib.check(lib.bgx_batch_create(len(self.groups), models, counts, ctypes.byref(handle)))
    self.handle = handle
    sizes = (ctypes.c_int64 * 6)()
    _lib.check(lib.bgx_batch_sizes(handle, sizes))
    self.q_floats, self.qd_floats, self.act_floats = int(sizes[1]), int(sizes[2]), int(sizes[3])
    self.n_items, self.grid = int(sizes[4]), int(sizes[5])
    # per-env offset table: env -> (group, row, q offset, qd offset, act offset)
    self.env_group = np.empty(self.n, np.int64)
    self.env_row = np.empty(self.n, np.int64)
    self._q_off = np.empty(self.n, np.int64)
    self._qd_off = np.empty(self.n, np.int64)
    self._act_off = np.empty(self.n, np.int64)
    self._width = np.empty((self.n, 3), np.int64)
    self.env_slot = np.empty(self.n, np.int64)  # position in the packed env order
    qo = qdo = ao = slot = 0
    for g, grp in enumerate(self.groups):
      m = grp.model
      for r, e in enumerate(grp.envs):
        self.env_group[e], self.env_row[e], self.env_slot[e] = g, r, slot + r
        self._q_off[e] = qo + r * m.nq
        self._qd_off[e] = qdo + r * m.nv
        self._act_off[e] = ao + r * m.nu
        self._width[e] = (m.nq, m.nv, m.nu)
      qo += grp.n * m.nq
      qdo += grp.n * m.nv
      ao += grp.n * m.nu
      slot += grp.n
    assert (qo, qdo, ao) == (self.q_floats, self.qd_floats, self.act_floats)

  def __del__(self):
    h = getattr(self, "handle", None)
    if h is not None and h.value:
      try:
        _lib.load().bgx_batch_destroy(h)
      except Exception:  # interpreter shutdown
        pass

  def env_rows(self, e: int) -> Tuple[slice, slice, slice]:
    """Slices of env e's q, qd and action row in the packed buffers."""
    nq, nv, nu = (int(x) for x in self._width[e])
    return (slice(int(self._q_off[e]), int(self._q_off[e]) + nq),
            slice(int(self._qd_off[e]), int(self._qd_off[e]) + nv),
            slice(int(self._act_off[e]), int(self._act_off[e]) + nu))

  def pack(self, per_env: Sequence, kind: str) -> torch.Tensor:
    """A packed float32 device buffer from per-env rows: kind "q", "qd" or
    "act" (for "act", rows may carry a leading step axis: (T, width))."""
    k = {"q": 0, "qd": 1, "act": 2}[kind]
    rows = [np.asarray(torch.as_tensor(r).cpu().numpy() if torch.is_tensor(r) else r,
                       np.float32) for r in per_env]
    lead = rows[0].shape[:-1] if rows else ()
    width = (self.q_floats, self.qd_floats, self.act_floats)[k]
    out = np.zeros(lead + (width,), np.float32)
    for e, r in enumerate(rows):
      out[..., self.env_rows(e)[k]] = r
    return torch.tensor(out, device=self.device)

  def unpack(self, packed: torch.Tensor, kind: str) -> List[torch.Tensor]:
    """Per-env rows of a packed buffer (views)."""
    k = {"q": 0, "qd": 1, "act": 2}[kind]
    return [packed[..., self.env_rows(e)[k]] for e in range(self.n)]

  def _check(self, t: Optional[torch.Tensor], width: int, name: str, lead: Tuple = ()):
    if t is None:
      return
    if (not torch.is_tensor(t) or t.dtype != torch.float32 or t.device != self.device or
        not t.is_contiguous() or tuple(t.shape) != tuple(lead) + (width,)):
      raise ValueError(f"{name} must be a contiguous float32 tensor of shape "
                       f"{tuple(lead) + (width,)} on {self.device}")

  def launch_rollout(self, q: torch.Tensor, qd: torch.Tensor, acts: Optional[torch.Tensor],
                     n_steps: int, n_frames: int, status: Optional[torch.Tensor] = None,
                     q_traj: Optional[torch.Tensor] = None,
                     qd_traj: Optional[torch.Tensor] = None):
    """Enqueues the batch's rollout on the current stream (one kernel launch,
    graph-capturable): q (q floats), qd (qd floats) packed, updated in place;
    acts (>= n_steps, act floats) packed per step; optional per-env status and
    (n_steps, ...) trajectories."""
    self._check(q, self.q_floats, "q")
    self._check(qd, self.qd_floats, "qd")
    if acts is not None:
      if acts.dim() != 2 or acts.shape[0] < n_steps:
        raise ValueError(f"acts must have shape (>= {n_steps}, {self.act_floats})")
      self._check(acts, self.act_floats, "acts", (acts.shape[0],))
    self._check(q_traj, self.q_floats, "q_traj", (n_steps,))
    self._check(qd_traj, self.qd_floats, "qd_traj", (n_steps,))
    if status is not None and (status.dtype != torch.int32 or status.shape != (self.n,) or
                               status.device != self.device):
      raise ValueError(f"status must be an int32 tensor of shape ({self.n},)")
    lib = _lib.load()
    _lib.check(lib.bgx_batch_rollout(
        self.handle, int(n_steps), int(n_frames), _lib.ptr(q), _lib.ptr(qd),
        _lib.ptr(acts) if (acts is not None and self.act_floats) else None,
        self.act_floats, _lib.ptr(q_traj), _lib.ptr(qd_traj), _lib.ptr(status),
        ctypes.c_void_p(torch.cuda.current_stream(self.device).cuda_stream)))

  def rollout(self, q, qd, acts, n_steps: int, n_frames: int, trajectory: bool = False):
    """Advances the packed states q, qd (in place) by n_steps env steps;
    returns the per-env status in packed order and, with trajectory=True,
    the packed (n_steps, ...) trajectories."""
    status = torch.zeros(self.n, dtype=torch.int32, device=self.device)
    qt = torch.empty(n_steps, self.q_floats, device=self.device) if trajectory else None
    qdt = torch.empty(n_steps, self.qd_floats, device=self.device) if trajectory else None
    self.launch_rollout(q, qd, acts, n_steps, n_frames, status, qt, qdt)
    return (status, qt, qdt) if trajectory else status
