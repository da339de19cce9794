"""Environments of different structure stepped together.

The reference vmaps one System over a batch (brax/envs/wrappers/training.py:
59-72); envs of different topologies need one compiled program per topology.
``HeteroBatch`` keeps a per-env offset table (env -> group, row) over groups
of envs that share a compiled topology.  A group of one shared System uses a
single model table; a group of per-env systems (``SystemBatch``, e.g. chains
with their own lengths and masses) uses one table per env.  ``rollout``
launches the groups' kernels on several streams so the groups run
concurrently; a capture of ``launch_rollout`` is one HIP graph, replayed as
one launch.

A group's kernel takes about as long as one of its waves (a lane per env,
the waves run side by side), so the batch takes as long as its busiest
stream.  There is one stream per hardware queue of the process (the HIP
runtime's GPU_MAX_HW_QUEUES, 4 by default; streams beyond the queues would
share them and run in turn); the groups go to the streams longest first,
each to the least loaded one.  More queues do not fit the long chains:
each queue reserves scratch for the largest private segment it runs (9.7 KB
per lane for the rolled 25-32-link kernels), and at 8 queues the runtime
fails to allocate them.
"""

from __future__ import annotations

import ctypes
import os
from typing import Dict, List, Sequence

import torch

from brax_amd import _lib
from brax_amd import codegen
from brax_amd import systems as S
from brax_amd.generalized import pipeline


# the step kernels unroll their dof loops up to this many dofs
# (bgx_step.h unroll_nv); longer trees run rolled loops through scratch
_UNROLL_MAX_DOFS = 24


def _group_cost(g) -> float:
  """Relative time of one wave of a group's kernel: dofs^3, four times that
  with contact or limit rows or with rolled dof loops (kernel traces: a
  14-dof quadruped 11 ms, a 24-link chain 16 ms, a 25-link chain 60 ms per
  20 env steps)."""
  s0 = g.sys.systems[0] if hasattr(g.sys, "systems") else g.sys
  nv = s0.qd_size()
  rows = s0.num_contacts() + s0.num_limit_rows()
  return float(nv ** 3) * (4.0 if rows > 0 or nv > _UNROLL_MAX_DOFS else 1.0)


def assign_streams(costs: Sequence[float], k: int) -> List[int]:
  """Stream of each job: longest first, each to the least loaded of k
  streams (ties to the lowest stream)."""
  load = [0.0] * k
  out = [0] * len(costs)
  for g in sorted(range(len(costs)), key=lambda g: -costs[g]):
    j = min(range(k), key=lambda j: load[j])
    out[g] = j
    load[j] += costs[g]
  return out


class _Group:

  def __init__(self, envs: List[int], sys_obj, device):
    self.envs = envs
    self.sys = sys_obj
    self.model = pipeline.device_model(sys_obj, device)
    self.n = len(envs)


class HeteroBatch:
  """One System per env; envs are grouped by compiled topology."""

  def __init__(self, systems: Sequence[S.System], device=None):
    self.device = torch.device(device if device is not None else "cuda")
    if self.device.index is None:
      self.device = torch.device("cuda", torch.cuda.current_device())
    compiled = _lib.topologies()
    keyed: Dict[str, List[int]] = {}
    for e, s in enumerate(systems):
      shared = codegen.Topology(s).signature()
      per_env = codegen.Topology(s, per_env=True).signature()
      if per_env in compiled:
        key = "pe:" + per_env
      elif shared in compiled:
        key = f"sh:{shared}:{id(s)}"
      else:
        raise NotImplementedError(f"topology of {s.name!r} is not compiled")
      keyed.setdefault(key, []).append(e)
    self.groups: List[_Group] = []
    # most expensive groups first (dofs^3 per env), so they start first
    order = sorted(keyed.items(), key=lambda kv: -systems[kv[1][0]].qd_size() ** 3 * len(kv[1]))
    for key, envs in order:
      obj = (pipeline.SystemBatch([systems[e] for e in envs]) if key.startswith("pe:")
             else systems[envs[0]])
      self.groups.append(_Group(envs, obj, self.device))
    self.n = len(systems)
    # per-env offset table: env -> (group, row)
    self.env_group = torch.empty(self.n, dtype=torch.int64)
    self.env_row = torch.empty(self.n, dtype=torch.int64)
    for g, grp in enumerate(self.groups):
      for r, e in enumerate(grp.envs):
        self.env_group[e], self.env_row[e] = g, r
    self._streams = None

  def split(self, per_env: Sequence[torch.Tensor]) -> List[torch.Tensor]:
    """Per-group (n_g, width_g) tensors from per-env rows."""
    return [torch.stack([torch.as_tensor(per_env[e], dtype=torch.float32)
                         for e in g.envs]).to(self.device).contiguous()
            for g in self.groups]

  def gather(self, per_group: Sequence[torch.Tensor], env: int) -> torch.Tensor:
    """Row of env from per-group tensors."""
    return per_group[int(self.env_group[env])][int(self.env_row[env])]

  def streams(self, k: int = 0):
    """The batch's streams (k of them; 0: one per hardware queue, at most one
    per group) and each group's stream."""
    if self._streams is None:
      if k <= 0:
        k = int(os.environ.get("GPU_MAX_HW_QUEUES", "") or 4)
      k = max(1, min(k, len(self.groups)))
      self._streams = [torch.cuda.Stream(device=self.device) for _ in range(k)]
      self._stream_of = assign_streams([_group_cost(g) for g in self.groups], k)
    return self._streams

  def launch_rollout(self, q: List[torch.Tensor], qd: List[torch.Tensor],
                     acts: List[torch.Tensor], n_steps: int, n_frames: int,
                     status: List[torch.Tensor] = None):
    """Enqueues every group's rollout (in place on q, qd) on its own stream,
    forked from and joined back into the current stream."""
    lib = _lib.load()
    cur = torch.cuda.current_stream(self.device)
    streams = self.streams()
    for s in streams:
      s.wait_stream(cur)
    for g, grp in enumerate(self.groups):
      s = streams[self._stream_of[g]]
      m = grp.model
      a = acts[g] if m.nu else None
      _lib.check(lib.bgx_rollout(
          m.handle, grp.n, n_steps, n_frames, _lib.ptr(q[g]), _lib.ptr(qd[g]),
          _lib.ptr(a), grp.n * m.nu, None, None, None, None, None, None,
          _lib.ptr(status[g]) if status is not None else None,
          ctypes.c_void_p(s.cuda_stream)))
    for s in streams:
      cur.wait_stream(s)

  def autotune(self, q, qd, acts, n_frames: int, n_steps: int = 2):
    """Tunes every group's envs per workgroup (autotune.tune_step) on a short
    rollout from its current state; q, qd are not modified.  Returns the
    choices, group by group."""
    from brax_amd.generalized import autotune
    return [autotune.tune_step(g.sys, q[i], qd[i], acts[i][:n_steps], n_frames)
            for i, g in enumerate(self.groups)]

  def rollout(self, q, qd, acts, n_steps: int, n_frames: int):
    """Advances per-group states q[g], qd[g] (in place) by n_steps env steps;
    acts[g]: (n_steps, n_g, nu_g)."""
    status = [torch.zeros(g.n, dtype=torch.int32, device=self.device) for g in self.groups]
    self.launch_rollout(q, qd, acts, n_steps, n_frames, status)
    return status
