"""Launch-shape and LDS-tile autotuner for the step and adjoint kernels.

The kernels run one env per lane.  A batch of N envs fills N / epw
workgroups of epw lanes (envs per workgroup): a large batch wants full
64-lane waves, while a small one (the 4096-env differentiable rollout, the
few hundred envs of one chain length in a heterogeneous batch) reaches more
of the 256 compute units with fewer lanes per workgroup.  Where the best
point lies depends on the topology's registers and LDS per lane, so the
tuner measures: it times a short rollout (or adjoint) per candidate on the
model's own topology at the caller's batch size, keeps the fastest and
remembers it per (topology, device, batch size, substeps).  The results of
the kernels are bitwise the same for every launch shape; the choice is only
about speed.  ``tune_tile`` picks the step kernels' LDS tile, the fast
contact slots per env, among the compiled ones (results agree within
rounding: an env with more active contacts than slots takes the
slot-per-candidate path).  The reference has no counterpart (XLA picks its
launch shapes).
"""

from __future__ import annotations

from typing import Dict, Sequence, Tuple

import torch

from brax_amd import _lib
from brax_amd.generalized import pipeline

STEP = 0
ADJOINT = 1
TILE = 2  # the step kernels' LDS tile: fast contact slots per env
LANES = 3  # the step kernels' lanes per env (1, or the multi-lane width)
ADJOINT_PARTS = 4  # env parts of the adjoint's backward pass, a stream each (0: one)
CANDIDATES: Tuple[int, ...] = (64, 32, 16, 8, 4)

_CACHE: Dict[tuple, int] = {}


def set_launch(m: pipeline.DeviceModel, kernels: int, envs_per_workgroup: int) -> None:
  """Sets the envs per workgroup (a power of two up to 64; 0 = 64) of a
  model's step (STEP) or adjoint (ADJOINT) kernels."""
  _lib.check(_lib.load().bgx_model_set_launch(m.handle, int(kernels), int(envs_per_workgroup)))


def get_launch(m: pipeline.DeviceModel, kernels: int) -> int:
  r = _lib.load().bgx_model_get_launch(m.handle, int(kernels))
  if r < 0:
    raise _lib.BgxError("bad launch query")
  return r


def _time_ms(fn, reps: int) -> float:
  fn()  # warm-up: code objects loaded, caches filled
  best = float("inf")
  for _ in range(reps):
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record()
    fn()
    e1.record()
    e1.synchronize()
    best = min(best, e0.elapsed_time(e1))
  return best


def _key(m, kernels, n, n_frames):
  return (m.topo_id, str(m.device), kernels, n, n_frames)


def tiles(sys, device=None) -> Tuple[int, ...]:
  """The fast-contact-slot counts (LDS tiles) the step kernels of sys's
  topology are compiled for, default first."""
  m = pipeline.device_model(sys, device)
  buf = (_lib.c_i32 * 3)()
  k = _lib.load().bgx_topology_tiles(m.topo_id, buf)
  return tuple(int(buf[i]) for i in range(max(k, 0)))


def _rollout_runner(m, q, qd, acts, n_frames):
  """A callable running bgx_rollout over acts from copies of (q, qd)."""
  lib = _lib.load()
  n = q.shape[0]
  q0, qd0 = q.detach().float().contiguous(), qd.detach().float().contiguous()
  a = acts.detach().float().contiguous()
  qw, qdw = torch.empty_like(q0), torch.empty_like(qd0)
  T = a.shape[0]

  def run():
    qw.copy_(q0)
    qdw.copy_(qd0)
    _lib.check(lib.bgx_rollout(m.handle, n, T, int(n_frames), _lib.ptr(qw), _lib.ptr(qdw),
                               _lib.ptr(a) if m.nu else None, n * m.nu, None, None, None,
                               None, None, None, None, pipeline._stream()))
  return run


def tune_tile(sys, q: torch.Tensor, qd: torch.Tensor, acts: torch.Tensor, n_frames: int,
              reps: int = 3) -> int:
  """Picks the step kernels' LDS tile (fast contact slots per env) for a
  batch of q.shape[0] envs by timing bgx_rollout over acts with each
  compiled tile.  More slots hold more envs' contacts on the fast path but
  take more LDS per env (fewer workgroups per compute unit); which wins
  depends on the batch's contact counts, which change as the envs settle,
  so acts should span the horizon of interest (a two-step window from
  standing starts picks too few slots).  Returns the choice, which is set
  on the model."""
  m = pipeline.device_model(sys, q.device)
  n = q.shape[0]
  key = _key(m, TILE, n, n_frames)
  if key in _CACHE:
    set_launch(m, TILE, _CACHE[key])
    return _CACHE[key]
  cands = tiles(sys, q.device)
  if len(cands) == 1:
    set_launch(m, TILE, cands[0])
    _CACHE[key] = cands[0]
    return cands[0]
  run = _rollout_runner(m, q, qd, acts, n_frames)
  times = {}
  for c in cands:
    set_launch(m, TILE, c)
    times[c] = _time_ms(run, reps)
  best = min(times, key=times.get)
  set_launch(m, TILE, best)
  _CACHE[key] = best
  return best


def tune_step(sys, q: torch.Tensor, qd: torch.Tensor, acts: torch.Tensor, n_frames: int,
              candidates: Sequence[int] = CANDIDATES, reps: int = 3) -> int:
  """Picks the envs per workgroup of the step kernels for a batch of
  q.shape[0] envs by timing bgx_rollout over acts (T, N, act_size) from
  copies of (q, qd); the inputs are not modified.  Returns the choice, which
  is set on the model."""
  m = pipeline.device_model(sys, q.device)
  n = q.shape[0]
  key = _key(m, STEP, n, n_frames)
  if key in _CACHE:
    set_launch(m, STEP, _CACHE[key])
    return _CACHE[key]
  run = _rollout_runner(m, q, qd, acts, n_frames)
  times = {}
  for c in candidates:
    set_launch(m, STEP, c)
    times[c] = _time_ms(run, reps)
  best = min(times, key=times.get)
  set_launch(m, STEP, best)
  _CACHE[key] = best
  return best


def tune_adjoint(sys, q: torch.Tensor, qd: torch.Tensor, acts: torch.Tensor, n_frames: int,
                 candidates: Sequence[int] = CANDIDATES, reps: int = 2) -> int:
  """Picks the envs per workgroup of the adjoint kernels for a batch of
  q.shape[0] envs by timing the backward pass of a rollout over acts
  (T, N, act_size) with one checkpoint.  Returns the choice, which is set on
  the model."""
  m = pipeline.device_model(sys, q.device)
  n = q.shape[0]
  key = _key(m, ADJOINT, n, n_frames)
  if key in _CACHE:
    set_launch(m, ADJOINT, _CACHE[key])
    return _CACHE[key]
  q0 = q.detach().float().contiguous()
  qd0 = qd.detach().float().contiguous().requires_grad_(True)
  a = acts.detach().float().contiguous().requires_grad_(True)
  T = a.shape[0]
  qt, qdt, _ = pipeline.rollout_grad(sys, q0, qd0, a, n_frames, ckpt_every=T)
  gq = torch.ones_like(qt)
  gqd = torch.ones_like(qdt)

  def run():
    torch.autograd.grad((qt, qdt), (qd0, a), (gq, gqd), retain_graph=True)

  times = {}
  for c in candidates:
    set_launch(m, ADJOINT, c)
    times[c] = _time_ms(run, reps)
  best = min(times, key=times.get)
  set_launch(m, ADJOINT, best)
  _CACHE[key] = best
  return best
