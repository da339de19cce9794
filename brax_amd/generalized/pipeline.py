"""Generalized-coordinates physics pipeline on MI355X.

Mirrors brax.generalized.pipeline (brax/generalized/pipeline.py:32-94): the
same ``init(sys, q, qd, ...)`` / ``step(sys, state, act, ...)`` signatures and
State layout, batched over a leading env axis.  Every call runs hand-written
HIP kernels from libbrax_gen.so through its C ABI (include/brax_gen.h); there
is no host or PyTorch fallback, and a missing library raises.

Also provided, for the callers of the reference's step:
  * ``pipeline_step``: n_frames substeps with one action
    (PipelineEnv.pipeline_step, brax/envs/base.py:128-137);
  * ``rollout``: n_steps env steps in one launch, optionally returning the
    (q, qd) trajectory.
"""

from __future__ import annotations

import collections
import ctypes
from typing import Dict, Optional, Tuple

import numpy as np
import torch

from brax_amd import _lib
from brax_amd import codegen
from brax_amd import systems as S
from brax_amd.generalized.base import Contact, Motion, State, Transform

STATUS_SLOW_PATH = 1
STATUS_NONFINITE = 2


class SystemBatch:
  """One System per env, all of one structure: the envs share a compiled
  topology and each has its own model table (the reference batches per-env
  System arrays with jax.vmap, brax/envs/wrappers/training.py:223-260)."""

  def __init__(self, systems):
    self.systems = list(systems)
    if not self.systems:
      raise ValueError("empty SystemBatch")
    sigs = {codegen.Topology(s, per_env=True).signature() for s in self.systems}
    if len(sigs) != 1:
      raise ValueError("the systems of a SystemBatch must share one structure")

  def __len__(self):
    return len(self.systems)


class DeviceModel:
  """A System (or a SystemBatch) uploaded to one GPU: topology id + device
  model table(s)."""

  def __init__(self, sys, device: torch.device):
    self.sys = sys
    self.device = device
    batch = isinstance(sys, SystemBatch)
    systems = sys.systems if batch else [sys]
    topo = codegen.Topology(systems[0], per_env=batch)
    compiled = _lib.topologies()
    if topo.signature() not in compiled:
      hint = ""
      if not batch and codegen.Topology(sys, per_env=True).signature() in compiled:
        hint = " (it is compiled with per-env tables: pass a SystemBatch)"
      raise NotImplementedError(
          f"the topology of system {systems[0].name!r} is not compiled into "
          "libbrax_gen.so; add it to brax_amd.codegen.registered_systems() "
          f"and rebuild{hint}")
    self.topo_id, _, dims = compiled[topo.signature()]
    self.L, self.nq, self.nv, self.nu = dims[0], dims[1], dims[2], dims[3]
    self.n_tables = len(systems)
    packed = {}  # a System repeated over envs is packed once
    for s in systems:
      if id(s) not in packed:
        packed[id(s)] = codegen.pack_model(s, topo)
    table = np.ascontiguousarray(np.concatenate([packed[id(s)] for s in systems]))
    handle = ctypes.c_void_p()
    lib = _lib.load()
    with torch.cuda.device(device):
      _lib.check(lib.bgx_model_create(self.topo_id, table.ctypes.data,
                                      table.size, ctypes.byref(handle)))
    self.handle = handle

  def __del__(self):
    h = getattr(self, "handle", None)
    if h is not None and h.value:
      try:
        _lib.load().bgx_model_destroy(h)
      except Exception:  # interpreter shutdown
        pass


# uploaded models, most recently used last; at most _MODELS_MAX are kept
# (a dropped model's device table is freed when its last user lets go)
_MODELS: "collections.OrderedDict[Tuple[int, str], Tuple[S.System, DeviceModel]]" = (
    collections.OrderedDict())
_MODELS_MAX = 256


def device_model(sys, device=None) -> DeviceModel:
  """The uploaded model of a System or SystemBatch (cached per object and
  device, least recently used first out)."""
  device = torch.device(device if device is not None else "cuda")
  if device.type != "cuda":
    raise ValueError("the generalized pipeline runs on the GPU only")
  if device.index is None:
    device = torch.device("cuda", torch.cuda.current_device())
  key = (id(sys), str(device))
  hit = _MODELS.get(key)
  if hit is None or hit[0] is not sys:
    for s in (sys.systems if isinstance(sys, SystemBatch) else [sys]):
      S.from_arrays_check(s)
    hit = (sys, DeviceModel(sys, device))
    _MODELS[key] = hit
    while len(_MODELS) > _MODELS_MAX:
      _MODELS.popitem(last=False)
  _MODELS.move_to_end(key)
  return hit[1]


def _as_batch(x, n_cols: int, device, name: str) -> torch.Tensor:
  t = torch.as_tensor(x, dtype=torch.float32, device=device)
  if t.dim() == 1:
    t = t.unsqueeze(0)
  if t.dim() != 2 or t.shape[1] != n_cols:
    raise ValueError(f"{name} must have shape (N, {n_cols}), got {tuple(t.shape)}")
  return t.contiguous()


def _stream():
  return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _frames(m: DeviceModel, n: int, device):
  return (torch.empty(n, m.L, 3, device=device),
          torch.empty(n, m.L, 4, device=device),
          torch.empty(n, m.L, 3, device=device),
          torch.empty(n, m.L, 3, device=device))


def contacts(sys: S.System, q) -> Optional[Contact]:
  """The contacts of states q (N, q_size): what the reference's pipeline
  attaches to its State with debug=True (brax/generalized/pipeline.py:58-59,
  91-92), from the GPU's collisions; None for a topology without contact
  candidates."""
  m = device_model(sys, q.device if torch.is_tensor(q) and q.is_cuda else None)
  qb = _as_batch(q, m.nq, m.device, "q").detach()
  n = qb.shape[0]
  topo = codegen.Topology(sys.systems[0] if isinstance(sys, SystemBatch) else sys,
                          per_env=isinstance(sys, SystemBatch))
  c = topo.ncand
  if c == 0:
    return None
  dist = torch.empty(n, c, device=m.device)
  pos = torch.empty(n, c, 3, device=m.device)
  frame = torch.empty(n, c, 3, 3, device=m.device)
  lib = _lib.load()
  _lib.check(lib.bgx_contacts(m.handle, n, _lib.ptr(qb), _lib.ptr(dist), _lib.ptr(pos),
                              _lib.ptr(frame), _stream()))
  srcs = [p for p, _ in topo.cand_src]
  t = lambda x: torch.tensor(np.asarray(x, np.float32), device=m.device)  # noqa: E731
  first = sys.systems[0] if isinstance(sys, SystemBatch) else sys
  la = [int(first.geom_bodyid[p.geom1]) - 1 for p in srcs]
  lb = [int(first.geom_bodyid[p.geom2]) - 1 for p in srcs]
  el = [0.5 * (float(first.elasticity[p.geom1]) + float(first.elasticity[p.geom2])) for p in srcs]
  return Contact(pos=pos, frame=frame, dist=dist,
                 friction=t([p.friction[0] for p in srcs]),
                 solref=t([p.solref for p in srcs]), solimp=t([p.solimp for p in srcs]),
                 link_idx=(torch.tensor(la, device=m.device), torch.tensor(lb, device=m.device)),
                 elasticity=t(el))


def _with_contacts(sys, state: State, debug: bool) -> State:
  return state.replace(contact=contacts(sys, state.q)) if debug else state


class _FramesFn(torch.autograd.Function):
  """Frames x, xd of states (q, qd) (bgx_init), differentiable: the backward
  pass is bgx_frames_vjp (the reference's State.x, State.xd carry jax.grad
  gradients back to q and qd)."""

  @staticmethod
  def forward(ctx, q, qd, m):
    qc, qdc = q.detach().contiguous(), qd.detach().contiguous()
    n = qc.shape[0]
    xp, xr, xa, xv = _frames(m, n, m.device)
    lib = _lib.load()
    _lib.check(lib.bgx_init(m.handle, n, _lib.ptr(qc), _lib.ptr(qdc), _lib.ptr(xp),
                            _lib.ptr(xr), _lib.ptr(xa), _lib.ptr(xv), _stream()))
    ctx.save_for_backward(qc, qdc)
    ctx.m = m
    return xp, xr, xa, xv

  @staticmethod
  def backward(ctx, gxp, gxr, gxa, gxv):
    qc, qdc = ctx.saved_tensors
    m = ctx.m
    n = qc.shape[0]
    gq = torch.empty(n, m.nq, device=m.device)
    gqd = torch.empty(n, m.nv, device=m.device)
    c = lambda g: None if g is None else g.contiguous().float()  # noqa: E731
    gxp, gxr, gxa, gxv = c(gxp), c(gxr), c(gxa), c(gxv)
    lib = _lib.load()
    _lib.check(lib.bgx_frames_vjp(m.handle, n, _lib.ptr(qc), _lib.ptr(qdc), _lib.ptr(gxp),
                                  _lib.ptr(gxr), _lib.ptr(gxa), _lib.ptr(gxv), _lib.ptr(gq),
                                  _lib.ptr(gqd), _stream()))
    return gq, gqd, None


def frames(sys, q, qd):
  """(x, xd) of states q (N, q_size), qd (N, qd_size): kinematics.forward
  (brax/kinematics.py:31-108) as pipeline.init computes them; differentiable
  with respect to q and qd."""
  m = device_model(sys, q.device)
  xp, xr, xa, xv = _FramesFn.apply(q, qd, m)
  return Transform(xp, xr), Motion(xa, xv)


def init(sys: S.System, q, qd, unused_act=None, unused_ctrl=None,
         debug: bool = False, device=None) -> State:
  """pipeline.init (brax/generalized/pipeline.py:32-61).  When q or qd
  require gradients, the frames x, xd carry them (bgx_frames_vjp)."""
  del unused_act, unused_ctrl
  if device is None:
    device = q.device if torch.is_tensor(q) and q.is_cuda else "cuda"
  m = device_model(sys, device)
  if _needs_grad(q, qd):
    qb = _as_batch(q, m.nq, m.device, "q")
    qdb = _as_batch(qd, m.nv, m.device, "qd")
    x, xd = frames(sys, qb, qdb)
    return _with_contacts(sys, State(q=qb, qd=qdb, x=x, xd=xd), debug)
  qb = _as_batch(q, m.nq, m.device, "q").clone()
  qdb = _as_batch(qd, m.nv, m.device, "qd").clone()
  n = qb.shape[0]
  xp, xr, xa, xv = _frames(m, n, m.device)
  lib = _lib.load()
  _lib.check(lib.bgx_init(m.handle, n, _lib.ptr(qb), _lib.ptr(qdb), _lib.ptr(xp),
                          _lib.ptr(xr), _lib.ptr(xa), _lib.ptr(xv), _stream()))
  return _with_contacts(sys, State(q=qb, qd=qdb, x=Transform(xp, xr), xd=Motion(xa, xv)),
                        debug)


def _act(m: DeviceModel, act, n: int):
  if act is None or m.nu == 0:
    return None
  a = _as_batch(act, m.nu, m.device, "act")
  if a.shape[0] == 1 and n > 1:
    a = a.expand(n, m.nu).contiguous()
  if a.shape[0] != n:
    raise ValueError(f"act has {a.shape[0]} rows for {n} envs")
  return a


def _needs_grad(*ts) -> bool:
  return torch.is_grad_enabled() and any(
      t is not None and torch.is_tensor(t) and t.requires_grad for t in ts)


def pipeline_step(sys: S.System, state: State, act, n_frames: int = 1,
                  debug: bool = False) -> State:
  """``n_frames`` substeps with one action (brax/envs/base.py:128-137).

  Differentiable like the reference's step under jax.grad: when q, qd or
  act require gradients, the new q and qd carry them through the
  hand-written adjoint (a one-step ``rollout_grad``), and the frames x, xd
  of the returned State carry them on to q and qd (``frames``).  With
  debug=True the State carries the contacts of the new state
  (``contacts``).
  """
  m = device_model(sys, state.q.device)
  if _needs_grad(state.q, state.qd, act):
    n = state.q.shape[0]
    if m.nu:
      a = torch.as_tensor(act, dtype=torch.float32, device=m.device)
      a = (a[None] if a.dim() == 1 else a).expand(n, m.nu)[None]
    else:
      a = torch.zeros(1, n, 0, device=m.device)
    qt, qdt, status = rollout_grad(sys, state.q, state.qd, a, n_frames, ckpt_every=1)
    q, qd = qt[0], qdt[0]
    x, xd = frames(sys, q, qd)
    return _with_contacts(sys, State(q=q, qd=qd, x=x, xd=xd, status=status), debug)
  q = state.q.contiguous().clone()
  qd = state.qd.contiguous().clone()
  n = q.shape[0]
  a = _act(m, act, n)
  xp, xr, xa, xv = _frames(m, n, m.device)
  status = torch.zeros(n, dtype=torch.int32, device=m.device)
  lib = _lib.load()
  _lib.check(lib.bgx_step(m.handle, n, int(n_frames), _lib.ptr(q), _lib.ptr(qd),
                          _lib.ptr(a), _lib.ptr(xp), _lib.ptr(xr), _lib.ptr(xa),
                          _lib.ptr(xv), _lib.ptr(status), _stream()))
  return _with_contacts(sys, State(q=q, qd=qd, x=Transform(xp, xr), xd=Motion(xa, xv),
                                   status=status), debug)


def step(sys: S.System, state: State, act, debug: bool = False) -> State:
  """pipeline.step (brax/generalized/pipeline.py:64-94): one substep."""
  return pipeline_step(sys, state, act, 1, debug)


def rollout(sys: S.System, state: State, acts, n_frames: int,
            trajectory: bool = False, frames: bool = True,
            stream=None) -> Tuple[State, Optional[Tuple[torch.Tensor, torch.Tensor]]]:
  """``acts.shape[0]`` env steps of ``n_frames`` substeps in one launch.

  acts: (T, N, act_size) float32 on the GPU.  Returns the final State and,
  with ``trajectory=True``, the (q, qd) after every env step as tensors of
  shape (T, N, q_size) and (T, N, qd_size).
  """
  m = device_model(sys, state.q.device)
  q = state.q.contiguous().clone()
  qd = state.qd.contiguous().clone()
  n = q.shape[0]
  acts = torch.as_tensor(acts, dtype=torch.float32, device=m.device)
  if acts.dim() != 3 or acts.shape[1] != n or acts.shape[2] != m.nu:
    raise ValueError(f"acts must have shape (T, {n}, {m.nu})")
  acts = acts.contiguous()
  T = acts.shape[0]
  qt = torch.empty(T, n, m.nq, device=m.device) if trajectory else None
  qdt = torch.empty(T, n, m.nv, device=m.device) if trajectory else None
  xp = xr = xa = xv = None
  if frames:
    xp, xr, xa, xv = _frames(m, n, m.device)
  status = torch.zeros(n, dtype=torch.int32, device=m.device)
  lib = _lib.load()
  s = ctypes.c_void_p(stream) if stream is not None else _stream()
  _lib.check(lib.bgx_rollout(m.handle, n, T, int(n_frames), _lib.ptr(q),
                             _lib.ptr(qd), _lib.ptr(acts) if m.nu else None,
                             n * m.nu, _lib.ptr(qt), _lib.ptr(qdt),
                             _lib.ptr(xp), _lib.ptr(xr), _lib.ptr(xa),
                             _lib.ptr(xv), _lib.ptr(status), s))
  st = State(q=q, qd=qd,
             x=Transform(xp, xr) if frames else None,
             xd=Motion(xa, xv) if frames else None, status=status)
  return st, ((qt, qdt) if trajectory else None)


def probe(sys: S.System, state: State, act, n_frames: int = 1) -> Dict[str, torch.Tensor]:
  """Diagnostics of the last substep: qf_smooth, qdd, active contacts,
  solver iterations, line-search trials (bgx_step_probe)."""
  m = device_model(sys, state.q.device)
  q = state.q.contiguous().clone()
  qd = state.qd.contiguous().clone()
  n = q.shape[0]
  a = _act(m, act, n)
  w = 2 * m.nv + 3
  out = torch.zeros(n, w, device=m.device)
  lib = _lib.load()
  _lib.check(lib.bgx_step_probe(m.handle, n, int(n_frames), _lib.ptr(q),
                                _lib.ptr(qd), _lib.ptr(a), _lib.ptr(out),
                                _stream()))
  nv = m.nv
  return {"q": q, "qd": qd, "qf_smooth": out[:, :nv], "qdd": out[:, nv:2 * nv],
          "n_contacts": out[:, 2 * nv], "iters": out[:, 2 * nv + 1],
          "trials": out[:, 2 * nv + 2]}


class _RolloutFn(torch.autograd.Function):
  """Differentiable rollout: forward = bgx_rollout with the trajectory,
  backward = bgx_rollout_adjoint (segments of ckpt_every env steps are
  recomputed from their checkpoints)."""

  @staticmethod
  def forward(ctx, q0, qd0, acts, m, n_frames, ckpt_every, min_workspace):
    n, T = q0.shape[0], acts.shape[0]
    q = q0.detach().contiguous().clone()
    qd = qd0.detach().contiguous().clone()
    a = acts.detach().contiguous()
    qt = torch.empty(T, n, m.nq, device=m.device)
    qdt = torch.empty(T, n, m.nv, device=m.device)
    status = torch.zeros(n, dtype=torch.int32, device=m.device)
    lib = _lib.load()
    _lib.check(lib.bgx_rollout(m.handle, n, T, int(n_frames), _lib.ptr(q), _lib.ptr(qd),
                               _lib.ptr(a) if m.nu else None, n * m.nu, _lib.ptr(qt),
                               _lib.ptr(qdt), None, None, None, None, _lib.ptr(status),
                               _stream()))
    idx = list(range(ckpt_every - 1, T - 1, ckpt_every))
    ckq = torch.cat([q0.detach()[None].contiguous(), qt[idx]], 0).contiguous()
    ckqd = torch.cat([qd0.detach()[None].contiguous(), qdt[idx]], 0).contiguous()
    ctx.save_for_backward(ckq, ckqd, a)
    ctx.m, ctx.n_frames, ctx.ckpt_every = m, int(n_frames), int(ckpt_every)
    ctx.min_workspace = bool(min_workspace)
    ctx.mark_non_differentiable(status)
    return qt, qdt, status

  @staticmethod
  def backward(ctx, gqt, gqdt, _gstatus):
    ckq, ckqd, a = ctx.saved_tensors
    m = ctx.m
    T, n = a.shape[0], ckq.shape[1]
    gqt = None if gqt is None else gqt.contiguous().float()
    gqdt = None if gqdt is None else gqdt.contiguous().float()
    lib = _lib.load()
    size = (lib.bgx_adjoint_workspace_min_bytes if ctx.min_workspace
            else lib.bgx_adjoint_workspace_bytes)
    nbytes = size(m.handle, n, ctx.ckpt_every, ctx.n_frames)
    if nbytes < 0:
      raise _lib.BgxError("bad adjoint workspace arguments")
    ws = torch.empty(nbytes // 4 + 1, device=m.device)
    gact = torch.empty(T, n, m.nu, device=m.device)
    gq0 = torch.empty(n, m.nq, device=m.device)
    gqd0 = torch.empty(n, m.nv, device=m.device)
    status = torch.zeros(n, dtype=torch.int32, device=m.device)
    _lib.check(lib.bgx_rollout_adjoint(
        m.handle, n, T, ctx.n_frames, ctx.ckpt_every, _lib.ptr(ckq), _lib.ptr(ckqd),
        _lib.ptr(a) if m.nu else None, n * m.nu, _lib.ptr(gqt), _lib.ptr(gqdt),
        _lib.ptr(gact) if m.nu else None, _lib.ptr(gq0), _lib.ptr(gqd0), _lib.ptr(ws),
        nbytes, _lib.ptr(status), _stream()))
    return gq0, gqd0, gact, None, None, None, None


def rollout_grad(sys: S.System, q0: torch.Tensor, qd0: torch.Tensor, acts: torch.Tensor,
                 n_frames: int, ckpt_every: int = 16, min_workspace: bool = False):
  """Differentiable rollout of ``acts.shape[0]`` env steps.

  Returns (q_traj (T, N, q_size), qd_traj (T, N, qd_size), status (N)); the
  trajectories carry gradients to q0, qd0 and acts through the hand-written
  adjoint (the reference obtains them with jax.grad through its scan of
  pipeline.step, brax/training/agents/apg/train.py:188-198).  The backward
  pass recomputes segments of ``ckpt_every`` env steps from checkpoints.
  With ``min_workspace`` the backward pass takes the least workspace (one
  substep's tape at a time) instead of a whole segment's tapes.
  """
  m = device_model(sys, q0.device)
  if q0.dim() != 2 or q0.shape[1] != m.nq or qd0.shape != (q0.shape[0], m.nv):
    raise ValueError("q0 must be (N, q_size) and qd0 (N, qd_size)")
  if acts.dim() != 3 or acts.shape[1:] != (q0.shape[0], m.nu):
    raise ValueError(f"acts must have shape (T, {q0.shape[0]}, {m.nu})")
  if ckpt_every < 1 or n_frames < 1:
    raise ValueError("ckpt_every and n_frames must be positive")
  return _RolloutFn.apply(q0.float(), qd0.float(), acts.float(), m, n_frames, ckpt_every,
                          min_workspace)
