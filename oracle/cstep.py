"""ctypes front end of oracle/cstep.cpp, the C++ restatement of the oracle.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py): the CPU baseline of
bench.py and the tests load it; the product never does.  The System is
flattened here from its fields, the way oracle/generalized.py reads them,
not from the product's model tables.
"""

from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "cstep.cpp")
LIB = os.path.join(HERE, "_build", "libcstep.so")

_P = ctypes.c_void_p
_I = ctypes.c_int


class _Sys(ctypes.Structure):
  _fields_ = [(n, _I) for n in ("L", "nq", "nv", "nu", "nc", "has_limit", "mat_inv_iters",
                                "solver_iters", "solver_maxls")] + [
      ("dt", ctypes.c_double), ("ls_eps", ctypes.c_double), ("gravity", ctypes.c_double * 3)] + [
      (n, _P) for n in ("ltype", "parent", "qadr", "dadr", "lt", "lj", "li", "inertia", "mass",
                        "l_invw", "motion", "dofp", "solp", "act_q", "act_d", "actp", "con")] + [
      ("capsule_frame", _I)]


def build(force: bool = False) -> str:
  """Compiles libcstep.so (g++ -O3, OpenMP, x86-64-v3 so that it runs on
  any AVX2 host)."""
  if force or not os.path.exists(LIB) or os.path.getmtime(LIB) < os.path.getmtime(SRC):
    os.makedirs(os.path.dirname(LIB), exist_ok=True)
    tmp = LIB + ".tmp"
    subprocess.run(["g++", "-O3", "-march=x86-64-v3", "-std=c++17", "-fopenmp", "-shared",
                    "-fPIC", SRC, "-o", tmp], check=True)
    os.replace(tmp, LIB)
  return LIB


_lib = None


def load():
  global _lib
  if _lib is None:
    if not os.path.exists(LIB):
      build()
    _lib = ctypes.CDLL(LIB)
    for name, typ in (("ocs_rollout_f32", ctypes.c_float), ("ocs_rollout_f64", ctypes.c_double)):
      fn = getattr(_lib, name)
      fn.restype = _I
      fn.argtypes = [_P, _I, _I, _I, _P, _P, _P, _P, _I]
    if _lib.ocs_sizeof_sys() != ctypes.sizeof(_Sys):
      raise RuntimeError("oracle/cstep.cpp and oracle/cstep.py disagree on the System layout")
  return _lib


class CStep:
  """The step of one System in C++ (float32 or float64), OpenMP over envs."""

  def __init__(self, sys, ls_eps: float = float(np.finfo(np.float32).eps),
               capsule_frame: str = "axis"):
    """ls_eps: the line-search slack; capsule_frame: "axis" (the tangent
    along the capsule axis projected on the plane) or "orthogonals" (the
    sphere's frame), the two readings of MJX's plane-capsule frame."""
    lib = load()
    s = sys
    L, nv = s.num_links(), s.qd_size()
    types = {"f": 0, "1": 1, "2": 2, "3": 3}
    qadr, dadr, qi, di = [], [], 0, 0
    from brax_amd.systems import Q_WIDTHS, QD_WIDTHS
    for t in s.link_types:
      qadr.append(qi), dadr.append(di)
      qi += Q_WIDTHS[t]
      di += QD_WIDTHS[t]
    a = lambda x, dt=np.float64: np.ascontiguousarray(np.asarray(x, dt))  # noqa: E731
    self._keep = keep = {}
    keep["ltype"] = a([types[t] for t in s.link_types], np.int32)
    keep["parent"] = a(s.link_parents, np.int32)
    keep["qadr"] = a(qadr, np.int32)
    keep["dadr"] = a(dadr, np.int32)
    tf = lambda t: np.concatenate([np.asarray(t.pos).reshape(L, 3),  # noqa: E731
                                   np.asarray(t.rot).reshape(L, 4)], 1)
    keep["lt"] = a(tf(s.link.transform))
    keep["lj"] = a(tf(s.link.joint))
    keep["li"] = a(tf(s.link.inertia.transform))
    keep["inertia"] = a(np.asarray(s.link.inertia.i).reshape(L, 9))
    keep["mass"] = a(s.link.inertia.mass)
    keep["l_invw"] = a(s.link.invweight)
    keep["motion"] = a(np.concatenate([np.asarray(s.dof.motion.ang).reshape(nv, 3),
                                       np.asarray(s.dof.motion.vel).reshape(nv, 3)], 1))
    lo = np.zeros(nv) if s.dof.limit is None else np.asarray(s.dof.limit[0])
    hi = np.zeros(nv) if s.dof.limit is None else np.asarray(s.dof.limit[1])
    keep["dofp"] = a(np.stack([s.dof.armature, s.dof.stiffness, s.dof.damping,
                               s.dof.invweight, lo, hi], 1))
    keep["solp"] = a(np.asarray(s.dof.solver_params).reshape(nv, 7))
    act = s.actuator
    nu = s.act_size()
    keep["act_q"] = a(act.q_id if nu else np.zeros(1), np.int32)
    keep["act_d"] = a(act.qd_id if nu else np.zeros(1), np.int32)
    keep["actp"] = a(np.stack([np.asarray(act.ctrl_range)[:, 0], np.asarray(act.ctrl_range)[:, 1],
                               np.asarray(act.force_range)[:, 0], np.asarray(act.force_range)[:, 1],
                               act.gain, act.gear, act.bias_q, act.bias_qd], 1)
                     if nu else np.zeros((1, 8)))
    recs = []
    for p in s.contact_pairs():
      for sign in ((0.0,) if p.kind == 0 else (1.0, -1.0)):
        g1, g2 = p.geom1, p.geom2
        recs.append(np.concatenate([
            [p.kind, sign, int(s.geom_bodyid[g1]) - 1], s.geom_pos[g1], s.geom_quat[g1],
            [int(s.geom_bodyid[g2]) - 1], s.geom_pos[g2], s.geom_quat[g2],
            [s.geom_size[g2][0], s.geom_size[g2][1] if p.kind else 0.0, p.friction[0]],
            p.solref, p.solimp,
            [int(s.geom_bodyid[g1]) - 1, int(s.geom_bodyid[g2]) - 1]]))
    rec_len = lib.ocs_con_record()
    keep["con"] = a(np.stack(recs) if recs else np.zeros((1, rec_len)))
    if keep["con"].shape[1] != rec_len:
      raise RuntimeError("contact record length mismatch")
    st = _Sys()
    st.L, st.nq, st.nv, st.nu, st.nc = L, s.q_size(), nv, nu, len(recs)
    st.has_limit = int(s.dof.limit is not None)
    st.mat_inv_iters = int(s.matrix_inv_iterations)
    st.solver_iters, st.solver_maxls = int(s.solver_iterations), int(s.solver_maxls)
    st.dt, st.ls_eps = float(s.opt.timestep), float(ls_eps)
    st.capsule_frame = {"axis": 0, "orthogonals": 1}[capsule_frame]
    for k in range(3):
      st.gravity[k] = float(s.gravity[k])
    for name, arr in keep.items():
      setattr(st, name, arr.ctypes.data)
    self.sys_struct = st
    self.sys = sys
    self.nq, self.nv, self.nu = s.q_size(), nv, nu

  def rollout(self, q, qd, acts, n_frames: int, dtype=np.float32, threads: int = None,
              rows: bool = False):
    """acts (T, N, nu): T env steps of n_frames substeps from (q, qd)
    (N, nq), (N, nv); returns the final (q, qd) as new arrays, and with
    rows=True the mean active constraint rows per substep."""
    lib = load()
    q = np.ascontiguousarray(np.array(q, dtype))
    qd = np.ascontiguousarray(np.array(qd, dtype))
    acts = np.ascontiguousarray(np.asarray(acts, dtype))
    n, T = q.shape[0], acts.shape[0]
    threads = threads or os.cpu_count() or 1
    fn = lib.ocs_rollout_f32 if dtype == np.float32 else lib.ocs_rollout_f64
    nr = np.zeros(n, np.float64)
    rc = fn(ctypes.byref(self.sys_struct), n, T, int(n_frames), q.ctypes.data,
            qd.ctypes.data, acts.ctypes.data if self.nu else None, nr.ctypes.data, int(threads))
    if rc:
      raise RuntimeError("the System exceeds the C++ oracle's fixed sizes")
    if rows:
      return q, qd, float(nr.sum() / max(1, n * T * n_frames))
    return q, qd
