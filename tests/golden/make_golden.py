"""Writes the golden vectors of the GPU parity tests (run on the CPU).

  python tests/golden/make_golden.py

Every expected value comes from the float64 oracle (oracle/generalized.py),
which the CPU tests pin against the reference's known answers and closed
forms.  Files:
  quadruped.npz  seeded synthetic-quadruped inputs (q noise U(-0.1, 0.1) on
                 all coordinates, qd 0.1 N(0,1), act U(-1,1)); states after
                 each of 20 env steps (5 substeps) with the exact inverse;
                 the first-substep qf_smooth/qdd; the 1-step state with the
                 reference's Newton-Schulz inverse and its reset flags; the
                 1-step state of the float32 oracle
  quadruped_ns.npz  the same inputs with the reference's literal
                 Newton-Schulz inverse over 20 env steps, and the first
                 substep at which each env's warm start reset
  quadruped_f32.npz the float32 oracle free-running over the same 20 env
                 steps
  pendulum3.npz  contact-free triple pendulum, 20 env steps
  ant.npz        ant, 1 env step (float64, and the float32 oracle)
  humanoid.npz   synthetic humanoid, q noise U(-0.01, 0.01) with the pelvis
                 lowered by U(0, 0.03) m (feet in contact), qd U(-0.01, 0.01),
                 act U(-0.4, 0.4); 3 env steps (float64), 1 step (float32)
  chains.npz     hinge chains of 2, 7, 16 and 32 links, 4 envs each with
                 their own model (seeds 0-3), q U(-0.3, 0.3), qd 0.3 N(0,1),
                 act U(-1, 1); 20 env steps (float64)
  stages.npz     the quadruped's first substep from the golden inputs (first
                 4 envs), stage by stage: frames x, xd, root COM, cdof, the
                 mass matrix, constraint rows J, d, aref, qf_smooth, the
                 solver's force and qf_constraint (float64)
"""

import os
import sys

import numpy as np
import torch

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from brax_amd import systems as S  # noqa: E402
from oracle.generalized import Oracle  # noqa: E402

F64 = torch.float64


def _inputs(sys, n, steps, seed, noise=0.1, act_scale=1.0):
  rng = np.random.default_rng(seed)
  q0 = sys.init_q + rng.uniform(-noise, noise, (n, sys.q_size()))
  qd0 = 0.1 * rng.standard_normal((n, sys.qd_size()))
  acts = rng.uniform(-act_scale, act_scale, (steps, n, max(sys.act_size(), 0)))
  return q0, qd0, acts


def _trajectory(o, q0, qd0, acts, n_frames):
  st = o.init(q0, qd0)
  qs, qds = [], []
  for a in acts:
    st = o.env_step(st, a if a.shape[-1] else None, n_frames)
    qs.append(st.q.numpy().copy())
    qds.append(st.qd.numpy().copy())
  return np.stack(qs), np.stack(qds)


def quadruped(path):
  sys = S.quadruped(0)
  n, steps = 32, 20
  q0, qd0, acts = _inputs(sys, n, steps, seed=0)
  ex = Oracle(sys, F64, inverse="exact")
  q_traj, qd_traj = _trajectory(ex, q0, qd0, acts, 5)
  st1 = ex.step(ex.init(q0, qd0), acts[0])
  o32 = Oracle(sys, torch.float32, inverse="exact")
  st32 = o32.env_step(o32.init(q0, qd0), acts[0], 5)
  ns = Oracle(sys, F64, inverse="ns")
  st = ns.init(q0, qd0)
  x_init = st.mass_mx_inv.clone()
  reset = None
  for f in range(5):
    st = ns.step(st, acts[0])
    if f == 0:
      eye = torch.eye(sys.qd_size(), dtype=F64)
      reset = (torch.linalg.norm((eye - st.mass_mx @ x_init).reshape(n, -1), dim=-1) > 1).numpy()
  np.savez_compressed(
      path, q0=q0, qd0=qd0, acts=acts, q_traj=q_traj, qd_traj=qd_traj,
      qf_smooth1=st1.qf_smooth.numpy(), qdd1=st1.qdd.numpy(),
      q1_ns=st.q.numpy(), qd1_ns=st.qd.numpy(), ns_reset=reset,
      q1_f32=st32.q.double().numpy(), qd1_f32=st32.qd.double().numpy())


def quadruped_ns(path):
  """The reference's literal Newton-Schulz inverse (warm-started, with its
  reset rule) over the 20 env steps of quadruped.npz's inputs: the state
  after every env step, and per env the first substep whose warm start
  reset (-1 if none)."""
  sys = S.quadruped(0)
  g = np.load(os.path.join(HERE, "quadruped.npz"))
  q0, qd0, acts = g["q0"], g["qd0"], g["acts"]
  n = q0.shape[0]
  ns = Oracle(sys, F64, inverse="ns")
  st = ns.init(q0, qd0)
  eye = torch.eye(sys.qd_size(), dtype=F64)
  first_reset = np.full(n, -1)
  qs, qds = [], []
  for t in range(acts.shape[0]):
    for f in range(5):
      x_prev = st.mass_mx_inv.clone()
      st = ns.step(st, acts[t])
      r = (torch.linalg.norm((eye - st.mass_mx @ x_prev).reshape(n, -1), dim=-1) > 1).numpy()
      first_reset = np.where((first_reset < 0) & r, 5 * t + f, first_reset)
    qs.append(st.q.numpy().copy())
    qds.append(st.qd.numpy().copy())
  np.savez_compressed(path, q_traj_ns=np.stack(qs), qd_traj_ns=np.stack(qds),
                      first_reset=first_reset)


def quadruped_f32(path):
  """The float32 oracle (exact inverse) free-running over the 20 env steps of
  quadruped.npz's inputs: the state after every env step.  Beside the float64
  trajectory it shows how far float32 arithmetic of the same algorithm drifts
  on this contact-rich workload."""
  sys = S.quadruped(0)
  g = np.load(os.path.join(HERE, "quadruped.npz"))
  q_traj, qd_traj = _trajectory(Oracle(sys, torch.float32, inverse="exact"), g["q0"],
                                g["qd0"], g["acts"], 5)
  np.savez_compressed(path, q_traj_f32=q_traj.astype(np.float64),
                      qd_traj_f32=qd_traj.astype(np.float64))


def pendulum3(path):
  sys = S.pendulum(3)
  n, steps = 16, 20
  rng = np.random.default_rng(1)
  q0 = rng.uniform(-1, 1, (n, 3))
  qd0 = rng.standard_normal((n, 3))
  acts = np.zeros((steps, n, 0))
  q_traj, qd_traj = _trajectory(Oracle(sys, F64, inverse="exact"), q0, qd0, acts, 5)
  np.savez_compressed(path, q0=q0, qd0=qd0, q_traj=q_traj, qd_traj=qd_traj)


def ant(path):
  sys = S.ant()
  n = 16
  q0, qd0, acts = _inputs(sys, n, 1, seed=2)
  q_traj, qd_traj = _trajectory(Oracle(sys, F64, inverse="exact"), q0, qd0, acts, 5)
  o32 = Oracle(sys, torch.float32, inverse="exact")
  st32 = o32.env_step(o32.init(q0, qd0), acts[0], 5)
  np.savez_compressed(path, q0=q0, qd0=qd0, acts=acts, q_traj=q_traj, qd_traj=qd_traj,
                      q1_f32=st32.q.double().numpy(), qd1_f32=st32.qd.double().numpy())


def humanoid(path):
  sys = S.humanoid(0)
  n, steps = 16, 3
  rng = np.random.default_rng(3)
  q0 = sys.init_q + rng.uniform(-0.01, 0.01, (n, sys.q_size()))
  q0[:, 2] -= rng.uniform(0.0, 0.03, n)
  qd0 = rng.uniform(-0.01, 0.01, (n, sys.qd_size()))
  acts = rng.uniform(-0.4, 0.4, (steps, n, sys.act_size()))
  q_traj, qd_traj = _trajectory(Oracle(sys, F64, inverse="exact"), q0, qd0, acts, 5)
  o32 = Oracle(sys, torch.float32, inverse="exact")
  st32 = o32.env_step(o32.init(q0, qd0), acts[0], 5)
  np.savez_compressed(path, q0=q0, qd0=qd0, acts=acts, q_traj=q_traj, qd_traj=qd_traj,
                      q1_f32=st32.q.double().numpy(), qd1_f32=st32.qd.double().numpy())


def chains(path):
  out = {}
  for L in (2, 7, 16, 32):
    rng = np.random.default_rng(L)
    n, steps = 4, 20
    q0 = rng.uniform(-0.3, 0.3, (n, L))
    qd0 = 0.3 * rng.standard_normal((n, L))
    acts = rng.uniform(-1, 1, (steps, n, L))
    qs, qds = [], []
    for e in range(n):
      q, qd = _trajectory(Oracle(S.chain(L, e), F64, inverse="exact"), q0[e:e + 1],
                          qd0[e:e + 1], acts[:, e:e + 1], 5)
      qs.append(q[:, 0]), qds.append(qd[:, 0])
    out.update({f"q0_{L}": q0, f"qd0_{L}": qd0, f"acts_{L}": acts,
                f"q_traj_{L}": np.stack(qs, 1), f"qd_traj_{L}": np.stack(qds, 1)})
  np.savez_compressed(path, **out)


def stages(path):
  sys = S.quadruped(0)
  g = np.load(os.path.join(HERE, "quadruped.npz"))
  q0, qd0, act = g["q0"][:4], g["qd0"][:4], g["acts"][0, :4]
  o = Oracle(sys, F64, inverse="exact")
  st0 = o.init(q0, qd0)
  st1 = o.step(st0, act)
  np.savez_compressed(
      path, q0=q0, qd0=qd0, act=act, x_pos=st0.x_pos.numpy(), x_rot=st0.x_rot.numpy(),
      xd_ang=st0.xd_ang.numpy(), xd_vel=st0.xd_vel.numpy(), root_com=st0.root_com.numpy(),
      cdof_ang=st0.cdof[0].numpy(), cdof_vel=st0.cdof[1].numpy(),
      mass_mx=st0.mass_mx.numpy(), con_jac=st0.con_jac.numpy(),
      con_diag=st0.con_diag.numpy(), con_aref=st0.con_aref.numpy(),
      qf_smooth=st1.qf_smooth.numpy(), force=st1.pg_info["force"].numpy(),
      qf_constraint=st1.qf_constraint.numpy(), q1=st1.q.numpy(), qd1=st1.qd.numpy())


if __name__ == "__main__":
  quadruped(os.path.join(HERE, "quadruped.npz"))
  quadruped_ns(os.path.join(HERE, "quadruped_ns.npz"))
  quadruped_f32(os.path.join(HERE, "quadruped_f32.npz"))
  pendulum3(os.path.join(HERE, "pendulum3.npz"))
  ant(os.path.join(HERE, "ant.npz"))
  humanoid(os.path.join(HERE, "humanoid.npz"))
  chains(os.path.join(HERE, "chains.npz"))
  stages(os.path.join(HERE, "stages.npz"))
  print("written")
