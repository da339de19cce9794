"""The committed golden vectors (tests/golden/, written by make_golden.py
from the float64 torch oracle) reproduced by both oracles: the torch oracle
stage by stage on the quadruped's first substep, and the C++ restatement
(oracle/cstep.cpp, written independently of the torch code) over every
golden trajectory: quadruped (20 env steps, contacts and limits), ant,
humanoid, triple pendulum and hinge chains of 2-32 links."""

import os

import numpy as np
import pytest
import torch

from brax_amd import systems as S
from oracle.cstep import CStep
from oracle.generalized import Oracle

G = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def test_torch_oracle_reproduces_quadruped_stages():
  g = np.load(os.path.join(G, "stages.npz"))
  o = Oracle(S.quadruped(0), torch.float64, inverse="exact")
  st0 = o.init(g["q0"], g["qd0"])
  st1 = o.step(st0, g["act"])
  got = {"x_pos": st0.x_pos, "x_rot": st0.x_rot, "xd_ang": st0.xd_ang, "xd_vel": st0.xd_vel,
         "root_com": st0.root_com, "cdof_ang": st0.cdof[0], "cdof_vel": st0.cdof[1],
         "mass_mx": st0.mass_mx, "con_jac": st0.con_jac, "con_diag": st0.con_diag,
         "con_aref": st0.con_aref, "qf_smooth": st1.qf_smooth, "force": st1.pg_info["force"],
         "qf_constraint": st1.qf_constraint, "q1": st1.q, "qd1": st1.qd}
  for k, v in got.items():
    np.testing.assert_allclose(v.numpy(), g[k], rtol=1e-12, atol=1e-12, err_msg=k)
  # the stages are not vacuous: the feet touch the floor
  assert (g["con_diag"] > 0).sum() > 0 and np.abs(g["force"]).max() > 0


CASES = [("quadruped", lambda: S.quadruped(0)), ("ant", S.ant),
         ("humanoid", lambda: S.humanoid(0)), ("pendulum3", lambda: S.pendulum(3))]


@pytest.mark.parametrize("name,make", CASES)
def test_cstep_reproduces_golden_trajectory(name, make):
  g = np.load(os.path.join(G, name + ".npz"))
  cs = CStep(make())
  q, qd = g["q0"], g["qd0"]
  steps = g["q_traj"].shape[0]
  acts = g["acts"] if "acts" in g.files else np.zeros((steps, q.shape[0], 0))
  for t in range(steps):
    q, qd = cs.rollout(q, qd, acts[t:t + 1], 5, dtype=np.float64, threads=4)
    np.testing.assert_allclose(q, g["q_traj"][t], rtol=1e-9, atol=1e-9)
    np.testing.assert_allclose(qd, g["qd_traj"][t], rtol=1e-9, atol=1e-9)


@pytest.mark.parametrize("L", [2, 7, 16, 32])
def test_cstep_reproduces_golden_chains(L):
  g = np.load(os.path.join(G, "chains.npz"))
  for e in range(4):
    cs = CStep(S.chain(L, e))
    q, qd = g[f"q0_{L}"][e:e + 1], g[f"qd0_{L}"][e:e + 1]
    for t in range(20):
      q, qd = cs.rollout(q, qd, g[f"acts_{L}"][t:t + 1, e:e + 1], 5, dtype=np.float64,
                         threads=1)
    np.testing.assert_allclose(q[0], g[f"q_traj_{L}"][19, e], rtol=1e-9, atol=1e-9)
    np.testing.assert_allclose(qd[0], g[f"qd_traj_{L}"][19, e], rtol=1e-9, atol=1e-9)
