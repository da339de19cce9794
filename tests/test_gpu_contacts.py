"""The debug contact output against the oracle's MJX plane collisions
(contact.py:28-67): distance, point and frame of every candidate, for the
quadruped (spheres and capsules), the ant and the humanoid."""

import numpy as np
import pytest
import torch

from brax_amd import systems as S
from brax_amd.generalized import pipeline
from oracle.generalized import Oracle

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda", 0)


@pytest.mark.parametrize("name,make", [("quadruped", lambda: S.quadruped(0)), ("ant", S.ant),
                                       ("humanoid", lambda: S.humanoid(0))])
def test_contacts_match_oracle(name, make):
  sys = make()
  rng = np.random.default_rng(4)
  n = 32
  q0 = (sys.init_q + rng.uniform(-0.1, 0.1, (n, sys.q_size()))).astype(np.float32)
  q0[:, 2] -= 0.1  # lower the root: some candidates touch the floor
  qd0 = np.zeros((n, sys.qd_size()), np.float32)
  st = pipeline.init(sys, torch.tensor(q0, device=DEV), torch.tensor(qd0, device=DEV),
                     debug=True)
  torch.cuda.synchronize()
  c = st.contact
  o = Oracle(sys, torch.float64, inverse="exact")
  ost = o.init(q0.astype(np.float64), qd0.astype(np.float64))
  ref = o.collide(ost.x_pos, ost.x_rot)
  np.testing.assert_allclose(c.dist.cpu().numpy(), ref["dist"].numpy(), atol=2e-5)
  np.testing.assert_allclose(c.pos.cpu().numpy(), ref["pos"].numpy(), atol=2e-5)
  np.testing.assert_allclose(c.frame.cpu().numpy(), ref["frame"].numpy(), atol=2e-5)
  assert list(c.link_idx[1].cpu().numpy()) == list(ref["link_b"])
  # some candidates touch the floor in these poses
  assert (c.dist < 0).any()


def test_step_debug_attaches_contacts_of_the_new_state():
  sys = S.quadruped(0)
  q0 = torch.tensor(np.tile(sys.init_q, (4, 1)).astype(np.float32), device=DEV)
  st = pipeline.init(sys, q0, torch.zeros(4, sys.qd_size(), device=DEV))
  st = pipeline.pipeline_step(sys, st, torch.zeros(4, sys.act_size(), device=DEV), 5,
                              debug=True)
  torch.cuda.synchronize()
  again = pipeline.contacts(sys, st.q)
  np.testing.assert_array_equal(st.contact.dist.cpu().numpy(), again.dist.cpu().numpy())
