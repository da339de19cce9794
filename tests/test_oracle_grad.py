"""The oracle's autograd gradient of a rollout against float64 central
differences: pins the derivative the adjoint kernel is checked against
(jax.grad through the reference's unrolled solver; the solver's discrete
decisions are constants of the derivative on both sides)."""

import numpy as np
import torch

from brax_amd import systems as S
from oracle.generalized import Oracle


def _loss_fn(o, wq, wqd, n_frames):
  def f(q0, qd0, acts):
    st = o.init(q0, qd0)
    loss = 0.0
    for t in range(acts.shape[0]):
      st = o.env_step(st, acts[t], n_frames)
      loss = loss + (wq[t] * st.q).sum() + (wqd[t] * st.qd).sum()
    return loss
  return f


def test_oracle_grad_matches_central_differences():
  sys = S.quadruped(0)
  o = Oracle(sys, torch.float64, inverse="exact")
  rng = np.random.default_rng(0)
  n, T = 3, 2
  q0 = torch.tensor(sys.init_q + rng.uniform(-0.1, 0.1, (n, 15)), requires_grad=True)
  qd0 = torch.tensor(0.1 * rng.standard_normal((n, 14)), requires_grad=True)
  acts = torch.tensor(rng.uniform(-1, 1, (T, n, 8)), requires_grad=True)
  wq = torch.tensor(rng.standard_normal((T, n, 15)))
  wqd = torch.tensor(rng.standard_normal((T, n, 14)))
  f = _loss_fn(o, wq, wqd, 5)
  grads = torch.autograd.grad(f(q0, qd0, acts), (q0, qd0, acts))
  eps = 1e-6
  base = [q0.detach(), qd0.detach(), acts.detach()]
  checked = 0
  for k, (x, g) in enumerate(zip(base, grads)):
    for _ in range(4):
      idx = tuple(int(rng.integers(0, s)) for s in x.shape)
      xp, xm = x.clone(), x.clone()
      xp[idx] += eps
      xm[idx] -= eps
      ap, am = list(base), list(base)
      ap[k], am[k] = xp, xm
      with torch.no_grad():
        fd = (f(*ap) - f(*am)).item() / (2 * eps)
      assert abs(fd - g[idx].item()) <= 1e-6 * max(1.0, abs(fd)), (k, idx, fd, g[idx].item())
      checked += 1
  assert checked == 12
