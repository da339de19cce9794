"""bgx_rollout_adjoint_noalloc (no workspace argument, a static per-GPU
scratch array) against bgx_rollout_adjoint with a caller workspace, through
the C ABI: bitwise the same gradients on two calls in a row, the second with
fewer envs and a shorter horizon than the first (the scratch keeps the first
call's contents), and a clear error for a problem larger than the scratch.
"""

import ctypes

import numpy as np
import pytest
import torch

from brax_amd import _lib
from brax_amd import systems as S
from brax_amd.generalized import pipeline

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda", 0)


def _case(m, n, T, ckpt, seed):
  rng = np.random.default_rng(seed)
  sys = m.sys
  t = lambda x: torch.tensor(np.asarray(x, np.float32), device=DEV)  # noqa: E731
  q0 = t(sys.init_q + rng.uniform(-0.1, 0.1, (n, m.nq)))
  qd0 = t(0.1 * rng.standard_normal((n, m.nv)))
  acts = t(rng.uniform(-1, 1, (T, n, m.nu)))
  lib = _lib.load()
  q, qd = q0.clone(), qd0.clone()
  qt = torch.empty(T, n, m.nq, device=DEV)
  qdt = torch.empty(T, n, m.nv, device=DEV)
  _lib.check(lib.bgx_rollout(m.handle, n, T, 5, _lib.ptr(q), _lib.ptr(qd), _lib.ptr(acts),
                             n * m.nu, _lib.ptr(qt), _lib.ptr(qdt), None, None, None, None,
                             None, None))
  idx = list(range(ckpt - 1, T - 1, ckpt))
  ckq = torch.cat([q0[None], qt[idx]]).contiguous()
  ckqd = torch.cat([qd0[None], qdt[idx]]).contiguous()
  gq = t(rng.standard_normal((T, n, m.nq)))
  gqd = t(rng.standard_normal((T, n, m.nv)))
  return ckq, ckqd, acts, gq, gqd


def _run(m, n, T, ckpt, case, noalloc):
  ckq, ckqd, acts, gq, gqd = case
  lib = _lib.load()
  ga = torch.full((T, n, m.nu), float("nan"), device=DEV)
  g0 = torch.full((n, m.nq), float("nan"), device=DEV)
  gd0 = torch.full((n, m.nv), float("nan"), device=DEV)
  status = torch.full((n,), -1, dtype=torch.int32, device=DEV)
  common = (m.handle, n, T, 5, ckpt, _lib.ptr(ckq), _lib.ptr(ckqd), _lib.ptr(acts), n * m.nu,
            _lib.ptr(gq), _lib.ptr(gqd), _lib.ptr(ga), _lib.ptr(g0), _lib.ptr(gd0))
  if noalloc:
    _lib.check(lib.bgx_rollout_adjoint_noalloc(*common, _lib.ptr(status), None))
  else:
    nbytes = lib.bgx_adjoint_workspace_bytes(m.handle, n, ckpt, 5)
    ws = torch.empty(nbytes // 4 + 1, device=DEV)
    _lib.check(lib.bgx_rollout_adjoint(*common, _lib.ptr(ws), nbytes, _lib.ptr(status), None))
  torch.cuda.synchronize()
  return [x.cpu().numpy() for x in (ga, g0, gd0, status)]


def test_noalloc_matches_workspace_entry_point():
  lib = _lib.load()
  m = pipeline.device_model(S.quadruped(0), DEV)
  cap = lib.bgx_adjoint_scratch_bytes()
  need = lib.bgx_adjoint_workspace_min_bytes(m.handle, 4096, 16, 5)
  print(f"\nscratch {cap} bytes; least workspace of 4096 quadrupeds, checkpoint 16: {need}")
  assert 0 < need <= cap
  # two calls in a row, the second smaller; each against the workspace entry point
  for n, T, ckpt, seed in ((1024, 24, 8, 1), (200, 7, 3, 2)):
    case = _case(m, n, T, ckpt, seed)
    ref = _run(m, n, T, ckpt, case, noalloc=False)
    got = _run(m, n, T, ckpt, case, noalloc=True)
    for a, b in zip(ref[:3], got[:3]):
      assert np.isfinite(a).all()
      np.testing.assert_array_equal(a, b)
    np.testing.assert_array_equal(ref[3], got[3])


def test_noalloc_rejects_a_problem_larger_than_its_scratch():
  lib = _lib.load()
  m = pipeline.device_model(S.quadruped(0), DEV)
  n = 65536
  assert lib.bgx_adjoint_workspace_min_bytes(m.handle, n, 64, 5) > lib.bgx_adjoint_scratch_bytes()
  z = torch.zeros(1, device=DEV)
  rc = lib.bgx_rollout_adjoint_noalloc(m.handle, n, 64, 5, 64, _lib.ptr(z), _lib.ptr(z),
                                       _lib.ptr(z), n * m.nu, None, None, _lib.ptr(z),
                                       _lib.ptr(z), _lib.ptr(z), None, None)
  assert rc == 1
  assert b"scratch holds" in lib.bgx_last_error()
