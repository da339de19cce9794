"""Launch shapes: the results of the step and adjoint kernels do not depend
on the envs per workgroup (every lane runs the same instructions on the same
values whatever the workgroup's lane count), and the autotuner picks one of
its candidates and sets it on the model."""

import numpy as np
import pytest
import torch

from brax_amd import systems as S
from brax_amd.generalized import autotune, pipeline

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda", 0)


def _inputs(sys, n, T, seed):
  rng = np.random.default_rng(seed)
  q0 = (sys.init_q + rng.uniform(-0.1, 0.1, (n, sys.q_size()))).astype(np.float32)
  qd0 = (0.1 * rng.standard_normal((n, sys.qd_size()))).astype(np.float32)
  acts = rng.uniform(-1, 1, (T, n, sys.act_size())).astype(np.float32)
  return (torch.tensor(q0, device=DEV), torch.tensor(qd0, device=DEV),
          torch.tensor(acts, device=DEV))


def _rollout(sys, q0, qd0, acts, epw):
  m = pipeline.device_model(sys, DEV)
  autotune.set_launch(m, autotune.STEP, epw)
  try:
    st = pipeline.State(q=q0.clone(), qd=qd0.clone(), x=None, xd=None)
    st, (qt, qdt) = pipeline.rollout(sys, st, acts, 5, trajectory=True)
    torch.cuda.synchronize()
    return [t.cpu().numpy() for t in (qt, qdt, st.q, st.qd, st.x.pos, st.xd.vel)]
  finally:
    autotune.set_launch(m, autotune.STEP, 0)


@pytest.mark.parametrize("name,n", [("quadruped", 200), ("humanoid", 70)])
def test_rollout_bitwise_across_envs_per_workgroup(name, n):
  """A ragged batch (not a multiple of 64) over 6 env steps: bitwise the
  same trajectory and frames for 64, 16, 4 and 1 envs per workgroup."""
  sys = getattr(S, name)(0)
  q0, qd0, acts = _inputs(sys, n, 6, seed=3)
  ref = _rollout(sys, q0, qd0, acts, 64)
  for epw in (16, 4, 1):
    got = _rollout(sys, q0, qd0, acts, epw)
    for a, b in zip(ref, got):
      np.testing.assert_array_equal(a, b)


def test_chain_batch_bitwise_across_envs_per_workgroup():
  systems = [S.chain(9, seed) for seed in range(37)]
  batch = pipeline.SystemBatch(systems)
  rng = np.random.default_rng(9)
  q0 = torch.tensor(rng.uniform(-0.3, 0.3, (37, 9)).astype(np.float32), device=DEV)
  qd0 = torch.tensor((0.3 * rng.standard_normal((37, 9))).astype(np.float32), device=DEV)
  acts = torch.tensor(rng.uniform(-1, 1, (4, 37, 9)).astype(np.float32), device=DEV)
  ref = _rollout(batch, q0, qd0, acts, 64)
  got = _rollout(batch, q0, qd0, acts, 8)
  for a, b in zip(ref, got):
    np.testing.assert_array_equal(a, b)


def test_adjoint_bitwise_across_envs_per_workgroup():
  sys = S.quadruped(0)
  q0, qd0, acts = _inputs(sys, 100, 6, seed=4)
  m = pipeline.device_model(sys, DEV)
  out = []
  for epw in (64, 8):
    autotune.set_launch(m, autotune.ADJOINT, epw)
    try:
      qd = qd0.clone().requires_grad_(True)
      a = acts.clone().requires_grad_(True)
      qt, qdt, _ = pipeline.rollout_grad(sys, q0, qd, a, 5, ckpt_every=4)
      (-qdt[:, :, 0].sum() + 1e-3 * (a * a).sum()).backward()
      torch.cuda.synchronize()
      out.append((a.grad.cpu().numpy(), qd.grad.cpu().numpy()))
    finally:
      autotune.set_launch(m, autotune.ADJOINT, 0)
  np.testing.assert_array_equal(out[0][0], out[1][0])
  np.testing.assert_array_equal(out[0][1], out[1][1])
  assert np.isfinite(out[0][0]).all()


@pytest.mark.parametrize("name,n,parts", [("quadruped", 300, 4), ("humanoid", 130, 2)])
def test_adjoint_bitwise_across_stream_parts(name, n, parts):
  """The adjoint's backward pass split into env parts on several streams
  (300 quadrupeds: parts of 128, 128 and 44 envs; 130 humanoids: 128 and 2)
  gives the gradients of one part bit for bit."""
  sys = getattr(S, name)(0)
  q0, qd0, acts = _inputs(sys, n, 6, seed=5)
  if name == "humanoid":
    acts = 0.4 * acts
  m = pipeline.device_model(sys, DEV)
  out = []
  for k in (1, parts):
    autotune.set_launch(m, autotune.ADJOINT_PARTS, k)
    try:
      q = q0.clone().requires_grad_(True)
      qd = qd0.clone().requires_grad_(True)
      a = acts.clone().requires_grad_(True)
      qt, qdt, _ = pipeline.rollout_grad(sys, q, qd, a, 5, ckpt_every=4)
      (-qdt[:, :, 0].sum() + (qt * qt).sum() + 1e-3 * (a * a).sum()).backward()
      torch.cuda.synchronize()
      out.append([t.grad.cpu().numpy() for t in (q, qd, a)])
    finally:
      autotune.set_launch(m, autotune.ADJOINT_PARTS, 0)
  for x, y in zip(*out):
    np.testing.assert_array_equal(x, y)
  assert all(np.isfinite(x).all() for x in out[0])
  assert autotune.get_launch(m, autotune.ADJOINT_PARTS) == 0


def test_autotune_sets_a_candidate():
  sys = S.quadruped(0)
  q0, qd0, acts = _inputs(sys, 512, 2, seed=5)
  m = pipeline.device_model(sys, DEV)
  try:
    epw = autotune.tune_step(sys, q0, qd0, acts, 5)
    assert epw in autotune.CANDIDATES
    assert autotune.get_launch(m, autotune.STEP) == epw
    epw = autotune.tune_adjoint(sys, q0, qd0, acts, 5)
    assert epw in autotune.CANDIDATES
    assert autotune.get_launch(m, autotune.ADJOINT) == epw
  finally:
    autotune.set_launch(m, autotune.STEP, 0)
    autotune.set_launch(m, autotune.ADJOINT, 0)


def test_lds_tiles_agree():
  """The quadruped's step kernels compiled for 4, 6 (default) and 8 fast
  contact slots (the LDS tile the autotuner picks among): 64 envs with the
  torso low enough that 2-10 candidates touch the floor, 4 env steps.  An
  env takes the slot-per-candidate path exactly when its active contacts
  exceed the tile (status bit), and every tile's states agree with the
  default tile's within the band."""
  from brax_amd.generalized import autotune
  sys = S.quadruped(0)
  assert autotune.tiles(sys) == (6, 4, 8)
  rng = np.random.default_rng(41)
  n = 64
  q0 = sys.init_q + rng.uniform(-0.1, 0.1, (n, 15))
  q0[:, 2] = rng.uniform(0.05, 0.3, n)
  qd0 = 0.1 * rng.standard_normal((n, 14))
  acts = rng.uniform(-1, 1, (4, n, 8))
  t = lambda x: torch.tensor(np.asarray(x, np.float32), device="cuda")  # noqa: E731
  m = pipeline.device_model(sys)
  out = {}
  for k in (6, 4, 8):
    autotune.set_launch(m, autotune.TILE, k)
    st = pipeline.init(sys, t(q0), t(qd0))
    st, _ = pipeline.rollout(sys, st, t(acts), 5)
    torch.cuda.synchronize()
    out[k] = (st.q.cpu().numpy(), st.qd.cpu().numpy(), st.status.cpu().numpy())
  autotune.set_launch(m, autotune.TILE, 0)
  slow = {k: int(((v[2] & 1) > 0).sum()) for k, v in out.items()}
  print(f"\nenvs on the slot-per-candidate path per tile {slow}", flush=True)
  assert slow[4] >= slow[6] >= slow[8] and slow[4] > slow[8]
  for k in (4, 8):
    for a, b in zip(out[k][:2], out[6][:2]):
      r = np.abs(a.astype(np.float64) - b) / (1e-5 + 1e-4 * np.abs(b))
      eq = int(np.all(a == b, axis=1).sum())
      print(f"tile {k} vs 6: max ratio {r.max():.3g}, envs bitwise equal {eq}/{n}", flush=True)
      assert r.max() <= 1.0
