"""Per-env model tables and heterogeneous batches on the GPU.

Chains (hinge chains with per-env lengths and masses) are contact-free, so
float32 trajectories track the float64 oracle tightly over 20 env steps.
A mixed batch must give every env the result it gets in a homogeneous batch
of its own kind.
"""

import numpy as np
import pytest
import torch

from brax_amd import systems as S
from brax_amd.generalized import hetero, pipeline
from oracle.generalized import Oracle

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda", 0)


def _ratio(gpu, ref):
  gpu = np.asarray(gpu, np.float64)
  ref = np.asarray(ref, np.float64)
  return np.abs(gpu - ref) / (1e-5 + 1e-4 * np.abs(ref))


def _oracle_traj(sys, q0, qd0, acts, n_frames=5, dtype=torch.float64):
  o = Oracle(sys, dtype, inverse="exact")
  st = o.init(q0, qd0)
  qs, qds = [], []
  for a in acts:
    st = o.env_step(st, a, n_frames)
    qs.append(st.q.double().numpy().copy())
    qds.append(st.qd.double().numpy().copy())
  return np.stack(qs), np.stack(qds)


@pytest.mark.parametrize("n_links", [2, 7])
def test_chain_batch_per_env_tables_20_steps(n_links):
  """Per-env model tables, short chains: 20 env steps free-running within
  1e-4 relative (1e-5 absolute) of the float64 oracle."""
  n, T = 6, 20
  systems = [S.chain(n_links, seed) for seed in range(n)]
  batch = pipeline.SystemBatch(systems)
  rng = np.random.default_rng(n_links)
  q0 = rng.uniform(-0.3, 0.3, (n, n_links)).astype(np.float32)
  qd0 = (0.3 * rng.standard_normal((n, n_links))).astype(np.float32)
  acts = rng.uniform(-1, 1, (T, n, n_links)).astype(np.float32)
  st = pipeline.init(batch, torch.tensor(q0, device=DEV), torch.tensor(qd0, device=DEV))
  _, (qt, qdt) = pipeline.rollout(batch, st, torch.tensor(acts, device=DEV), 5,
                                  trajectory=True)
  torch.cuda.synchronize()
  qt, qdt = qt.cpu().numpy(), qdt.cpu().numpy()
  worst = 0.0
  for e in range(n):
    rq, rqd = _oracle_traj(systems[e], q0[e:e + 1].astype(np.float64),
                           qd0[e:e + 1].astype(np.float64), acts[:, e:e + 1].astype(np.float64))
    worst = max(worst, _ratio(qt[:, e], rq[:, 0]).max(), _ratio(qdt[:, e], rqd[:, 0]).max())
  print(f"\nchain{n_links}: worst tolerance ratio over 20 steps {worst:.3f}", flush=True)
  assert worst <= 1.0


@pytest.mark.parametrize("n_links", [16, 32])
def test_long_chain_teacher_forced(n_links):
  """Long chains with link masses spanning 0.1-10 kg are ill-conditioned:
  the float32 restatement of the algorithm (the oracle in float32) leaves
  float64 by more than 1e-4 relative within one env step.  Each env step of
  the GPU trajectory is checked from the GPU's own state: its deviation from
  float64 must be of the float32 oracle's size (within 4x on 90% of the env
  steps, within 20x on all), or within the tolerance."""
  n, T = 4, 10
  systems = [S.chain(n_links, seed) for seed in range(n)]
  batch = pipeline.SystemBatch(systems)
  rng = np.random.default_rng(n_links)
  q0 = rng.uniform(-0.3, 0.3, (n, n_links)).astype(np.float32)
  qd0 = (0.3 * rng.standard_normal((n, n_links))).astype(np.float32)
  acts = rng.uniform(-1, 1, (T, n, n_links)).astype(np.float32)
  st = pipeline.init(batch, torch.tensor(q0, device=DEV), torch.tensor(qd0, device=DEV))
  _, (qt, qdt) = pipeline.rollout(batch, st, torch.tensor(acts, device=DEV), 5,
                                  trajectory=True)
  torch.cuda.synchronize()
  qt, qdt = qt.double().cpu().numpy(), qdt.double().cpu().numpy()
  pairs = []
  for e in range(n):
    o64 = Oracle(systems[e], torch.float64, inverse="exact")
    o32 = Oracle(systems[e], torch.float32, inverse="exact")
    qs, qds = q0[e:e + 1].astype(np.float64), qd0[e:e + 1].astype(np.float64)
    for t in range(T):
      a = acts[t, e:e + 1].astype(np.float64)
      r64 = o64.env_step(o64.init(qs, qds), a, 5)
      r32 = o32.env_step(o32.init(qs, qds), a, 5)
      rq, rqd = r64.q.numpy(), r64.qd.numpy()
      g = max(_ratio(qt[t, e], rq[0]).max(), _ratio(qdt[t, e], rqd[0]).max())
      f = max(_ratio(r32.q.double().numpy()[0], rq[0]).max(),
              _ratio(r32.qd.double().numpy()[0], rqd[0]).max())
      pairs.append((g, f))
      qs, qds = qt[t, e:e + 1], qdt[t, e:e + 1]
  g = np.array([p[0] for p in pairs])
  f = np.array([p[1] for p in pairs])
  ok4 = (g <= np.maximum(1.0, 4.0 * f)).mean()
  print(f"\nchain{n_links} teacher-forced: gpu ratio median {np.median(g):.2f} max {g.max():.2f}; "
        f"float32 oracle median {np.median(f):.2f} max {f.max():.2f}; within 4x: {ok4:.2f}",
        flush=True)
  assert ok4 >= 0.9
  assert (g <= np.maximum(1.0, 20.0 * f)).all()


@pytest.mark.parametrize("n_links", [2, 7])
def test_chain_golden_20_steps(n_links):
  """The committed chain golden vectors (float64 oracle, 4 envs with their
  own models): 20 env steps free-running within 1e-4 relative (1e-5
  absolute)."""
  import os
  g = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "chains.npz"))
  L = n_links
  batch = pipeline.SystemBatch([S.chain(L, e) for e in range(4)])
  st = pipeline.init(batch, torch.tensor(g[f"q0_{L}"], dtype=torch.float32, device=DEV),
                     torch.tensor(g[f"qd0_{L}"], dtype=torch.float32, device=DEV))
  _, (qt, qdt) = pipeline.rollout(batch, st, torch.tensor(g[f"acts_{L}"], dtype=torch.float32,
                                                          device=DEV), 5, trajectory=True)
  torch.cuda.synchronize()
  worst = max(_ratio(qt.cpu().numpy(), g[f"q_traj_{L}"]).max(),
              _ratio(qdt.cpu().numpy(), g[f"qd_traj_{L}"]).max())
  print(f"\nchain{L} golden: worst tolerance ratio over 20 steps {worst:.3f}", flush=True)
  assert worst <= 1.0


def _homogeneous(sys, quad, q0, qd0, acts):
  """One env alone: a shared-table model for the quadruped, a one-env
  SystemBatch for a chain (the kernels of pipeline.rollout)."""
  obj = pipeline.SystemBatch([sys]) if sys is not quad else quad
  st = pipeline.init(obj, torch.tensor(q0[None], device=DEV), torch.tensor(qd0[None], device=DEV))
  st, (qt, qdt) = pipeline.rollout(obj, st, torch.tensor(acts[:, None], device=DEV), 5,
                                   trajectory=True)
  torch.cuda.synchronize()
  return st.q[0].cpu().numpy(), st.qd[0].cpu().numpy(), qt[:, 0].cpu().numpy(), qdt[:, 0].cpu().numpy()


def test_hetero_batch_matches_homogeneous_runs():
  """One persistent launch over quadrupeds and chains of several lengths:
  every env's final state and trajectory bitwise equal a run of its own."""
  rng = np.random.default_rng(11)
  quad = S.quadruped(0)
  chains = [S.chain(L, s) for s, L in enumerate([2, 5, 13, 32, 5, 32, 3, 17, 9])]
  systems = [quad] * 70 + chains
  perm = rng.permutation(len(systems))
  systems = [systems[i] for i in perm]
  hb = hetero.HeteroBatch(systems, DEV)
  assert len(hb.groups) == 8  # quadrupeds share one table; a group per chain length
  T = 4
  q0 = [(s.init_q + rng.uniform(-0.1, 0.1, s.q_size())).astype(np.float32) for s in systems]
  qd0 = [(0.1 * rng.standard_normal(s.qd_size())).astype(np.float32) for s in systems]
  acts = [rng.uniform(-1, 1, (T, s.act_size())).astype(np.float32) for s in systems]
  q, qd, ap = hb.pack(q0, "q"), hb.pack(qd0, "qd"), hb.pack(acts, "act")
  status, qt, qdt = hb.rollout(q, qd, ap, T, 5, trajectory=True)
  torch.cuda.synchronize()
  assert int(status.abs().sum()) == 0
  for e, s in enumerate(systems):
    rq, rqd, rqt, rqdt = _homogeneous(s, quad, q0[e], qd0[e], acts[e])
    sq, sqd, _ = hb.env_rows(e)
    np.testing.assert_array_equal(q[sq].cpu().numpy(), rq)
    np.testing.assert_array_equal(qd[sqd].cpu().numpy(), rqd)
    np.testing.assert_array_equal(qt[:, sq].cpu().numpy(), rqt)
    np.testing.assert_array_equal(qdt[:, sqd].cpu().numpy(), rqdt)


def test_hetero_batch_full_size():
  """The heterogeneous workload at its own size: 32768 envs, half
  quadrupeds, half chains of 2-32 links (8 seeded models per length), in one
  persistent launch of 20 env steps; a seeded sample of 64 envs bitwise
  equals runs of their own, and every env stays finite."""
  rng = np.random.default_rng(0)
  quad = S.quadruped(0)
  pool = {L: [S.chain(L, 1000 * L + j) for j in range(8)] for L in range(2, 33)}
  n = 32768
  lengths = rng.integers(2, 33, n // 2)
  systems = [quad] * (n - n // 2) + [pool[int(L)][int(rng.integers(8))] for L in lengths]
  systems = [systems[i] for i in rng.permutation(n)]
  hb = hetero.HeteroBatch(systems, DEV)
  T = 20
  q0 = [(s.init_q + rng.uniform(-0.1, 0.1, s.q_size())).astype(np.float32) for s in systems]
  qd0 = [(0.1 * rng.standard_normal(s.qd_size())).astype(np.float32) for s in systems]
  acts = [rng.uniform(-1, 1, (T, s.act_size())).astype(np.float32) for s in systems]
  q, qd, ap = hb.pack(q0, "q"), hb.pack(qd0, "qd"), hb.pack(acts, "act")
  status = hb.rollout(q, qd, ap, T, 5)
  torch.cuda.synchronize()
  print(f"\n{n} mixed envs: {len(hb.groups)} models, {hb.n_items} work items, "
        f"{hb.grid} workgroups; non-finite envs {int((status.cpu().numpy() & 2).astype(bool).sum())}",
        flush=True)
  assert torch.isfinite(q).all() and torch.isfinite(qd).all()
  sample = np.sort(np.random.default_rng(1).choice(n, 64, replace=False))
  for e in sample:
    s = systems[e]
    rq, rqd, _, _ = _homogeneous(s, quad, q0[e], qd0[e], acts[e])
    sq, sqd, _ = hb.env_rows(e)
    np.testing.assert_array_equal(q[sq].cpu().numpy(), rq)
    np.testing.assert_array_equal(qd[sqd].cpu().numpy(), rqd)


def test_hetero_batch_graph_replays():
  """The batch launch captured in a HIP graph and replayed three times (the
  work counter resets itself): each replay from the same state gives the
  eager result."""
  rng = np.random.default_rng(5)
  quad = S.quadruped(0)
  systems = [quad] * 100 + [S.chain(L, L) for L in range(2, 33)]
  hb = hetero.HeteroBatch(systems, DEV)
  T = 3
  q0 = hb.pack([(s.init_q + rng.uniform(-0.1, 0.1, s.q_size())).astype(np.float32)
                for s in systems], "q")
  qd0 = hb.pack([(0.1 * rng.standard_normal(s.qd_size())).astype(np.float32)
                 for s in systems], "qd")
  ap = hb.pack([rng.uniform(-1, 1, (T, s.act_size())).astype(np.float32) for s in systems],
               "act")
  q, qd = q0.clone(), qd0.clone()
  hb.launch_rollout(q, qd, ap, T, 5)
  torch.cuda.synchronize()
  ref_q, ref_qd = q.clone(), qd.clone()
  side = torch.cuda.Stream(device=DEV)
  qg, qdg = q0.clone(), qd0.clone()
  graph = torch.cuda.CUDAGraph()
  with torch.cuda.graph(graph, stream=side):
    hb.launch_rollout(qg, qdg, ap, T, 5)
  for _ in range(3):
    qg.copy_(q0)
    qdg.copy_(qd0)
    torch.cuda.synchronize()
    graph.replay()
    torch.cuda.synchronize()
    assert torch.equal(qg, ref_q) and torch.equal(qdg, ref_qd)
