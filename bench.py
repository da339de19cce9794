"""Throughput of the generalized-coordinates step on one MI355X.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--envs N]

Workload: the synthetic quadruped (free torso + 4 legs x 2 hinged capsules,
qd size 14), N = 65536 envs, seeded q = default + U(-0.1, 0.1),
qd = 0.1 N(0,1), actions U(-1, 1) pre-generated for every env step and
resident in HBM.  One bench step = one env step of every env = 5 substeps
(dt 0.01) of pipeline.step with one action (PipelineEnv.pipeline_step).  The
timed region replays one HIP graph holding a single rollout launch of K env
steps (the state stays in registers between steps).

Prints one JSON line: env-steps/s, the FP32 roofline of the step kernel and a
CPU baseline (the oracle restatement in float32 on this host's cores, on a
bounded sample of the same workload).
"""

from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FP32_PEAK_TFLOPS = 157.3  # MI355X published FP32 (vector) peak
HBM_PEAK_GBS = 8000.0
N_FRAMES = 5


def flop_per_substep(L, nv, C, n_lim, G, K, R_a):
  """Algorithmic FLOP of one substep (op-count model; FMA = 2 FLOP).

  Minimal result-preserving formulation: tree passes, CRBA + Cholesky,
  contact Jacobians over all candidates, A and b over the R_a active rows,
  K projected-gradient updates with one line-search trial each, output.
  Frozen before measuring; R_a is the oracle's mean active-row count.
  """
  tree = 720 * L + 190 * nv
  mass = 5.5 * nv * (nv + 1) + 36 * nv + nv ** 3 / 3 + nv ** 2
  jac = C * (70 * nv + 60) + 10 * n_lim + 100 * G
  a_b = R_a * nv ** 2 + 2 * nv ** 2 + 2 * R_a * nv
  pg = K * (20 * nv * R_a + 20 * R_a)
  out = 2 * nv * R_a + 2 * nv ** 2 + 2 * nv + 60
  return tree + mass + jac + a_b + pg + out


def workload(n, steps, seed=0):
  from brax_amd import systems
  sys_ = systems.quadruped(0)
  rng = np.random.default_rng(seed)
  q0 = (sys_.init_q + rng.uniform(-0.1, 0.1, (n, sys_.q_size()))).astype(np.float32)
  qd0 = (0.1 * rng.standard_normal((n, sys_.qd_size()))).astype(np.float32)
  return sys_, q0, qd0, rng


def cpu_baseline(sys_, q0, qd0, rng_seed=1, n=256, steps=8):
  """Oracle (float32, torch CPU) on the first n envs for `steps` env steps:
  env-steps/s on this host, and the mean active constraint-row count."""
  import torch
  from oracle.generalized import Oracle
  threads = torch.get_num_threads()
  o = Oracle(sys_, dtype=torch.float32, inverse="exact")
  rng = np.random.default_rng(rng_seed)
  acts = rng.uniform(-1, 1, (steps, n, sys_.act_size()))
  st = o.init(q0[:n].astype(np.float64), qd0[:n].astype(np.float64))
  rows = []
  t0 = time.perf_counter()
  for s in range(steps):
    for _ in range(N_FRAMES):
      st = o.step(st, acts[s])
      rows.append((st.con_jac.abs().sum(-1) > 0).sum(-1).float().mean().item())
  dt = time.perf_counter() - t0
  return {
      "value": n * steps / dt, "unit": "env-steps/s", "cores": threads,
      "kind": "port",
      "sample": f"{n} quadruped envs x {steps} env steps ({N_FRAMES} substeps), "
                f"oracle/generalized.py float32 (torch CPU), {dt:.1f} s",
  }, float(np.mean(rows))


def main():
  ap = argparse.ArgumentParser()
  ap.add_argument("--gpus", type=int, default=1)
  ap.add_argument("--steps", type=int, default=1000)
  ap.add_argument("--warmup", type=int, default=50)
  ap.add_argument("--envs", type=int, default=65536)
  ap.add_argument("--no-graph", action="store_true")
  ap.add_argument("--no-cpu", action="store_true")
  args = ap.parse_args()
  if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
    print("bench.py measures one GPU: the step has no cross-GPU exchange; run "
          "with --gpus 1", file=sys.stderr)
    sys.exit(2)

  import torch
  from brax_amd import _lib
  from brax_amd.generalized import pipeline

  dev = torch.device("cuda", 0)
  torch.cuda.set_device(dev)
  n, K, W = args.envs, args.steps, args.warmup
  sys_, q0, qd0, rng = workload(n, K)
  m = pipeline.device_model(sys_, dev)
  lib = _lib.load()
  acts = torch.empty(max(K, W), n, sys_.act_size(), device=dev)
  g = torch.Generator(device=dev)
  g.manual_seed(0)
  acts.uniform_(-1.0, 1.0, generator=g)
  q_init = torch.tensor(q0, device=dev)
  qd_init = torch.tensor(qd0, device=dev)
  q = q_init.clone()
  qd = qd_init.clone()
  status = torch.zeros(n, dtype=torch.int32, device=dev)
  stream = torch.cuda.Stream(device=dev)

  def launch(n_steps):
    _lib.check(lib.bgx_rollout(
        m.handle, n, n_steps, N_FRAMES, _lib.ptr(q), _lib.ptr(qd),
        _lib.ptr(acts), n * sys_.act_size(), None, None, None, None, None,
        None, _lib.ptr(status), ctypes.c_void_p(stream.cuda_stream)))

  with torch.cuda.stream(stream):
    launch(W)  # warmup
  torch.cuda.synchronize()
  q.copy_(q_init)
  qd.copy_(qd_init)
  torch.cuda.synchronize()

  graph = None
  if not args.no_graph:
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph, stream=stream):
      launch(K)
    torch.cuda.synchronize()
    q.copy_(q_init)
    qd.copy_(qd_init)
  torch.cuda.synchronize()

  ev0 = torch.cuda.Event(enable_timing=True)
  ev1 = torch.cuda.Event(enable_timing=True)
  t0 = time.perf_counter()
  with torch.cuda.stream(stream):
    ev0.record(stream)
    if graph is not None:
      graph.replay()
    else:
      launch(K)
    ev1.record(stream)
  torch.cuda.synchronize()
  wall = time.perf_counter() - t0
  kernel_ms = ev0.elapsed_time(ev1)

  st = status.cpu().numpy()
  finite = bool(torch.isfinite(q).all() and torch.isfinite(qd).all())
  env_steps_per_s = n * K / wall

  cpu = None
  r_a = 9.9
  if not args.no_cpu:
    cpu, r_a = cpu_baseline(sys_, q0, qd0)
  nv, L = sys_.qd_size(), sys_.num_links()
  fl = flop_per_substep(L, nv, sys_.num_contacts(), sys_.num_limit_rows(),
                        len(sys_.geom_type), sys_.solver_iterations, r_a)
  flop_launch = fl * n * K * N_FRAMES
  achieved = flop_launch / (kernel_ms * 1e-3) / 1e12
  hbm_bytes = n * (4 * (sys_.q_size() + sys_.qd_size()) * 2 + 4 * sys_.act_size() * K)

  out = {
      "metric": "env-steps/s @65536 quadrupeds",
      "value": env_steps_per_s,
      "unit": "env-steps/s",
      "n_gpus": 1,
      "steps": K,
      "warmup": W,
      "ms_per_step": wall * 1e3 / K,
      "higher_is_better": True,
      "scaling": "weak",
      "vs_baseline": None,
      "dtype": "fp32",
      "data": "synthetic (seeded quadruped states and U(-1,1) actions in HBM)",
      "config": {"workload": "synthetic quadruped, generalized pipeline, "
                             "5 substeps x dt 0.01 per env step, one HIP-graph "
                             "rollout launch",
                 "envs": n, "env_steps": K, "substeps_per_env_step": N_FRAMES,
                 "qd_size": nv, "parallelism": "none (1 GPU)"},
      "roofline": {"bound": "valu", "achieved": achieved,
                   "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                   "frac": achieved / FP32_PEAK_TFLOPS, "traffic": None,
                   "flop_per_substep": fl, "mean_active_rows": r_a,
                   "kernel_ms": kernel_ms,
                   "hbm_algorithmic_gbs": hbm_bytes / (kernel_ms * 1e-3) / 1e9},
      "cpu_baseline": cpu,
      "slow_path_envs": int((st & 1).sum()),
      "finite": finite,
  }
  print(json.dumps(out))


if __name__ == "__main__":
  main()
