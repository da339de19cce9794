"""Throughput of the generalized-coordinates step on one MI355X.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--envs N]
                  [--workload quadruped|humanoid|apg]

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

Other workloads (one JSON line each, same keys):
  humanoid  16384 synthetic humanoids (23 dofs, hinge triples), 1000 env
            steps of 5 substeps (dt 0.003), actions U(-0.4, 0.4);
  hetero    32768 envs in one HIP graph: 50% hinge chains of 2-32 links
            with per-env model tables (drawn from 8 seeded chain models per
            length), 50% quadrupeds; 500 env steps;
  apg       differentiable rollout: 4096 quadrupeds, horizon 128, loss
            -sum_t v_torso,x + 1e-3 sum |act|^2, gradient w.r.t. every action
            and the initial qd by the adjoint with a checkpoint every 16 env
            steps; value = (forward + backward) env-steps/s, reported beside
            the forward-only rate at the same size.
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


WORKLOADS = {
    # name: (system builder, default envs, q noise, qd scale, action range)
    "quadruped": ("quadruped", 65536, 0.1, 0.1, 1.0),
    "humanoid": ("humanoid", 16384, 0.01, 0.01, 0.4),
    "apg": ("quadruped", 4096, 0.1, 0.1, 1.0),
    "hetero": ("quadruped", 32768, 0.1, 0.1, 1.0),
}
DEFAULT_STEPS = {"quadruped": 1000, "humanoid": 1000, "apg": 128, "hetero": 500}


def workload(name, n, seed=0):
  from brax_amd import systems
  builder, _, qn, qds, _ = WORKLOADS[name]
  sys_ = getattr(systems, builder)(0)
  rng = np.random.default_rng(seed)
  q0 = (sys_.init_q + rng.uniform(-qn, qn, (n, sys_.q_size()))).astype(np.float32)
  if name == "humanoid":
    qd0 = rng.uniform(-qds, qds, (n, sys_.qd_size())).astype(np.float32)
  else:
    qd0 = (qds * rng.standard_normal((n, sys_.qd_size()))).astype(np.float32)
  return sys_, q0, qd0, rng


def measured_traffic(name, n, steps):
  """HBM bytes per launch of `steps` env steps from the committed rocprofv3
  summary of this workload (FETCH_SIZE + WRITE_SIZE, calibrated), or None."""
  path = os.path.join(ROOT, "profiles", f"k_rollout_{name}_{n}.md")
  if not os.path.exists(path):
    return None
  for line in open(path):
    if line.startswith("traffic_bytes_per_env_step = "):
      return float(line.split("=")[1]) * steps
  return None


def cpu_baseline(sys_, q0, qd0, act_range, rng_seed=1, n=4096, steps=100):
  """The oracle's C++ restatement (oracle/cstep.cpp, float32, OpenMP over
  envs on up to 16 host threads) on the first n envs for `steps` env steps:
  env-steps/s on this host, and the mean active constraint-row count."""
  import os
  from oracle.cstep import CStep
  threads = min(os.cpu_count() or 1, 16)
  n = min(n, q0.shape[0])
  cs = CStep(sys_)
  rng = np.random.default_rng(rng_seed)
  acts = rng.uniform(-act_range, act_range, (steps, n, sys_.act_size()))
  cs.rollout(q0[:4], qd0[:4], acts[:1, :4], N_FRAMES, threads=1)  # load, warm
  t0 = time.perf_counter()
  _, _, rows = cs.rollout(q0[:n], qd0[:n], acts, N_FRAMES, threads=threads, rows=True)
  dt = time.perf_counter() - t0
  return {
      "value": n * steps / dt, "unit": "env-steps/s", "cores": threads,
      "kind": "port",
      "sample": f"{n} {sys_.name} envs x {steps} env steps ({N_FRAMES} substeps), "
                f"oracle/cstep.cpp float32, OpenMP {threads} threads, {dt:.1f} s",
  }, rows


def forward_bench(args, name):
  import torch
  from brax_amd import _lib
  from brax_amd.generalized import pipeline

  dev = torch.device("cuda", 0)
  torch.cuda.set_device(dev)
  n, K, W = args.envs or WORKLOADS[name][1], args.steps, args.warmup
  act_range = WORKLOADS[name][4]
  sys_, q0, qd0, _ = workload(name, n)
  m = pipeline.device_model(sys_, dev)
  lib = _lib.load()
  nu = sys_.act_size()
  acts = torch.empty(max(K, W), n, nu, device=dev)
  g = torch.Generator(device=dev)
  g.manual_seed(0)
  acts.uniform_(-act_range, act_range, generator=g)
  q_init = torch.tensor(q0, device=dev)
  qd_init = torch.tensor(qd0, device=dev)
  epw = 64
  if not args.no_tune:
    from brax_amd.generalized import autotune
    epw = autotune.tune_step(sys_, q_init, qd_init, acts[:2], N_FRAMES)
  q = q_init.clone()
  qd = qd_init.clone()
  status = torch.zeros(n, dtype=torch.int32, device=dev)
  stream = torch.cuda.Stream(device=dev)

  def launch(n_steps):
    _lib.check(lib.bgx_rollout(
        m.handle, n, n_steps, N_FRAMES, _lib.ptr(q), _lib.ptr(qd),
        _lib.ptr(acts), n * nu, None, None, None, None, None,
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

  if getattr(args, "calib_copy", False):
    src = torch.ones(64 << 20, device=dev)
    dst = torch.empty_like(src)
    torch.add(src, 0.0, out=dst)  # reads 256 MiB, writes 256 MiB
    torch.cuda.synchronize()
    del src, dst

  st = status.cpu().numpy()
  fin_env = torch.isfinite(q).all(1) & torch.isfinite(qd).all(1)
  finite = bool(fin_env.all())
  nonfinite_envs = int((~fin_env).sum())
  env_steps_per_s = n * K / wall

  cpu = None
  r_a = 9.9
  if not args.no_cpu:
    cpu, r_a = cpu_baseline(sys_, q0, qd0, act_range,
                            n=4096 if name == "quadruped" else 2048,
                            steps=200 if name == "quadruped" else 100)
  nv, L = sys_.qd_size(), sys_.num_links()
  fl = flop_per_substep(L, nv, sys_.num_contacts(), sys_.num_limit_rows(),
                        len(sys_.geom_type), sys_.solver_iterations, r_a)
  flop_launch = fl * n * K * N_FRAMES
  achieved = flop_launch / (kernel_ms * 1e-3) / 1e12
  hbm_bytes = n * (4 * (sys_.q_size() + sys_.qd_size()) * 2 + 4 * nu * K)
  return {
      "metric": f"env-steps/s @{n} {sys_.name}s",
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
      "data": f"synthetic (seeded {sys_.name} states and U(-{act_range:g},{act_range:g}) "
              "actions in HBM)",
      "config": {"workload": f"synthetic {sys_.name}, generalized pipeline, "
                             f"{N_FRAMES} substeps x dt {sys_.opt.timestep:g} per env "
                             "step, one HIP-graph rollout launch",
                 "envs": n, "env_steps": K, "substeps_per_env_step": N_FRAMES,
                 "qd_size": nv, "envs_per_workgroup": epw,
                 "parallelism": "none (1 GPU)"},
      "roofline": {"bound": "valu", "achieved": achieved,
                   "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                   "frac": achieved / FP32_PEAK_TFLOPS,
                   "traffic": measured_traffic(name, n, K),
                   "flop_per_substep": fl, "mean_active_rows": r_a,
                   "kernel_ms": kernel_ms,
                   "hbm_algorithmic_gbs": hbm_bytes / (kernel_ms * 1e-3) / 1e9},
      "cpu_baseline": cpu,
      "slow_path_envs": int((st & 1).sum()),
      "finite": finite,
      "nonfinite_envs": nonfinite_envs,
  }


def apg_bench(args):
  """Forward + adjoint of a 128-step rollout of 4096 quadrupeds."""
  import torch
  from brax_amd.generalized import pipeline

  dev = torch.device("cuda", 0)
  torch.cuda.set_device(dev)
  n = args.envs or WORKLOADS["apg"][1]
  T, ckpt = args.horizon, args.ckpt
  sys_, q0, qd0, rng = workload("apg", n)
  acts0 = torch.tensor(rng.uniform(-1, 1, (T, n, sys_.act_size())).astype(np.float32),
                       device=dev)
  q0t = torch.tensor(q0, device=dev)
  qd0t = torch.tensor(qd0, device=dev)

  def fwd_bwd():
    acts = acts0.clone().requires_grad_(True)
    qd_init = qd0t.clone().requires_grad_(True)
    qt, qdt, _ = pipeline.rollout_grad(sys_, q0t, qd_init, acts, N_FRAMES, ckpt)
    loss = -qdt[:, :, 0].sum() + 1e-3 * (acts * acts).sum()
    loss.backward()
    return acts.grad, qd_init.grad

  def fwd_only():
    st = pipeline.State(q=q0t, qd=qd0t, x=None, xd=None)
    return pipeline.rollout(sys_, st, acts0, N_FRAMES, trajectory=True, frames=False)

  from brax_amd import _lib
  epw = [64, 64]
  if not args.no_tune:
    from brax_amd.generalized import autotune
    epw[0] = autotune.tune_step(sys_, q0t, qd0t, acts0[:2], N_FRAMES)
    try:
      epw[1] = autotune.tune_adjoint(sys_, q0t, qd0t, acts0[:2], N_FRAMES)
    except _lib.BgxError as e:
      print(f"apg workload unavailable: {e}", file=sys.stderr)
      sys.exit(3)
  try:
    fwd_bwd()
  except _lib.BgxError as e:
    print(f"apg workload unavailable: {e}", file=sys.stderr)
    sys.exit(3)
  for _ in range(max(1, args.warmup // 50)):
    fwd_bwd()
    fwd_only()
  torch.cuda.synchronize()
  reps = 3
  t0 = time.perf_counter()
  for _ in range(reps):
    fwd_only()
  torch.cuda.synchronize()
  t_fwd = (time.perf_counter() - t0) / reps
  t0 = time.perf_counter()
  for _ in range(reps):
    ga, gqd = fwd_bwd()
  torch.cuda.synchronize()
  t_fb = (time.perf_counter() - t0) / reps
  finite = bool(torch.isfinite(ga).all() and torch.isfinite(gqd).all())
  fwd_rate, fb_rate = n * T / t_fwd, n * T / t_fb
  return {
      "metric": f"fwd+adjoint env-steps/s @{n} quadrupeds, horizon {T}",
      "value": fb_rate,
      "unit": "env-steps/s",
      "n_gpus": 1,
      "steps": T,
      "warmup": args.warmup,
      "ms_per_step": t_fb * 1e3 / T,
      "higher_is_better": True,
      "scaling": "weak",
      "vs_baseline": None,
      "dtype": "fp32",
      "data": "synthetic (seeded quadruped states and U(-1,1) actions)",
      "config": {"workload": "differentiable quadruped rollout, loss -sum v_torso,x + "
                             "1e-3 sum |act|^2, gradient w.r.t. actions and qd0, "
                             f"checkpoint every {ckpt} env steps",
                 "envs": n, "horizon": T, "ckpt_every": ckpt,
                 "envs_per_workgroup": {"step": epw[0], "adjoint": epw[1]},
                 "parallelism": "none (1 GPU)"},
      "forward_value": fwd_rate,
      "fwd_adjoint_over_forward": fb_rate / fwd_rate,
      "roofline": None,
      "cpu_baseline": None,
      "finite": finite,
  }


def hetero_bench(args):
  """Mixed batch: half chains (2-32 links, per-env tables), half quadrupeds."""
  import torch
  from brax_amd import systems
  from brax_amd.generalized import hetero

  dev = torch.device("cuda", 0)
  torch.cuda.set_device(dev)
  n = args.envs or WORKLOADS["hetero"][1]
  K, W = args.steps, args.warmup
  rng = np.random.default_rng(0)
  quad = systems.quadruped(0)
  pool = {L: [systems.chain(L, 1000 * L + j) for j in range(8)] for L in range(2, 33)}
  n_chain = n // 2
  lengths = rng.integers(2, 33, n_chain)
  envs = [quad] * (n - n_chain) + [pool[int(L)][int(rng.integers(8))] for L in lengths]
  envs = [envs[i] for i in rng.permutation(n)]
  hb = hetero.HeteroBatch(envs, dev)
  q0, qd0, acts = [], [], []
  for g in hb.groups:
    s0 = envs[g.envs[0]]
    q0.append(torch.tensor((s0.init_q + rng.uniform(-0.1, 0.1, (g.n, s0.q_size())))
                           .astype(np.float32), device=dev))
    qd0.append(torch.tensor((0.1 * rng.standard_normal((g.n, s0.qd_size())))
                            .astype(np.float32), device=dev))
    a = torch.empty(max(K, W), g.n, s0.act_size(), device=dev)
    a.uniform_(-1.0, 1.0)
    acts.append(a)
  q = [x.clone() for x in q0]
  qd = [x.clone() for x in qd0]
  side = torch.cuda.Stream(device=dev)
  epw = None
  if not args.no_tune:
    epw = hb.autotune(q0, qd0, acts, N_FRAMES)

  def reset():
    for a, b in zip(q, q0):
      a.copy_(b)
    for a, b in zip(qd, qd0):
      a.copy_(b)

  with torch.cuda.stream(side):
    hb.launch_rollout(q, qd, acts, W, N_FRAMES)
  torch.cuda.synchronize()
  graph = torch.cuda.CUDAGraph()
  with torch.cuda.graph(graph, stream=side):
    hb.launch_rollout(q, qd, acts, K, N_FRAMES)
  torch.cuda.synchronize()
  reset()
  torch.cuda.synchronize()
  ev0 = torch.cuda.Event(enable_timing=True)
  ev1 = torch.cuda.Event(enable_timing=True)
  t0 = time.perf_counter()
  with torch.cuda.stream(side):
    ev0.record(side)
    graph.replay()
    ev1.record(side)
  torch.cuda.synchronize()
  wall = time.perf_counter() - t0
  kernel_ms = ev0.elapsed_time(ev1)
  nonfinite_envs = sum(int((~(torch.isfinite(a).all(1) & torch.isfinite(b).all(1))).sum())
                       for a, b in zip(q, qd))
  finite = nonfinite_envs == 0
  fl = 0.0
  for g in hb.groups:
    s0 = envs[g.envs[0]]
    fl += g.n * flop_per_substep(s0.num_links(), s0.qd_size(), s0.num_contacts(),
                                 s0.num_limit_rows(), len(s0.geom_type),
                                 s0.solver_iterations, 9.9 if s0 is quad else 0.0)
  achieved = fl * K * N_FRAMES / (kernel_ms * 1e-3) / 1e12
  return {
      "metric": f"env-steps/s @{n} mixed envs (50% chains of 2-32 links, 50% quadrupeds)",
      "value": n * K / wall,
      "unit": "env-steps/s",
      "n_gpus": 1,
      "steps": K,
      "warmup": W,
      "ms_per_step": wall * 1e3 / K,
      "higher_is_better": True,
      "scaling": "weak",
      "vs_baseline": None,
      "dtype": "fp32",
      "data": "synthetic (seeded chain and quadruped states, U(-1,1) actions in HBM)",
      "config": {"workload": "heterogeneous batch: one HIP graph of per-topology "
                             f"rollout kernels on {len(hb.streams())} streams",
                 "hw_queues": os.environ.get("GPU_MAX_HW_QUEUES", "4"), "envs": n,
                 "env_steps": K, "groups": len(hb.groups),
                 "envs_per_workgroup": None if epw is None else
                 [[envs[g.envs[0]].name, g.n, e] for g, e in zip(hb.groups, epw)],
                 "parallelism": "none (1 GPU)"},
      "roofline": {"bound": "valu", "achieved": achieved, "peak": FP32_PEAK_TFLOPS,
                   "unit": "TFLOP/s", "frac": achieved / FP32_PEAK_TFLOPS,
                   "traffic": None, "kernel_ms": kernel_ms,
                   "note": "quadruped rows at the 65536-env mean 9.9"},
      "cpu_baseline": None,
      "finite": finite,
      "nonfinite_envs": nonfinite_envs,
  }


def main():
  ap = argparse.ArgumentParser()
  ap.add_argument("--gpus", type=int, default=1)
  ap.add_argument("--steps", type=int, default=0)
  ap.add_argument("--warmup", type=int, default=50)
  ap.add_argument("--envs", type=int, default=0)
  ap.add_argument("--workload", default="quadruped", choices=sorted(WORKLOADS))
  ap.add_argument("--ckpt", type=int, default=16)
  ap.add_argument("--no-graph", action="store_true")
  ap.add_argument("--no-cpu", action="store_true")
  ap.add_argument("--no-tune", action="store_true",
                  help="64 envs per workgroup instead of the autotuner's choice")
  ap.add_argument("--calib-copy", action="store_true",
                  help="run a 256 MiB elementwise copy before the timed region "
                       "(a known-size kernel for scaling HBM counters)")
  args = ap.parse_args()
  if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
    print("bench.py measures one GPU: the step has no cross-GPU exchange; run "
          "with --gpus 1", file=sys.stderr)
    sys.exit(2)
  args.steps = args.steps or DEFAULT_STEPS[args.workload]
  if args.workload == "apg":
    args.horizon = args.steps
    out = apg_bench(args)
  elif args.workload == "hetero":
    out = hetero_bench(args)
  else:
    out = forward_bench(args, args.workload)
  print(json.dumps(out))


if __name__ == "__main__":
  main()
