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
CPU baseline (the oracle restatement in float32 on this host's CPU share, on
the timed region's own inputs, all envs when they take a few seconds, else a
leading sample); the mean active constraint-row count of the roofline's FLOP
model comes from the same CPU run.  The line also carries ``fwd_adjoint``:
the differentiable 128-step rollout of 4096 quadrupeds (checkpoint every 16
env steps, the apg workload below) with its own roofline and CPU baseline
(the torch oracle with autograd); --no-adjoint leaves it out.

Other workloads (one JSON line each, same keys):
  humanoid  16384 synthetic humanoids (23 dofs, hinge triples), 1000 env
            steps of 5 substeps (dt 0.003), actions U(-0.4, 0.4);
  hetero    32768 envs in one persistent kernel launch (captured in a HIP
            graph): 50% hinge chains of 2-32 links with per-env model tables
            (drawn from 8 seeded chain models per length), 50% quadrupeds;
            500 env steps;
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
# the headline metric: forward env-steps/s at 65536 quadrupeds (value), the
# fwd+adjoint rate (fwd_adjoint.value) and the roofline fractions
METRIC = "env-steps/s @65536 quadrupeds; fwd+adjoint steps/s; % HBM/VALU roofline"


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


def profile_path(kernel, name, n):
  """The committed rocprofv3 summary of `kernel` on workload `name` at `n` envs."""
  return os.path.join("profiles", f"{kernel}_{name}_{n}.md")


def measured_traffic(kernel, name, n, steps, rel=None):
  """HBM bytes of one rollout launch of `steps` env steps of `n` envs, from
  the committed rocprofv3 --pmc summary of that same command (FETCH_SIZE +
  WRITE_SIZE of the rollout kernel, calibrated with a copy kernel), or None
  when no profile of this kernel, workload, batch and step count is
  committed (rel: the profile's path, if not the kernel's default)."""
  path = os.path.join(ROOT, rel or profile_path(kernel, name, n))
  if not os.path.exists(path):
    return None
  for line in open(path):
    if line.startswith("traffic_per_launch:"):
      f = dict(kv.split("=") for kv in line.split(":", 1)[1].split())
      if int(f["envs"]) == n and int(f["env_steps"]) == steps:
        return float(f["bytes"])
  return None


def host_cpu():
  """Threads the CPU baselines use (the host share this process may run on:
  OMP_NUM_THREADS when set, else the affinity mask) and a description of the
  host."""
  aff = len(os.sched_getaffinity(0))
  threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or aff
  model = "unknown"
  try:
    for line in open("/proc/cpuinfo"):
      if line.startswith("model name"):
        model = line.split(":", 1)[1].strip()
        break
  except OSError:
    pass
  return threads, {"nproc": os.cpu_count(), "affinity_cpus": aff, "cpu_model": model}


def cpu_baseline(sys_, q0, qd0, acts, n_frames=N_FRAMES):
  """The oracle's C++ restatement (oracle/cstep.cpp, float32, OpenMP over
  envs on the host's CPU share) on the given inputs: acts (K, n_s, nu) are
  the timed region's actions of its first n_s envs.  Returns the CPU
  baseline object and the mean active constraint-row count per substep
  over these inputs."""
  from oracle.cstep import CStep
  threads, host = host_cpu()
  K, n = acts.shape[0], acts.shape[1]
  cs = CStep(sys_)
  cs.rollout(q0[:4], qd0[:4], acts[:1, :4], n_frames, threads=1)  # load, warm
  t0 = time.perf_counter()
  _, _, rows = cs.rollout(q0[:n], qd0[:n], acts, n_frames, threads=threads, rows=True)
  dt = time.perf_counter() - t0
  return {
      "value": n * K / dt, "unit": "env-steps/s", "cores": threads,
      "kind": "port",
      "sample": f"the timed region's inputs for its first {n} {sys_.name} envs x {K} env steps "
                f"({n_frames} substeps), oracle/cstep.cpp float32, OpenMP {threads} threads, "
                f"{dt:.1f} s",
      "host": host,
  }, rows


def timed_sample(n, K, budget=2_600_000):
  """Envs of the CPU baseline's sample: all n when n x K env steps fit the
  budget (a few seconds of the C++ oracle on 16 cores), else fewer."""
  return max(1, min(n, budget // max(K, 1)))


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
  epw, kc = 64, None
  if args.lanes:
    from brax_amd.generalized import autotune
    autotune.set_launch(m, autotune.LANES, args.lanes)
  lanes = lib.bgx_model_get_launch(m.handle, 3)
  kernel = "k_mtree" if lanes > 1 else "k_rollout"
  if lanes > 1:
    # the multi-lane kernel's workgroup is one wavefront of 64 / lanes envs,
    # and it keeps one contact row per candidate lane (no LDS tile)
    epw = 64 // lanes
  elif not args.no_tune:
    from brax_amd.generalized import autotune
    epw = autotune.tune_step(sys_, q_init, qd_init, acts[:2], N_FRAMES)
    # the tile trades occupancy against slow-path substeps, which grow as
    # the envs settle onto the floor: time it over the timed steps (up to 20)
    kc = autotune.tune_tile(sys_, q_init, qd_init, acts[:min(K, 20)], N_FRAMES)
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
  r_a = None
  if not args.no_cpu:
    # the CPU baseline runs the timed region's own inputs (a leading sample
    # of its envs when the whole batch would take too long) and gives the
    # mean active constraint-row count over the timed steps
    n_s = timed_sample(n, K)
    cpu, r_a = cpu_baseline(sys_, q0[:n_s], qd0[:n_s], acts[:K, :n_s].cpu().numpy())
  nv, L = sys_.qd_size(), sys_.num_links()
  r_a_src = "oracle/cstep.cpp float32 over the timed inputs" if r_a is not None else \
      "not measured (--no-cpu): 9.9, the 65536-env mean of earlier runs"
  if r_a is None:
    r_a = 9.9
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
                 "lanes_per_env": lanes,
                 "fast_contact_slots": (lib.bgx_model_get_launch(m.handle, 2)
                                        if lanes == 1 else None),
                 "parallelism": "none (1 GPU)"},
      "roofline": {"bound": "valu", "achieved": achieved,
                   "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                   "frac": achieved / FP32_PEAK_TFLOPS,
                   "traffic": measured_traffic(kernel, name, n, K),
                   "traffic_source": f"{profile_path(kernel, name, n)} (rocprofv3 --pmc "
                                     "FETCH_SIZE, WRITE_SIZE of this command)",
                   "flop_per_substep": fl, "mean_active_rows": r_a,
                   "mean_active_rows_source": r_a_src,
                   "kernel_ms": kernel_ms,
                   "hbm_algorithmic_gbs": hbm_bytes / (kernel_ms * 1e-3) / 1e9},
      "cpu_baseline": cpu,
      "slow_path_envs": int((st & 1).sum()),
      "finite": finite,
      "nonfinite_envs": nonfinite_envs,
  }


def adjoint_cpu_baseline(sys_, q0, qd0, acts, n_frames=N_FRAMES):
  """fwd+adjoint on the CPU: the torch oracle (oracle/generalized.py,
  float32, the exact inverse the kernel uses) with torch autograd through
  the unrolled solver, as the reference differentiates with jax.grad
  (brax/training/agents/apg/train.py:188-198), on the given sample of the
  workload's inputs, same loss."""
  import torch
  from oracle.generalized import Oracle
  threads, host = host_cpu()
  torch.set_num_threads(threads)
  K, n = acts.shape[0], acts.shape[1]
  o = Oracle(sys_, torch.float32, inverse="exact")
  t0 = time.perf_counter()
  a = torch.tensor(acts, dtype=torch.float32, requires_grad=True)
  qd_init = torch.tensor(qd0, dtype=torch.float32, requires_grad=True)
  st = o.init(torch.tensor(q0, dtype=torch.float32), qd_init)
  loss = 1e-3 * (a * a).sum()
  for t in range(K):
    st = o.env_step(st, a[t], n_frames)
    loss = loss - st.qd[:, 0].sum()
  loss.backward()
  dt = time.perf_counter() - t0
  return {
      "value": n * K / dt, "unit": "env-steps/s", "cores": threads, "kind": "port",
      "sample": f"{n} quadruped envs x {K} env steps of the workload's inputs, loss and "
                f"gradients (actions, qd0), torch float32 oracle with autograd, "
                f"{threads} threads, {dt:.1f} s",
      "host": host,
  }


def fwd_adjoint_leg(n, T, ckpt, warmup, tune=True, cpu=True, parts=0):
  """Forward + adjoint of a T-step rollout of n quadrupeds (loss
  -sum_t v_torso,x + 1e-3 sum |act|^2, gradient w.r.t. every action and the
  initial qd, checkpoint every ckpt env steps), timed with HIP events on the
  stream the kernels run on, beside the forward-only rollout of the same
  inputs.

  FLOP model of one fwd+adjoint env step: 4 F, F the forward op count of
  flop_per_substep times the substeps (the forward pass, and 3 F for the
  gradient: the reverse-mode bound of Griewank and Walther's cheap gradient
  principle); the checkpoint recompute is not counted."""
  import torch
  from brax_amd import _lib
  from brax_amd.generalized import pipeline

  dev = torch.device("cuda", 0)
  torch.cuda.set_device(dev)
  sys_, q0, qd0, rng = workload("apg", n)
  acts_np = rng.uniform(-1, 1, (T, n, sys_.act_size())).astype(np.float32)
  acts0 = torch.tensor(acts_np, device=dev)
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

  from brax_amd.generalized import autotune
  m = pipeline.device_model(sys_, dev)
  autotune.set_launch(m, autotune.ADJOINT_PARTS, parts)
  epw = [64, 64]
  if tune:
    epw[0] = autotune.tune_step(sys_, q0t, qd0t, acts0[:2], N_FRAMES)
    # the backward pass over one checkpoint segment: a shorter window times
    # the launch overheads rather than the kernels
    epw[1] = autotune.tune_adjoint(sys_, q0t, qd0t, acts0[:ckpt], N_FRAMES)
  fwd_bwd()
  for _ in range(max(1, warmup // 50)):
    fwd_bwd()
    fwd_only()
  torch.cuda.synchronize()

  def timed(fn, reps=3):
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ev0.record()
    for _ in range(reps):
      out = fn()
    ev1.record()
    torch.cuda.synchronize()
    return out, (time.perf_counter() - t0) / reps, ev0.elapsed_time(ev1) / reps

  _, t_fwd, ms_fwd = timed(fwd_only)
  (ga, gqd), t_fb, ms_fb = timed(fwd_bwd)
  finite = bool(torch.isfinite(ga).all() and torch.isfinite(gqd).all())
  fwd_rate, fb_rate = n * T / t_fwd, n * T / t_fb
  out = {
      "value": fb_rate, "unit": "env-steps/s", "envs": n, "horizon": T, "ckpt_every": ckpt,
      "ms_per_step": t_fb * 1e3 / T,
      "forward_value": fwd_rate,
      "fwd_adjoint_over_forward": fb_rate / fwd_rate,
      "envs_per_workgroup": {"step": epw[0], "adjoint": epw[1]},
      # env parts of the backward pass, one stream each (0: one)
      "adjoint_parts": parts or 1,
      "finite": finite,
  }
  cpu_b, r_a = None, None
  if cpu:
    # active rows over this workload's own forward inputs (all n envs, T steps)
    from oracle.cstep import CStep
    threads, _ = host_cpu()
    _, _, r_a = CStep(sys_).rollout(q0, qd0, acts_np, N_FRAMES, threads=threads, rows=True)
    m = min(n, 4096)
    cpu_b = adjoint_cpu_baseline(sys_, q0[:m], qd0[:m], acts_np[:2, :m])
  r_a_src = "oracle/cstep.cpp float32 over this workload's inputs"
  if r_a is None:
    r_a, r_a_src = 9.9, "not measured (--no-cpu): 9.9"
  fl = flop_per_substep(sys_.num_links(), sys_.qd_size(), sys_.num_contacts(),
                        sys_.num_limit_rows(), len(sys_.geom_type), sys_.solver_iterations, r_a)
  achieved = 4 * fl * N_FRAMES * n * T / (ms_fb * 1e-3) / 1e12
  out["roofline"] = {
      "bound": "valu", "achieved": achieved, "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
      "frac": achieved / FP32_PEAK_TFLOPS, "traffic": None,
      "flop_model": "4 x forward op count per env step (forward + reverse-mode bound 3F; "
                    "recompute not counted)",
      "flop_per_substep_forward": fl, "mean_active_rows": r_a,
      "mean_active_rows_source": r_a_src, "gpu_ms_fwd_adjoint": ms_fb,
      "gpu_ms_forward": ms_fwd,
  }
  out["cpu_baseline"] = cpu_b
  return out


def apg_bench(args):
  """Forward + adjoint of a 128-step rollout of 4096 quadrupeds."""
  from brax_amd import _lib
  n = args.envs or WORKLOADS["apg"][1]
  try:
    leg = fwd_adjoint_leg(n, args.horizon, args.ckpt, args.warmup, tune=not args.no_tune,
                          cpu=not args.no_cpu, parts=args.adj_parts)
  except _lib.BgxError as e:
    print(f"apg workload unavailable: {e}", file=sys.stderr)
    sys.exit(3)
  return {
      "metric": f"fwd+adjoint env-steps/s @{n} quadrupeds, horizon {args.horizon}",
      "value": leg["value"],
      "unit": "env-steps/s",
      "n_gpus": 1,
      "steps": args.horizon,
      "warmup": args.warmup,
      "ms_per_step": leg["ms_per_step"],
      "higher_is_better": True,
      "scaling": "weak",
      "vs_baseline": None,
      "dtype": "fp32",
      "data": "synthetic (seeded quadruped states and U(-1,1) actions)",
      "config": {"workload": "differentiable quadruped rollout, loss -sum v_torso,x + "
                             "1e-3 sum |act|^2, gradient w.r.t. actions and qd0, "
                             f"checkpoint every {args.ckpt} env steps",
                 "envs": n, "horizon": args.horizon, "ckpt_every": args.ckpt,
                 "envs_per_workgroup": leg["envs_per_workgroup"],
                 "parallelism": "none (1 GPU)"},
      "forward_value": leg["forward_value"],
      "fwd_adjoint_over_forward": leg["fwd_adjoint_over_forward"],
      "roofline": leg["roofline"],
      "cpu_baseline": leg["cpu_baseline"],
      "finite": leg["finite"],
  }


def hetero_cpu_baseline(envs, hb, q0, qd0, acts, K, n_frames=N_FRAMES, chunk=64):
  """The oracle's C++ restatement on a bounded sample of the mixed batch's
  timed inputs: its first n_s envs for the K timed env steps, each env
  stepped with its own System.  Envs of one System run together in chunks
  of up to `chunk` envs, the chunks on a pool of the host's threads (one
  OpenMP thread per chunk: most Systems of the mix have a few envs each).
  Returns the baseline object and the quadrupeds' mean active constraint
  rows per substep over the sample."""
  from concurrent.futures import ThreadPoolExecutor
  from oracle.cstep import CStep
  threads, host = host_cpu()
  n_s = timed_sample(len(envs), K)
  q0n, qd0n, an = q0.cpu().numpy(), qd0.cpu().numpy(), acts[:K].cpu().numpy()
  by_sys = {}
  for e in range(n_s):
    by_sys.setdefault(id(envs[e]), []).append(e)
  tasks = []
  for ids in by_sys.values():
    sys_ = envs[ids[0]]
    cs = CStep(sys_)
    for c in range(0, len(ids), chunk):
      rows = [hb.env_rows(e) for e in ids[c:c + chunk]]
      q = np.stack([q0n[r[0]] for r in rows])
      qd = np.stack([qd0n[r[1]] for r in rows])
      a = np.stack([an[:, r[2]] for r in rows], 1)
      tasks.append((cs, q, qd, a, sys_.name == "quadruped"))
  tasks.sort(key=lambda t: -t[1].shape[0] * t[1].shape[1])  # longest first
  cs0, q, qd, a, _ = tasks[0]
  cs0.rollout(q[:1], qd[:1], a[:1, :1], n_frames, threads=1)  # load, warm

  def run(t):
    cs, q, qd, a, quad = t
    out = cs.rollout(q, qd, a, n_frames, threads=1, rows=quad)
    return (out[2] * q.shape[0], q.shape[0]) if quad else (0.0, 0)

  t0 = time.perf_counter()
  with ThreadPoolExecutor(max_workers=threads) as pool:
    res = list(pool.map(run, tasks))
  dt = time.perf_counter() - t0
  n_quad = sum(r[1] for r in res)
  rows = sum(r[0] for r in res) / n_quad if n_quad else None
  return {
      "value": n_s * K / dt, "unit": "env-steps/s", "cores": threads, "kind": "port",
      "sample": f"the timed region's inputs for its first {n_s} envs ({n_quad} quadrupeds, "
                f"{n_s - n_quad} chains) x {K} env steps ({n_frames} substeps), "
                f"oracle/cstep.cpp float32, each env with its own System, {len(tasks)} "
                f"chunks on {threads} threads, {dt:.1f} s",
      "host": host,
  }, rows


def hetero_bench(args):
  """Mixed batch: half chains (2-32 links, per-env tables), half quadrupeds,
  in one persistent kernel launch."""
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
  q0 = hb.pack([(s.init_q + rng.uniform(-0.1, 0.1, s.q_size())).astype(np.float32)
                for s in envs], "q")
  qd0 = hb.pack([(0.1 * rng.standard_normal(s.qd_size())).astype(np.float32) for s in envs],
                "qd")
  acts = torch.empty(max(K, W), hb.act_floats, device=dev)
  g = torch.Generator(device=dev)
  g.manual_seed(0)
  acts.uniform_(-1.0, 1.0, generator=g)
  q, qd = q0.clone(), qd0.clone()
  status = torch.zeros(n, dtype=torch.int32, device=dev)
  side = torch.cuda.Stream(device=dev)
  with torch.cuda.stream(side):
    hb.launch_rollout(q, qd, acts, W, N_FRAMES)  # warmup
  torch.cuda.synchronize()
  graph = torch.cuda.CUDAGraph()
  with torch.cuda.graph(graph, stream=side):
    hb.launch_rollout(q, qd, acts, K, N_FRAMES, status)
  torch.cuda.synchronize()
  q.copy_(q0)
  qd.copy_(qd0)
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
  if getattr(args, "calib_copy", False):
    src = torch.ones(64 << 20, device=dev)
    dst = torch.empty_like(src)
    torch.add(src, 0.0, out=dst)  # reads 256 MiB, writes 256 MiB
    torch.cuda.synchronize()
    del src, dst
  nonfinite_envs = int(((status & 2) != 0).sum())
  n_quad = n - n_chain
  cpu, r_a = (None, None)
  if not args.no_cpu:
    cpu, r_a = hetero_cpu_baseline(envs, hb, q0, qd0, acts, K)
  r_a_src = ("oracle/cstep.cpp float32 over the CPU baseline's quadrupeds" if r_a is not None
             else "not measured (--no-cpu): 9.9, the 65536-env mean of earlier runs")
  if r_a is None:
    r_a = 9.9
  fl = n_quad * flop_per_substep(quad.num_links(), quad.qd_size(), quad.num_contacts(),
                                 quad.num_limit_rows(), len(quad.geom_type),
                                 quad.solver_iterations, r_a)
  for L in lengths:
    fl += flop_per_substep(int(L), int(L), 0, 0, int(L), 4, 0.0)
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
      "config": {"workload": "heterogeneous batch: one persistent kernel launch (bgx_batch_rollout) "
                             "over a device work table, captured in a HIP graph",
                 "envs": n, "env_steps": K, "models": len(hb.groups),
                 "work_items": hb.n_items, "workgroups": hb.grid,
                 "parallelism": "none (1 GPU)"},
      "roofline": {"bound": "valu", "achieved": achieved, "peak": FP32_PEAK_TFLOPS,
                   "unit": "TFLOP/s", "frac": achieved / FP32_PEAK_TFLOPS,
                   "traffic": measured_traffic("k_rollout", "hetero", n, K,
                                               rel=f"profiles/hetero_{n}.md"),
                   "traffic_source": f"profiles/hetero_{n}.md (rocprofv3 --pmc FETCH_SIZE, "
                                     "WRITE_SIZE of this command)",
                   "kernel_ms": kernel_ms,
                   "note": "op-count model: quadrupeds at R_a active rows, chains without rows",
                   "mean_active_rows": r_a, "mean_active_rows_source": r_a_src},
      "cpu_baseline": cpu,
      "finite": nonfinite_envs == 0,
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
  ap.add_argument("--lanes", type=int, default=0,
                  help="lanes per env of the step kernels (1: one thread per env; 0: the "
                       "topology's default)")
  ap.add_argument("--adj-parts", type=int, default=0,
                  help="env parts (streams) of the adjoint's backward pass, 1-4 (0: automatic)")
  ap.add_argument("--no-graph", action="store_true")
  ap.add_argument("--no-cpu", action="store_true")
  ap.add_argument("--no-adjoint", action="store_true",
                  help="omit the fwd+adjoint part of the default quadruped line")
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
    if args.workload == "quadruped" and not args.no_adjoint:
      # the metric's second part: fwd+adjoint of the 4096-quadruped,
      # 128-step differentiable rollout with a checkpoint every 16 env steps
      out["fwd_adjoint"] = fwd_adjoint_leg(4096, 128, 16, args.warmup,
                                           tune=not args.no_tune, cpu=not args.no_cpu,
                                           parts=args.adj_parts)
      if out["config"]["envs"] == 65536:
        out["metric"] = METRIC
  print(json.dumps(out))


if __name__ == "__main__":
  main()
