// Heterogeneous batches in one persistent launch: the chain items.
//
// A batch (bgx_batch_create) holds envs of several models: chains (any
// length up to 32 links, per-env model tables) and envs of one tree topology
// T (shared or per-env tables).  Its envs live in packed buffers: model m's
// envs are a contiguous range, each env's q, qd and action rows at fixed
// strides.  The host cuts the envs into work items (64 tree envs, one per
// lane; or 64 / G chains of G lanes each), ordered longest first, and
// uploads the item table.  The batch rollout kernel (k_rollout<T, true> in
// bgx_kernels.h; k_rollout<NoTree, true> for chains alone) then runs a grid of as
// many workgroups as fit on the chip at once: each takes the next item from
// a device counter until none is left and runs that item's envs through
// all the rollout's substeps.  The last workgroup to leave resets the
// counter, so the launch can be replayed (HIP graphs) without a memset.
//
// Homogeneous launches run k_rollout<T, false> (k_rollout<NoTree, true> for
// chains), forming one item per workgroup from blockIdx.  Both instances
// reach an env's substeps through the one call site (tree_item, chain_item),
// so an env's arithmetic is the same compiled code whether it runs alone, in
// a batch of its kind, or in a mixed batch: results are bitwise equal (the
// GPU tests compare them).
#pragma once

#include <hip/hip_runtime.h>

#include "bgx_chain.h"
#include "bgx_launch.h"
#include "bgx_step.h"

namespace bgx {

// no tree topology: the kernel of chain launches (homogeneous chain
// batches, and heterogeneous batches of chains only)
struct NoTree {
  static constexpr int BATCH = 1;
};

// Runs the chains of one item (64 / G per wavefront; lane wl of the
// wavefront is link wl % G of env wl / G of the item) through n_steps x
// n_frames substeps.  An env past the item's count computes on the item's
// last env and writes nothing.  Status: BGX nonfinite bit per env; probe
// (homogeneous launches): qf_smooth and qdd of the last substep.
//
// Not inlined: in the batch kernel of a tree topology the tree step's
// register pressure otherwise reaches into the chain loops (spill reloads
// in the scans, a mixed batch 50% slower than its two halves run apart); as
// a function of its own the item gets its own register allocation, once
// per item, and every kernel calls the same code.
#ifdef BGX_HOST_CHECK
#define BGX_ITEM inline
#else
#define BGX_ITEM __device__ __attribute__((noinline))
#endif
template <int G>
BGX_ITEM void chain_item(const Launch& a, const BatchItem& b, float* lds) {
  const int wl = threadIdx.x % WAVE;
  const int l = wl & (G - 1), gb = wl - l, e = wl / G;
  const int L = b.L;
  const bool live = e < b.n;
  const int ec = live ? e : b.n - 1;
  const float* M = a.M + b.m_off + (long long)ec * b.M_stride;
  const ChainLink c = chain_link(M, L, l);
  const V3 grav = ldv3(M + H_GRAVITY);
  const float dt = M[H_DT];
  const bool dof = l < L;
  float* qp = a.q + b.q + (long long)ec * L;
  float* qdp = a.qd + b.qd + (long long)ec * L;
  float q = dof ? qp[l] : 0.f;
  float qd = dof ? qdp[l] : 0.f;
  float* probe = (live && a.probe) ? a.probe + (b.env + e) * (2 * L + 3) : nullptr;
  for (int s = 0; s < a.n_steps; ++s) {
    const float* ar = a.act ? a.act + s * a.act_stride + b.act + (long long)ec * L : nullptr;
    const float act = (dof && ar) ? ar[l] : 0.f;
    for (int f = 0; f < a.n_frames; ++f) {
      const bool last = probe && dof && s == a.n_steps - 1 && f == a.n_frames - 1;
      chain_substep<G>(c, grav, dt, l, wl, gb, q, qd, act, lds, last ? probe + l : nullptr,
                       last ? probe + L + l : nullptr);
    }
    if (live && dof) {
      if (a.q_traj) a.q_traj[s * a.q_step + b.q + (long long)e * L + l] = q;
      if (a.qd_traj) a.qd_traj[s * a.qd_step + b.qd + (long long)e * L + l] = qd;
    }
  }
  if (live && dof) qp[l] = q, qdp[l] = qd;
  const bool bad = dof && !(isfinite(q) && isfinite(qd));
  const unsigned long long mask = __ballot(bad);
  const unsigned long long gm = (G >= 64 ? ~0ull : ((1ull << G) - 1ull)) << gb;
  if (live && l == 0) {
    if (a.status) a.status[b.env + e] = (mask & gm) ? ST_NONFINITE : 0;
    if (probe) probe[2 * L] = probe[2 * L + 1] = probe[2 * L + 2] = 0.f;
  }
}

// a chain item by its lane count
BGX_D void run_chain_item(const Launch& a, const BatchItem& b, float* lds) {
  const int L = b.L;
  if (L <= 2) chain_item<2>(a, b, lds);
  else if (L <= 4) chain_item<4>(a, b, lds);
  else if (L <= 8) chain_item<8>(a, b, lds);
  else if (L <= 16) chain_item<16>(a, b, lds);
  else chain_item<32>(a, b, lds);
}

// the next item of a batch launch (-1 when the table is exhausted)
BGX_D int next_item(const Launch& a) {
  int it = 0;
  if (threadIdx.x == 0) it = atomicAdd(a.counter, 1);
  it = __builtin_amdgcn_readfirstlane(__shfl(it, 0));
  return it < a.n_items ? it : -1;
}

// at the end of a batch launch: every workgroup is past its last fetch once
// `done` reaches the grid size; the last one resets both counters
BGX_D void batch_done(const Launch& a) {
  if (threadIdx.x == 0 && atomicAdd(a.counter + 1, 1) == (int)gridDim.x - 1) {
    atomicExch(a.counter, 0);
    atomicExch(a.counter + 1, 0);
  }
}

}  // namespace bgx
