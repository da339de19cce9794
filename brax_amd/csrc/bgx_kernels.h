// Kernels of the generalized-coordinates step and its adjoint, and the
// per-topology launch table.  bgx_topo.hip instantiates them for one
// topology per object file; bgx.hip holds the C ABI and dispatches through
// the tables.
//
// Step kernels: one thread per environment, 64-thread workgroups (one
// wavefront each).  A launch advances every env by n_steps x n_frames
// substeps with the whole state in registers; HBM traffic is the initial
// (q, qd), one action per env step, optional per-step trajectory rows, and
// the final state.
#pragma once

#include <hip/hip_runtime.h>

#include <type_traits>

#include "bgx_adjoint.h"
#include "bgx_batch.h"
#include "bgx_launch.h"
#include "bgx_mtree.h"
#include "bgx_step.h"

namespace bgx {

// one adjoint substep: state (q, qd) of substep i, its action, the running
// cotangent adj (n, NQ + NV) of the substep's output state
struct AdjLaunch {
  const float* M;
  int n;
  const float* q;    // (n, NQ)
  const float* qd;   // (n, NV)
  const float* act;  // (n, NU) or null
  const float* gq_add;   // (n, NQ) cotangent added to adj's q part, or null
  const float* gqd_add;  // (n, NV)
  float* adj;        // (n, NQ + NV), read by stage A, written by stage B
  float* rec;        // (SIZE, n) stage A record
  float* gact;       // (n, NU) accumulated
  int epw;           // envs (lanes) per workgroup of A1/A2, 1..WAVE; 0 = WAVE
  // stage A1 tapes n_sub substeps sub_lo.. of a segment in one launch: the
  // start state of substep i at q + i n NQ (qd alike), its action at
  // act + (i / n_frames) act_stride, its tape at rec + (i - sub_lo) rec_stride
  int sub_lo, n_sub, n_frames;
  long long act_stride, rec_stride;
};

// Workgroups of fewer than WAVE lanes spread a small batch over more CUs.
// The LDS of a workgroup keeps the WAVE-lane stride whatever its lane count,
// so every lane computes the same instructions on the same values: results
// do not depend on the choice.
inline int block_lanes(int epw) { return epw > 0 && epw < WAVE ? epw : WAVE; }
inline dim3 env_grid(long long n, int b) { return dim3((unsigned)((n + b - 1) / b)); }

// the model table of env: shared, or one per env (PER_ENV topologies)
template <class T>
BGX_D const float* env_model(const float* M, int env) {
  return T::PER_ENV ? M + (long long)env * T::SIZE : M;
}

template <class T>
BGX_D void write_frames(const Launch& a, int env, const float* q, const float* qd) {
  V3 xp[T::L], xa[T::L], xv[T::L];
  Q4 xr[T::L];
  forward_kinematics<T, true>(env_model<T>(a.M, env), q, qd, xp, xr, xa, xv);
#pragma unroll
  for (int l = 0; l < T::L; ++l) {
    const long long o3 = ((long long)env * T::L + l) * 3;
    const long long o4 = ((long long)env * T::L + l) * 4;
    if (a.x_pos) {
      a.x_pos[o3] = xp[l].x, a.x_pos[o3 + 1] = xp[l].y, a.x_pos[o3 + 2] = xp[l].z;
    }
    if (a.x_rot) {
      a.x_rot[o4] = xr[l].w, a.x_rot[o4 + 1] = xr[l].x;
      a.x_rot[o4 + 2] = xr[l].y, a.x_rot[o4 + 3] = xr[l].z;
    }
    if (a.xd_ang) {
      a.xd_ang[o3] = xa[l].x, a.xd_ang[o3 + 1] = xa[l].y, a.xd_ang[o3 + 2] = xa[l].z;
    }
    if (a.xd_vel) {
      a.xd_vel[o3] = xv[l].x, a.xd_vel[o3 + 1] = xv[l].y, a.xd_vel[o3 + 2] = xv[l].z;
    }
  }
}

template <class T>
__global__ __launch_bounds__(WAVE) void k_init(Launch a) {
  const int env = blockIdx.x * blockDim.x + threadIdx.x;
  if (env >= a.n) return;
  float q[T::NQ], qd[T::NV];
#pragma unroll
  for (int k = 0; k < T::NQ; ++k) q[k] = a.q[(long long)env * T::NQ + k];
#pragma unroll
  for (int k = 0; k < T::NV; ++k) qd[k] = a.qd[(long long)env * T::NV + k];
  write_frames<T>(a, env, q, qd);
}

// every contact candidate of the state (the reference's debug contact
// output, brax/generalized/pipeline.py:58-59): MJX collisions at q, frames
// in the reference's orientation (normal from geom 1 to geom 2)
template <class T>
__global__ __launch_bounds__(WAVE) void k_contacts(Launch a) {
  const int env = blockIdx.x * blockDim.x + threadIdx.x;
  if (env >= a.n) return;
  float q[T::NQ], qd[T::NV];
#pragma unroll
  for (int k = 0; k < T::NQ; ++k) q[k] = a.q[(long long)env * T::NQ + k];
#pragma unroll
  for (int k = 0; k < T::NV; ++k) qd[k] = 0.f;
  const float* M = env_model<T>(a.M, env);
  V3 xp[T::L];
  Q4 xr[T::L];
  forward_kinematics<T, false, float>(M, q, qd, xp, xr, nullptr, nullptr);
#pragma unroll
  for (int c = 0; c < T::NCAND; ++c) {
    V3 ctr, axis, pos, n, t1, t2;
    float dist;
    if (T::PAIRS && T::CAND_KIND[c] >= KIND_SS) {
      pair_contact<T>(M, c, xp, xr, -1, dist, pos, n, t1, t2);
    } else {
      candidate_center<T>(M, c, xp, xr, ctr, axis);
      dist = candidate_dist<T>(M, c, ctr);
      contact_frame<T>(M, c, ctr, axis, dist, -1, pos, t1, t2);
      n = T::PLANE_Z ? v3(0.f, 0.f, 1.f) : ldv3(cand_rec<T>(M, c) + CD_PN);
    }
    const long long o = (long long)env * T::NCAND + c;
    a.c_dist[o] = dist;
    a.c_pos[3 * o] = pos.x, a.c_pos[3 * o + 1] = pos.y, a.c_pos[3 * o + 2] = pos.z;
    float* f = a.c_frame + 9 * o;
    f[0] = n.x, f[1] = n.y, f[2] = n.z;
    f[3] = t1.x, f[4] = t1.y, f[5] = t1.z;
    f[6] = t2.x, f[7] = t2.y, f[8] = t2.z;
  }
}

// The envs of one tree item, one per lane (a homogeneous launch's
// workgroup, or a batch item): n_steps x n_frames substeps with the state in
// registers.  The one call site of the step's substep in the rollout
// kernel.
template <class T>
BGX_D void tree_item(const Launch& a, const BatchItem& b, float* smem) {
  constexpr int NQ = T::NQ, NV = T::NV, NU = T::NU;
  const int i = threadIdx.x;
  if (i >= b.n) return;
  const long long env = b.env + i;
  // the stride is the topology's (compile time): a shared table stays one
  // address for the wavefront, read with scalar loads
  const float* M = a.M + b.m_off + (T::PER_ENV ? (long long)i * T::SIZE : 0);
  float* qp = a.q + b.q + (long long)i * NQ;
  float* qdp = a.qd + b.qd + (long long)i * NV;
  const float* actp = a.act ? a.act + b.act + (long long)i * NU : nullptr;
  // substep start states (the adjoint's recompute; homogeneous launches)
  const long long n_sub = (long long)a.n_steps * a.n_frames;
  const long long seg = blockIdx.y;
  float* subq = a.sub_q ? a.sub_q + (seg * n_sub * a.n + env) * NQ : nullptr;
  float* subqd = a.sub_qd ? a.sub_qd + (seg * n_sub * a.n + env) * NV : nullptr;
  float* lds = smem + threadIdx.x;
  float q[NQ], qd[NV];
#pragma unroll
  for (int k = 0; k < NQ; ++k) q[k] = qp[k];
#pragma unroll
  for (int k = 0; k < NV; ++k) qd[k] = qdp[k];
  float* probe = a.probe ? a.probe + env * (2 * NV + 3) : nullptr;
  int st = 0;
  for (int s = 0; s < a.n_steps; ++s) {
    float act[NU > 0 ? NU : 1];
#pragma unroll
    for (int k = 0; k < NU; ++k) act[k] = actp ? actp[s * a.act_stride + k] : 0.f;
    for (int f = 0; f < a.n_frames; ++f) {
      if (subq) {
        const long long o = ((long long)s * a.n_frames + f) * a.n;
#pragma unroll
        for (int k = 0; k < NQ; ++k) subq[o * NQ + k] = q[k];
#pragma unroll
        for (int k = 0; k < NV; ++k) subqd[o * NV + k] = qd[k];
      }
      st |= substep<T>(M, q, qd, act, lds, probe);
    }
    if (a.q_traj) {
#pragma unroll
      for (int k = 0; k < NQ; ++k) a.q_traj[s * a.q_step + b.q + (long long)i * NQ + k] = q[k];
    }
    if (a.qd_traj) {
#pragma unroll
      for (int k = 0; k < NV; ++k)
        a.qd_traj[s * a.qd_step + b.qd + (long long)i * NV + k] = qd[k];
    }
  }
  bool finite = true;
#pragma unroll
  for (int k = 0; k < NQ; ++k) {
    finite = finite && isfinite(q[k]);
    qp[k] = q[k];
  }
#pragma unroll
  for (int k = 0; k < NV; ++k) {
    finite = finite && isfinite(qd[k]);
    qdp[k] = qd[k];
  }
  if (!finite) st |= ST_NONFINITE;
  if (a.x_pos || a.x_rot || a.xd_ang || a.xd_vel) write_frames<T>(a, (int)env, q, qd);
  if (a.status) {
    if (a.n_seg > 1) atomicOr(a.status + env, st);
    else a.status[env] = a.status_or ? (a.status[env] | st) : st;
  }
}

template <class T>
constexpr bool is_chain_kernel() {
  return std::is_same<T, NoTree>::value;
}

// LDS of the rollout kernel: the tree step's, or a chain wavefront's, or
// the larger of the two where the kernel runs both
template <class T, bool CHAINS>
constexpr int rollout_lds_floats() {
  if constexpr (is_chain_kernel<T>()) {
    return CHAIN_LDS_FLOATS;
  } else if constexpr (CHAINS) {
    return lds_floats<T>() * WAVE > CHAIN_LDS_FLOATS ? lds_floats<T>() * WAVE
                                                      : CHAIN_LDS_FLOATS;
  } else {
    return lds_floats<T>() * WAVE;
  }
}

// the item of a homogeneous launch's workgroup: its envs (blockIdx.x) of
// checkpoint segment blockIdx.y
template <class T>
BGX_D BatchItem home_item(const Launch& a) {
  BatchItem b{};
  const long long seg = blockIdx.y;
  if constexpr (is_chain_kernel<T>()) {
    const int L = a.chain_L, size = chain_table_size(L);
    const int per = (int)blockDim.x / chain_lanes(L);
    const long long env0 = (long long)blockIdx.x * per;
    b.kind = BATCH_CHAIN, b.L = L;
    b.n = (int)(a.n - env0 < per ? a.n - env0 : per);
    b.env = env0;
    b.m_off = env0 * size, b.M_stride = size;
    b.q = env0 * L, b.qd = env0 * L, b.act = env0 * L;
  } else {
    const long long env0 = (long long)blockIdx.x * blockDim.x;
    b.kind = BATCH_TREE;
    b.n = (int)(a.n - env0 < (long long)blockDim.x ? a.n - env0 : blockDim.x);
    b.env = env0;
    b.m_off = T::PER_ENV ? env0 * T::SIZE : 0;
    b.M_stride = T::PER_ENV ? T::SIZE : 0;
    b.q = (seg * a.n + env0) * T::NQ;
    b.qd = (seg * a.n + env0) * T::NV;
    b.act = seg * a.seg_act_stride + env0 * T::NU;
  }
  return b;
}

// The rollout kernel.  Homogeneous launch (items null): workgroup
// (blockIdx.x, blockIdx.y) runs its own item once.  Batch launch: a
// persistent grid; workgroups take items from the counter until the table
// is exhausted.  T: a tree topology, or NoTree (chains only).  CHAINS: the
// kernel also runs chain items (the batch kernel of a tree topology); its
// tree path is the same single call site as the homogeneous kernel's, so
// both compile it to the same arithmetic.
template <class T, bool CHAINS>
__global__ __launch_bounds__(WAVE) void k_rollout(Launch a) {
  __shared__ float smem[rollout_lds_floats<T, CHAINS>()];
  const bool batch = a.items != nullptr;
  for (int round = 0;; ++round) {
    BatchItem b;
    if (batch) {
      const int it = next_item(a);
      if (it < 0) break;
      b = a.items[it];
    } else {
      if (round > 0) break;
      b = home_item<T>(a);
    }
    if constexpr (CHAINS) {
      if (b.kind == BATCH_CHAIN) {
        run_chain_item(a, b, smem);
        __syncthreads();  // the next item reuses the LDS
        continue;
      }
    }
    if constexpr (!is_chain_kernel<T>()) tree_item<T>(a, b, smem);
    __syncthreads();
  }
  if (batch) batch_done(a);
}

// cotangents of the frames of states (q, qd) (pipeline.init's x, xd:
// kinematics.forward, kinematics.py:31-108) carried to q and qd, as
// jax.grad differentiates the reference's frames; one thread per (env,
// input coordinate), coordinate-major so a wavefront shares its coordinate:
// the kinematics in dual numbers along that coordinate, contracted with the
// frames' cotangents
struct FramesVjp {
  const float* M;
  int n;
  const float* q;        // (n, NQ)
  const float* qd;       // (n, NV)
  const float* gx_pos;   // (n, L, 3) or null (zero)
  const float* gx_rot;   // (n, L, 4) or null
  const float* gxd_ang;  // (n, L, 3) or null
  const float* gxd_vel;  // (n, L, 3) or null
  float* gq;             // (n, NQ)
  float* gqd;            // (n, NV)
};

template <class T>
__global__ __launch_bounds__(WAVE) void k_frames_vjp(FramesVjp a) {
  constexpr int NQ = T::NQ, NV = T::NV, ND = NQ + NV, L = T::L;
  const int n_pad = (a.n + WAVE - 1) / WAVE * WAVE;
  const long long g = (long long)blockIdx.x * WAVE + threadIdx.x;
  const int dir = (int)(g / n_pad), env = (int)(g % n_pad);
  if (dir >= ND || env >= a.n) return;
  Dual q[NQ], qd[NV];
#pragma unroll
  for (int k = 0; k < NQ; ++k) q[k] = dual(a.q[(long long)env * NQ + k], k == dir ? 1.f : 0.f);
#pragma unroll
  for (int k = 0; k < NV; ++k)
    qd[k] = dual(a.qd[(long long)env * NV + k], NQ + k == dir ? 1.f : 0.f);
  V3T<Dual> xp[L], xa[L], xv[L];
  Q4T<Dual> xr[L];
  forward_kinematics<T, true, Dual>(env_model<T>(a.M, env), q, qd, xp, xr, xa, xv);
  float phi = 0.f;
  auto dot3 = [](const float* c, V3T<Dual> v) {
    return c[0] * v.x.d + c[1] * v.y.d + c[2] * v.z.d;
  };
#pragma unroll
  for (int l = 0; l < L; ++l) {
    const long long o = (long long)env * L + l;
    if (a.gx_pos) phi += dot3(a.gx_pos + 3 * o, xp[l]);
    if (a.gx_rot) {
      const float* c = a.gx_rot + 4 * o;
      phi += c[0] * xr[l].w.d + c[1] * xr[l].x.d + c[2] * xr[l].y.d + c[3] * xr[l].z.d;
    }
    if (a.gxd_ang) phi += dot3(a.gxd_ang + 3 * o, xa[l]);
    if (a.gxd_vel) phi += dot3(a.gxd_vel + 3 * o, xv[l]);
  }
  if (dir < NQ) a.gq[(long long)env * NQ + dir] = phi;
  else a.gqd[(long long)env * NV + dir - NQ] = phi;
}

// adjoint stage A1: the substep with a slot per candidate, recording its
// tape; blockIdx.y picks the substep of the batch
template <class T>
__global__ __launch_bounds__(WAVE) void k_adj_fwd(AdjLaunch a) {
  __shared__ float smem[lds_floats<T>() * WAVE];
  constexpr int NQ = T::NQ, NV = T::NV, KS = TapeLayout<T>::KS;
  const int env = blockIdx.x * blockDim.x + threadIdx.x;
  if (env >= a.n) return;
  const int i = a.sub_lo + (int)blockIdx.y;
  const float* qs = a.q + (long long)i * a.n * NQ;
  const float* qds = a.qd + (long long)i * a.n * NV;
  const float* as = a.act ? a.act + (long long)(i / a.n_frames) * a.act_stride : nullptr;
  float q[NQ], qd[NV], act[T::NU > 0 ? T::NU : 1];
#pragma unroll
  for (int k = 0; k < NQ; ++k) q[k] = qs[(long long)env * NQ + k];
#pragma unroll
  for (int k = 0; k < NV; ++k) qd[k] = qds[(long long)env * NV + k];
#pragma unroll
  for (int k = 0; k < T::NU; ++k) act[k] = as ? as[(long long)env * T::NU + k] : 0.f;
  // the step's own order: the fast path (KC slots in LDS) records the tape
  // when the active contacts fit, else the path with a slot per candidate
  // records it (the tape's layout has a slot per candidate either way)
  const RecPtr tape{a.rec + (long long)blockIdx.y * a.rec_stride + env, a.n};
  float* lds = smem + threadIdx.x;
  bool done = false;
  if constexpr (T::KC < KS) {
    LdsRows<T::KC> lrows{lds + T::L * 6 * WAVE};
    done = substep_ks<T, T::KC>(a.M, q, qd, act, lds, lrows, nullptr, tape);
  }
  if (!done) {
    PrivRows<KS> rows;
    substep_ks<T, KS>(launder(a.M), q, qd, act, lds, rows, nullptr, tape);
  }
}

// adjoint stage A2: the reverse sweep from the tape; its row vectors live in
// LDS, rev_lanes<T>() lanes per workgroup, rev_pair<T>() lanes per env
template <class T>
__global__ __launch_bounds__(WAVE) void k_adj_rev(AdjLaunch a) {
  constexpr int NQ = T::NQ, NV = T::NV, AL = rev_lanes<T>(), P = rev_pair<T>();
  __shared__ float smem[rev_lds_floats<T>(AL)];
  const long long lane = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const int env = (int)(lane / P), sub = (int)(lane % P);
  if (env >= a.n) return;
  float q[NQ], qd[NV], gq1[NQ], gqd1[NV];
  float act[T::NU > 0 ? T::NU : 1], gact[T::NU > 0 ? T::NU : 1];
  const float* adj = a.adj + (long long)env * (NQ + NV);
#pragma unroll
  for (int k = 0; k < NQ; ++k) {
    q[k] = a.q[(long long)env * NQ + k];
    gq1[k] = adj[k] + (a.gq_add ? a.gq_add[(long long)env * NQ + k] : 0.f);
  }
#pragma unroll
  for (int k = 0; k < NV; ++k) {
    qd[k] = a.qd[(long long)env * NV + k];
    gqd1[k] = adj[NQ + k] + (a.gqd_add ? a.gqd_add[(long long)env * NV + k] : 0.f);
  }
#pragma unroll
  for (int k = 0; k < T::NU; ++k) {
    act[k] = a.act ? a.act[(long long)env * T::NU + k] : 0.f;
    gact[k] = 0.f;
  }
  adjoint_rev<T, AL, P>(a.M, q, qd, act, gq1, gqd1, RecPtr{a.rec + env, a.n},
                        smem + threadIdx.x / P,
                        smem + RevLds<T>::SHARED * (AL / P) + threadIdx.x, gact, sub);
  if (sub == 0) {
#pragma unroll
    for (int k = 0; k < T::NU; ++k) a.gact[(long long)env * T::NU + k] += gact[k];
  }
}

// adjoint stage B: one thread per (env, q coordinate), coordinate-major over
// env blocks of WAVE (a wave shares its coordinate; tape reads coalesce),
// and after them one thread per env for all qd coordinates (one reverse
// pass)
template <class T>
__global__ __launch_bounds__(WAVE) void k_adj_b(AdjLaunch a) {
  constexpr int NQ = T::NQ, NV = T::NV, ND = NQ + NV;
  const int n_pad = (a.n + WAVE - 1) / WAVE * WAVE;
  const long long g = (long long)blockIdx.x * WAVE + threadIdx.x;
  const int dir = (int)(g / n_pad), env = (int)(g % n_pad);
  if (dir > NQ || env >= a.n) return;
  float act[T::NU > 0 ? T::NU : 1];
#pragma unroll
  for (int k = 0; k < T::NU; ++k) act[k] = a.act ? a.act[(long long)env * T::NU + k] : 0.f;
  const float* q = a.q + (long long)env * NQ;
  const float* qd = a.qd + (long long)env * NV;
  float* out = a.adj + (long long)env * ND;
  if (dir < NQ) {
    out[dir] = adjoint_b<T>(a.M, q, qd, act, RecPtr{a.rec + env, a.n}, dir);
  } else {
    float gqd[NV];
    adjoint_b_qd_all<T>(a.M, q, qd, act, RecPtr{a.rec + env, a.n}, gqd);
#pragma unroll
    for (int j = 0; j < NV; ++j) out[NQ + j] = gqd[j];
  }
}

// a batch launch of the rollout kernel: a persistent grid of `grid`
// one-wavefront workgroups (a.items set)
template <class T>
hipError_t launch_batch(const Launch& a, int grid, hipStream_t s) {
  hipLaunchKernelGGL((k_rollout<T, true>), dim3((unsigned)grid), dim3(WAVE), 0, s, a);
  return hipGetLastError();
}

// resident workgroups of the batch kernel per compute unit
template <class T>
int batch_blocks_per_cu() {
  int nb = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_rollout<T, true>, WAVE, 0) !=
      hipSuccess)
    return 1;
  return nb > 0 ? nb : 1;
}

// A tree topology with K fast contact slots in place of its default KC: the
// LDS tile of the step kernel (6 L + 16 K floats per lane).  Every template
// of the step reads the slot count as T::KC, so Tile<T, K> instantiates the
// same code with the other tile.
template <class T, int K>
struct Tile : T {
  static constexpr int KC = K;
};

// per-topology host entry points

struct TopoOps {
  const char* name;
  const char* sig;
  int dims[8];  // L, NQ, NV, NU, NLIM, NCAND, KC, table size
  int adj_rec;  // floats per env of the adjoint record
  int per_env;  // one model table per env
  hipError_t (*init)(const Launch&, hipStream_t);
  hipError_t (*rollout)(const Launch&, hipStream_t);
  hipError_t (*adj_fwd)(const AdjLaunch&, hipStream_t);
  hipError_t (*adj_rev)(const AdjLaunch&, hipStream_t);
  hipError_t (*adj_b)(const AdjLaunch&, hipStream_t);
  hipError_t (*contacts)(const Launch&, hipStream_t);
  hipError_t (*frames_vjp)(const FramesVjp&, hipStream_t);
  int chain;     // a hinge-chain topology (bgx_chain.h kernels)
  int mlanes;    // lanes per env of the multi-lane tree kernel (0: not compiled)
  int tiles[3];  // fast contact slots the step kernels are compiled for (0: none)
  // the persistent batch kernel that runs this topology's envs beside
  // chains (null: the topology cannot join a heterogeneous batch), and its
  // resident workgroups per compute unit
  hipError_t (*batch)(const Launch&, int grid, hipStream_t);
  int (*batch_occupancy)();
};

template <class T>
struct Ops {
  static hipError_t init(const Launch& a, hipStream_t s) {
    const int b = block_lanes(a.epw);
    hipLaunchKernelGGL(k_init<T>, env_grid(a.n, b), dim3(b), 0, s, a);
    return hipGetLastError();
  }
  template <class TT>
  static hipError_t rollout_tile(const Launch& a, hipStream_t s) {
    const int b = block_lanes(a.epw);
    const dim3 g = env_grid(a.n, b);
    Launch h = a;
    h.items = nullptr;
    hipLaunchKernelGGL((k_rollout<TT, false>), dim3(g.x, a.n_seg > 1 ? (unsigned)a.n_seg : 1u),
                       dim3(b), 0, s, h);
    return hipGetLastError();
  }
  static hipError_t rollout(const Launch& a, hipStream_t s) {
    if constexpr (T::MLANE) {
      static_assert(mtree_ok<T>(), "topology outside the multi-lane tree kernel's family");
      if (a.lanes != 1) {
        constexpr int E = MTab<T>::E;
        hipLaunchKernelGGL(k_mtree<T>,
                           dim3((unsigned)((a.n + E - 1) / E), a.n_seg > 1 ? (unsigned)a.n_seg : 1u),
                           dim3(WAVE), 0, s, a);
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess || !(a.x_pos || a.x_rot || a.xd_ang || a.xd_vel)) return e;
        return init(a, s);  // frames of the final state
      }
    }
    if constexpr (T::TILE1 > 0) {
      if (a.kc == T::TILE1) return rollout_tile<Tile<T, T::TILE1>>(a, s);
    }
    if constexpr (T::TILE2 > 0) {
      if (a.kc == T::TILE2) return rollout_tile<Tile<T, T::TILE2>>(a, s);
    }
    return rollout_tile<T>(a, s);
  }
  static hipError_t adj_fwd(const AdjLaunch& a, hipStream_t s) {
    const int b = block_lanes(a.epw);
    const dim3 g = env_grid(a.n, b);
    hipLaunchKernelGGL(k_adj_fwd<T>, dim3(g.x, (unsigned)a.n_sub), dim3(b), 0, s, a);
    return hipGetLastError();
  }
  static hipError_t adj_rev(const AdjLaunch& a, hipStream_t s) {
    // rev_lanes<T>() lanes per workgroup: the LDS layout's strides
    const int b = rev_lanes<T>();
    hipLaunchKernelGGL(k_adj_rev<T>, env_grid((long long)a.n * rev_pair<T>(), b), dim3(b), 0, s,
                       a);
    return hipGetLastError();
  }
  static hipError_t adj_b(const AdjLaunch& a, hipStream_t s) {
    // NQ q coordinates, then the qd reverse pass
    const long long threads = (long long)(a.n + WAVE - 1) / WAVE * WAVE * (T::NQ + 1);
    hipLaunchKernelGGL(k_adj_b<T>, dim3((unsigned)((threads + WAVE - 1) / WAVE)), dim3(WAVE),
                       0, s, a);
    return hipGetLastError();
  }
  static hipError_t frames_vjp(const FramesVjp& a, hipStream_t s) {
    const long long threads = (long long)(a.n + WAVE - 1) / WAVE * WAVE * (T::NQ + T::NV);
    hipLaunchKernelGGL(k_frames_vjp<T>, dim3((unsigned)((threads + WAVE - 1) / WAVE)),
                       dim3(WAVE), 0, s, a);
    return hipGetLastError();
  }
  static hipError_t contacts(const Launch& a, hipStream_t s) {
    const int b = block_lanes(a.epw);
    hipLaunchKernelGGL(k_contacts<T>, env_grid(a.n, b), dim3(b), 0, s, a);
    return hipGetLastError();
  }
  static TopoOps table() {
    TopoOps t{T::NAME,
              T::SIG,
              {T::L, T::NQ, T::NV, T::NU, T::NLIM, T::NCAND, T::KC, T::SIZE},
              AdjRec<T>::SIZE,
              T::PER_ENV,
              &init,
              &rollout,
              nullptr,
              nullptr,
              nullptr,
              nullptr};
    if constexpr (T::NCAND > 0) t.contacts = &contacts;
    t.frames_vjp = &frames_vjp;
    t.mlanes = T::MLANE ? MTab<T>::G : 0;
    t.tiles[0] = T::KC, t.tiles[1] = T::TILE1, t.tiles[2] = T::TILE2;
    if constexpr (T::BATCH) t.batch = &launch_batch<T>, t.batch_occupancy = &batch_blocks_per_cu<T>;
    // the adjoint is compiled for topologies with a shared model table
    if constexpr (!T::PER_ENV) t.adj_fwd = &adj_fwd, t.adj_rev = &adj_rev, t.adj_b = &adj_b;
    return t;
  }
};

}  // namespace bgx
