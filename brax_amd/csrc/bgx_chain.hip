// Kernels and launch tables of the hinge-chain topologies (bgx_chain.h: G
// lanes per env).  One object file for every chain length: their rollouts
// run the chain kernel k_rollout<NoTree, true> (the link count a kernel argument),
// which also runs heterogeneous batches of chains alone; their frames a
// kernel per lane count.  bgx.hip reaches each chain topology's table
// through bgx_ops_<struct>().
#include <hip/hip_runtime.h>

#include "bgx_kernels.h"
#include "topologies.inc"

namespace bgx {
namespace {

// pipeline.init of a chain batch: world frames and velocities
template <int G>
__global__ __launch_bounds__(WAVE) void k_chain_init(Launch a, int L) {
  const int wl = threadIdx.x, l = wl & (G - 1), e = wl / G;
  const long long env = (long long)blockIdx.x * (blockDim.x / G) + e;
  const bool live = env < a.n;
  const long long ec = live ? env : a.n - 1;
  const float* M = a.M + ec * chain_table_size(L);
  const ChainLink c = chain_link(M, L, l);
  const bool dof = l < L;
  const float q = dof ? a.q[ec * L + l] : 0.f;
  const float qd = dof ? a.qd[ec * L + l] : 0.f;
  V3 xp, xa, xv;
  Q4 xr;
  chain_frames_xd<G>(l, c, q, qd, xp, xr, xa, xv);
  if (!(live && dof)) return;
  const long long o = env * L + l;
  if (a.x_pos) a.x_pos[3 * o] = xp.x, a.x_pos[3 * o + 1] = xp.y, a.x_pos[3 * o + 2] = xp.z;
  if (a.x_rot) {
    a.x_rot[4 * o] = xr.w, a.x_rot[4 * o + 1] = xr.x;
    a.x_rot[4 * o + 2] = xr.y, a.x_rot[4 * o + 3] = xr.z;
  }
  if (a.xd_ang) a.xd_ang[3 * o] = xa.x, a.xd_ang[3 * o + 1] = xa.y, a.xd_ang[3 * o + 2] = xa.z;
  if (a.xd_vel) a.xd_vel[3 * o] = xv.x, a.xd_vel[3 * o + 1] = xv.y, a.xd_vel[3 * o + 2] = xv.z;
}

// envs per workgroup of a chain launch: the setting (envs_per_workgroup),
// at most one wavefront's worth
template <int G>
int chain_envs_per_wg(int epw) {
  const int full = WAVE / G;
  const int e = block_lanes(epw);
  return e < full ? e : full;
}

template <int L>
hipError_t chain_init(const Launch& a, hipStream_t s) {
  constexpr int G = chain_lanes(L);
  const int per = chain_envs_per_wg<G>(a.epw);
  hipLaunchKernelGGL(k_chain_init<G>, dim3((unsigned)((a.n + per - 1) / per)),
                     dim3((unsigned)(per * G)), 0, s, a, L);
  return hipGetLastError();
}

// a homogeneous chain batch: the chain kernel (k_rollout<NoTree, true>), one item
// of per envs per workgroup
template <int L>
hipError_t chain_rollout(const Launch& a, hipStream_t s) {
  constexpr int G = chain_lanes(L);
  // chains have no adjoint: the recompute's fields are never set for them
  if (a.n_seg > 1 || a.sub_q || a.sub_qd) return hipErrorInvalidValue;
  const int per = chain_envs_per_wg<G>(a.epw);
  Launch h = a;
  h.items = nullptr, h.chain_L = L;
  hipLaunchKernelGGL((k_rollout<NoTree, true>), dim3((unsigned)((a.n + per - 1) / per)),
                     dim3((unsigned)(per * G)), 0, s, h);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  // frames at the new state (the rollout updated q, qd in place)
  if (a.x_pos || a.x_rot || a.xd_ang || a.xd_vel) return chain_init<L>(a, s);
  return hipSuccess;
}

template <class T>
TopoOps chain_table() {
  static_assert(T::CHAIN, "not a chain topology");
  static_assert(T::L >= 1 && T::L <= CHAIN_MAX, "chain length");
  static_assert(T::SIZE == chain_table_size(T::L), "chain model table layout");
  TopoOps t{T::NAME,
            T::SIG,
            {T::L, T::NQ, T::NV, T::NU, T::NLIM, T::NCAND, T::KC, T::SIZE},
            0,
            T::PER_ENV,
            &chain_init<T::L>,
            &chain_rollout<T::L>,
            nullptr,
            nullptr,
            nullptr,
            nullptr};
  t.chain = 1;
  t.tiles[0] = T::KC;
  t.batch = &launch_batch<NoTree>;
  t.batch_occupancy = &batch_blocks_per_cu<NoTree>;
  return t;
}

}  // namespace
}  // namespace bgx

#define BGX_CHAIN_OPS(ID, T)                                      \
  const bgx::TopoOps* bgx_ops_##T() {                             \
    static const bgx::TopoOps ops = bgx::chain_table<T>();        \
    return &ops;                                                  \
  }
BGX_FOR_EACH_CHAIN_TOPOLOGY(BGX_CHAIN_OPS)
#undef BGX_CHAIN_OPS
