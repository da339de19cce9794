// Kernels and C-ABI entry points of the generalized-coordinates step.
//
// One thread per environment, 64-thread workgroups (one wavefront each).  A
// launch advances every env by n_steps x n_frames substeps with the whole
// state in registers; HBM traffic is the initial (q, qd), one action per env
// step, optional per-step trajectory rows, and the final state.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>

#include "../../include/brax_gen.h"
#include "bgx_step.h"
#include "topologies.inc"

namespace bgx {

struct Launch {
  const float* M;
  int n;
  int n_steps;
  int n_frames;
  float* q;
  float* qd;
  const float* act;
  long long act_stride;
  float* q_traj;
  float* qd_traj;
  float* x_pos;
  float* x_rot;
  float* xd_ang;
  float* xd_vel;
  int* status;
  float* probe;
};

template <class T>
BGX_D void write_frames(const Launch& a, int env, const float* q,
                        const float* qd) {
  V3 xp[T::L], xa[T::L], xv[T::L];
  Q4 xr[T::L];
  forward_kinematics<T, true>(a.M, q, qd, xp, xr, xa, xv);
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
  const int env = blockIdx.x * WAVE + threadIdx.x;
  if (env >= a.n) return;
  float q[T::NQ], qd[T::NV];
#pragma unroll
  for (int k = 0; k < T::NQ; ++k) q[k] = a.q[(long long)env * T::NQ + k];
#pragma unroll
  for (int k = 0; k < T::NV; ++k) qd[k] = a.qd[(long long)env * T::NV + k];
  write_frames<T>(a, env, q, qd);
}

template <class T>
__global__ __launch_bounds__(WAVE) void k_rollout(Launch a) {
  // link scratch (6 per link), then the fast-path slot rows (5 vectors x 4
  // rows), which also hold a second link vector while the rows are unused
  constexpr int kRowsArea =
      RV_N * 4 * T::KC > T::L * 6 ? RV_N * 4 * T::KC : T::L * 6;
  __shared__ float smem[(T::L * 6 + kRowsArea) * WAVE];
  const int env = blockIdx.x * WAVE + threadIdx.x;
  if (env >= a.n) return;
  float* lds = smem + threadIdx.x;
  float q[T::NQ], qd[T::NV];
#pragma unroll
  for (int k = 0; k < T::NQ; ++k) q[k] = a.q[(long long)env * T::NQ + k];
#pragma unroll
  for (int k = 0; k < T::NV; ++k) qd[k] = a.qd[(long long)env * T::NV + k];
  float* probe = a.probe ? a.probe + (long long)env * (2 * T::NV + 3) : nullptr;
  int st = 0;
  for (int s = 0; s < a.n_steps; ++s) {
    float act[T::NU > 0 ? T::NU : 1];
#pragma unroll
    for (int k = 0; k < T::NU; ++k)
      act[k] = a.act ? a.act[s * a.act_stride + (long long)env * T::NU + k] : 0.f;
    for (int f = 0; f < a.n_frames; ++f)
      st |= substep<T>(a.M, q, qd, act, lds, probe);
    if (a.q_traj) {
#pragma unroll
      for (int k = 0; k < T::NQ; ++k)
        a.q_traj[((long long)s * a.n + env) * T::NQ + k] = q[k];
    }
    if (a.qd_traj) {
#pragma unroll
      for (int k = 0; k < T::NV; ++k)
        a.qd_traj[((long long)s * a.n + env) * T::NV + k] = qd[k];
    }
  }
  bool finite = true;
#pragma unroll
  for (int k = 0; k < T::NQ; ++k) {
    finite = finite && isfinite(q[k]);
    a.q[(long long)env * T::NQ + k] = q[k];
  }
#pragma unroll
  for (int k = 0; k < T::NV; ++k) {
    finite = finite && isfinite(qd[k]);
    a.qd[(long long)env * T::NV + k] = qd[k];
  }
  if (!finite) st |= ST_NONFINITE;
  if (a.x_pos || a.x_rot || a.xd_ang || a.xd_vel) write_frames<T>(a, env, q, qd);
  if (a.status) a.status[env] = st;
}

}  // namespace bgx

// ---------------------------------------------------------------------------
// C ABI

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

struct TopoInfo {
  const char* name;
  const char* sig;
  int dims[8];
};

#define BGX_TOPO_INFO(ID, T)                                                 \
  TopoInfo{T::NAME, T::SIG,                                                  \
           {T::L, T::NQ, T::NV, T::NU, T::NLIM, T::NCAND, T::KC, T::SIZE}},
const TopoInfo kTopos[] = {BGX_FOR_EACH_TOPOLOGY(BGX_TOPO_INFO)};
#undef BGX_TOPO_INFO

int check_hip(hipError_t e, const char* what) {
  if (e == hipSuccess) return BGX_OK;
  return fail(BGX_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

}  // namespace

struct bgx_model {
  int topo;
  float* table;  // device copy of the model table
  int64_t size;
};

extern "C" {

int bgx_version(void) { return 100; }

const char* bgx_last_error(void) { return g_err.c_str(); }

int bgx_num_topologies(void) { return BGX_NUM_TOPOLOGIES; }

const char* bgx_topology_name(int id) {
  return (id >= 0 && id < BGX_NUM_TOPOLOGIES) ? kTopos[id].name : nullptr;
}

const char* bgx_topology_signature(int id) {
  return (id >= 0 && id < BGX_NUM_TOPOLOGIES) ? kTopos[id].sig : nullptr;
}

int bgx_topology_dims(int id, int32_t* dims) {
  if (id < 0 || id >= BGX_NUM_TOPOLOGIES)
    return fail(BGX_ERR_TOPOLOGY, "unknown topology id");
  if (!dims) return fail(BGX_ERR_ARG, "dims is null");
  for (int k = 0; k < 8; ++k) dims[k] = kTopos[id].dims[k];
  return BGX_OK;
}

int bgx_model_create(int id, const float* host_table, int64_t n_floats,
                     bgx_model** out) {
  if (id < 0 || id >= BGX_NUM_TOPOLOGIES)
    return fail(BGX_ERR_TOPOLOGY, "unknown topology id");
  if (!host_table || !out) return fail(BGX_ERR_ARG, "null argument");
  if (n_floats != kTopos[id].dims[7])
    return fail(BGX_ERR_ARG, "model table size does not match the topology");
  float* d = nullptr;
  int rc = check_hip(hipMalloc(&d, n_floats * sizeof(float)), "hipMalloc");
  if (rc) return rc;
  rc = check_hip(hipMemcpy(d, host_table, n_floats * sizeof(float),
                           hipMemcpyHostToDevice),
                 "hipMemcpy");
  if (rc) {
    (void)hipFree(d);
    return rc;
  }
  *out = new bgx_model{id, d, n_floats};
  return BGX_OK;
}

int bgx_model_destroy(bgx_model* m) {
  if (!m) return BGX_OK;
  int rc = check_hip(hipFree(m->table), "hipFree");
  delete m;
  return rc;
}

int bgx_model_topology(const bgx_model* m) { return m ? m->topo : -1; }

static int launch(const bgx_model* m, bool init_only, bgx::Launch a,
                  void* stream) {
  if (!m) return fail(BGX_ERR_ARG, "model is null");
  if (a.n < 0) return fail(BGX_ERR_ARG, "negative env count");
  if (a.n_steps < 0 || a.n_frames < 0)
    return fail(BGX_ERR_ARG, "negative step count");
  if (a.n == 0) return BGX_OK;
  if (!a.q || !a.qd) return fail(BGX_ERR_ARG, "q/qd must be non-null");
  a.M = m->table;
  const dim3 grid((a.n + bgx::WAVE - 1) / bgx::WAVE), block(bgx::WAVE);
  hipStream_t s = static_cast<hipStream_t>(stream);
  switch (m->topo) {
#define BGX_CASE(ID, T)                                                \
  case ID:                                                             \
    if (init_only)                                                     \
      hipLaunchKernelGGL(bgx::k_init<T>, grid, block, 0, s, a);        \
    else                                                               \
      hipLaunchKernelGGL(bgx::k_rollout<T>, grid, block, 0, s, a);     \
    break;
    BGX_FOR_EACH_TOPOLOGY(BGX_CASE)
#undef BGX_CASE
    default:
      return fail(BGX_ERR_TOPOLOGY, "unknown topology id");
  }
  return check_hip(hipGetLastError(), "kernel launch");
}

int bgx_init(const bgx_model* m, int32_t n, const float* q, const float* qd,
             float* x_pos, float* x_rot, float* xd_ang, float* xd_vel,
             void* stream) {
  bgx::Launch a{};
  a.n = n;
  a.q = const_cast<float*>(q);
  a.qd = const_cast<float*>(qd);
  a.x_pos = x_pos, a.x_rot = x_rot, a.xd_ang = xd_ang, a.xd_vel = xd_vel;
  return launch(m, true, a, stream);
}

int bgx_step(const bgx_model* m, int32_t n, int32_t n_frames, float* q,
             float* qd, const float* act, float* x_pos, float* x_rot,
             float* xd_ang, float* xd_vel, int32_t* status, void* stream) {
  bgx::Launch a{};
  a.n = n, a.n_steps = 1, a.n_frames = n_frames;
  a.q = q, a.qd = qd, a.act = act, a.act_stride = 0;
  a.x_pos = x_pos, a.x_rot = x_rot, a.xd_ang = xd_ang, a.xd_vel = xd_vel;
  a.status = status;
  return launch(m, false, a, stream);
}

int bgx_rollout(const bgx_model* m, int32_t n, int32_t n_steps,
                int32_t n_frames, float* q, float* qd, const float* act,
                int64_t act_step_stride, float* q_traj, float* qd_traj,
                float* x_pos, float* x_rot, float* xd_ang, float* xd_vel,
                int32_t* status, void* stream) {
  bgx::Launch a{};
  a.n = n, a.n_steps = n_steps, a.n_frames = n_frames;
  a.q = q, a.qd = qd, a.act = act, a.act_stride = act_step_stride;
  a.q_traj = q_traj, a.qd_traj = qd_traj;
  a.x_pos = x_pos, a.x_rot = x_rot, a.xd_ang = xd_ang, a.xd_vel = xd_vel;
  a.status = status;
  return launch(m, false, a, stream);
}

int bgx_step_probe(const bgx_model* m, int32_t n, int32_t n_frames, float* q,
                   float* qd, const float* act, float* probe, void* stream) {
  if (!probe) return fail(BGX_ERR_ARG, "probe is null");
  bgx::Launch a{};
  a.n = n, a.n_steps = 1, a.n_frames = n_frames;
  a.q = q, a.qd = qd, a.act = act, a.act_stride = 0;
  a.probe = probe;
  return launch(m, false, a, stream);
}

}  // extern "C"
