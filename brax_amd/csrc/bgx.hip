// C ABI of the generalized-coordinates step (declared in include/brax_gen.h).
//
// The kernels live in one object file per topology (bgx_topo.hip); this file
// validates arguments, owns the model handles and error strings, and
// dispatches through the per-topology launch tables.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/brax_gen.h"
#include "bgx_kernels.h"
#include "topologies.inc"

#define BGX_DECL_OPS(ID, T) const bgx::TopoOps* bgx_ops_##T();
BGX_FOR_EACH_TOPOLOGY(BGX_DECL_OPS)
#undef BGX_DECL_OPS

namespace {

// most env parts of the adjoint's backward pass (one stream each)
constexpr int kMaxAdjParts = 4;

#define BGX_OPS_PTR(ID, T) bgx_ops_##T(),
const bgx::TopoOps* const kOps[] = {BGX_FOR_EACH_TOPOLOGY(BGX_OPS_PTR)};
#undef BGX_OPS_PTR

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

int check_hip(hipError_t e, const char* what) {
  if (e == hipSuccess) return BGX_OK;
  return fail(BGX_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

bool valid_topo(int id) { return id >= 0 && id < BGX_NUM_TOPOLOGIES; }

// fill / strided copy helpers for the adjoint driver (no memset nodes, so the
// call can be captured in a graph)
__global__ void k_fill(float* p, long long n, float v) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}
__global__ void k_fill_i(int* p, long long n, int v) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}
// dst (rows, dw) <- src columns [c0, c0 + dw) of (rows, sw)
__global__ void k_cols(float* dst, const float* src, long long rows, int dw, int sw, int c0) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < rows * dw) dst[i] = src[(i / dw) * sw + c0 + (i % dw)];
}
// dst (rows, w) <- src (rows, w)
__global__ void k_copy(float* dst, const float* src, long long n) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = src[i];
}
dim3 grid1(long long n) { return dim3((unsigned)((n + 255) / 256)); }

}  // namespace

struct bgx_batch {
  int kernel_topo;            // topology whose batch kernel runs the batch
  float* tables;              // device copy of the models' tables, back to back
  bgx::BatchItem* items;      // device work table, longest items first
  int n_items;
  int* counter;               // device [2], zero between launches
  int grid;                   // persistent workgroups
  int64_t n_envs, q_floats, qd_floats, act_floats;
};

struct bgx_model {
  int topo;
  float* table;      // device copy of the model table(s)
  int64_t size;      // floats
  int maxiter;       // solver iterations (from the table header)
  int64_t n_tables;  // tables (one per env for per-env topologies)
  int epw[2];        // envs per workgroup: step kernels, adjoint kernels (0 = 64)
  int kc;            // fast contact slots of the step kernels (0: the default)
  int lanes;         // lanes per env of the step kernels (0: the default)
  int adj_parts;     // env parts of the adjoint's backward pass (0: automatic)
};

extern "C" {

int bgx_version(void) { return 200; }

const char* bgx_last_error(void) { return g_err.c_str(); }

int bgx_num_topologies(void) { return BGX_NUM_TOPOLOGIES; }

const char* bgx_topology_name(int id) { return valid_topo(id) ? kOps[id]->name : nullptr; }

const char* bgx_topology_signature(int id) { return valid_topo(id) ? kOps[id]->sig : nullptr; }

int bgx_topology_dims(int id, int32_t* dims) {
  if (!valid_topo(id)) return fail(BGX_ERR_TOPOLOGY, "unknown topology id");
  if (!dims) return fail(BGX_ERR_ARG, "dims is null");
  for (int k = 0; k < 8; ++k) dims[k] = kOps[id]->dims[k];
  return BGX_OK;
}

int bgx_topology_per_env(int id) { return valid_topo(id) ? kOps[id]->per_env : -1; }

int bgx_topology_tiles(int id, int32_t* tiles) {
  if (!valid_topo(id)) return -1;
  int n = 0;
  for (int k = 0; k < 3; ++k)
    if (kOps[id]->tiles[k] > 0) {
      if (tiles) tiles[n] = kOps[id]->tiles[k];
      ++n;
    }
  return n;
}

int bgx_model_create(int id, const float* host_table, int64_t n_floats, bgx_model** out) {
  if (!valid_topo(id)) return fail(BGX_ERR_TOPOLOGY, "unknown topology id");
  if (!host_table || !out) return fail(BGX_ERR_ARG, "null argument");
  const int64_t size = kOps[id]->dims[7];
  const bool per_env = kOps[id]->per_env != 0;
  if (per_env ? (n_floats < size || n_floats % size != 0) : n_floats != size)
    return fail(BGX_ERR_ARG, per_env ? "model table size is not a multiple of the "
                                       "topology's table size"
                                     : "model table size does not match the topology");
  float* d = nullptr;
  int rc = check_hip(hipMalloc(&d, n_floats * sizeof(float)), "hipMalloc");
  if (rc) return rc;
  rc = check_hip(hipMemcpy(d, host_table, n_floats * sizeof(float), hipMemcpyHostToDevice),
                 "hipMemcpy");
  if (rc) {
    (void)hipFree(d);
    return rc;
  }
  *out = new bgx_model{id, d, n_floats, (int)host_table[bgx::H_MAXITER], n_floats / size, {0, 0},
                       0};
  return BGX_OK;
}

int bgx_model_destroy(bgx_model* m) {
  if (!m) return BGX_OK;
  int rc = check_hip(hipFree(m->table), "hipFree");
  delete m;
  return rc;
}

int bgx_model_topology(const bgx_model* m) { return m ? m->topo : -1; }

int bgx_model_set_launch(bgx_model* m, int kernels, int32_t envs_per_workgroup) {
  if (!m) return fail(BGX_ERR_ARG, "model is null");
  if (kernels == BGX_KERNELS_TILE) {
    const int k = envs_per_workgroup;
    const int* t = kOps[m->topo]->tiles;
    if (k != 0 && k != t[0] && k != t[1] && k != t[2])
      return fail(BGX_ERR_ARG, "the step kernels are not compiled for this slot count");
    m->kc = k;
    return BGX_OK;
  }
  if (kernels == BGX_KERNELS_LANES) {
    const int k = envs_per_workgroup;
    if (k != 0 && k != 1 && k != kOps[m->topo]->mlanes)
      return fail(BGX_ERR_ARG, "lanes per env must be 0, 1 or the topology's multi-lane width");
    m->lanes = k;
    return BGX_OK;
  }
  if (kernels == BGX_KERNELS_ADJOINT_PARTS) {
    const int k = envs_per_workgroup;
    if (k < 0 || k > kMaxAdjParts)
      return fail(BGX_ERR_ARG, "adjoint parts must be 0 (automatic) to " +
                                   std::to_string(kMaxAdjParts));
    m->adj_parts = k;
    return BGX_OK;
  }
  if (kernels != BGX_KERNELS_STEP && kernels != BGX_KERNELS_ADJOINT)
    return fail(BGX_ERR_ARG, "unknown kernel group");
  const int e = envs_per_workgroup;
  if (e != 0 && (e < 1 || e > bgx::WAVE || (e & (e - 1)) != 0))
    return fail(BGX_ERR_ARG, "envs per workgroup must be 0 or a power of two up to 64");
  m->epw[kernels] = e;
  return BGX_OK;
}

int bgx_model_get_launch(const bgx_model* m, int kernels) {
  if (m && kernels == BGX_KERNELS_TILE) return m->kc ? m->kc : kOps[m->topo]->tiles[0];
  if (m && kernels == BGX_KERNELS_LANES) {
    const int w = kOps[m->topo]->mlanes;
    return m->lanes ? m->lanes : (w ? w : 1);
  }
  if (m && kernels == BGX_KERNELS_ADJOINT_PARTS) return m->adj_parts;
  if (!m || (kernels != BGX_KERNELS_STEP && kernels != BGX_KERNELS_ADJOINT)) return -1;
  return m->epw[kernels] ? m->epw[kernels] : bgx::WAVE;
}

static int launch(const bgx_model* m, bool init_only, bgx::Launch a, void* stream) {
  if (!m) return fail(BGX_ERR_ARG, "model is null");
  if (!valid_topo(m->topo)) return fail(BGX_ERR_TOPOLOGY, "unknown topology id");
  if (a.n < 0) return fail(BGX_ERR_ARG, "negative env count");
  if (a.n_steps < 0 || a.n_frames < 0) return fail(BGX_ERR_ARG, "negative step count");
  if (a.n == 0) return BGX_OK;
  if (!a.q || !a.qd) return fail(BGX_ERR_ARG, "q/qd must be non-null");
  if (kOps[m->topo]->per_env && a.n > m->n_tables)
    return fail(BGX_ERR_ARG, "more envs than per-env model tables");
  a.M = m->table;
  a.epw = m->epw[BGX_KERNELS_STEP];
  a.kc = m->kc;
  a.lanes = m->lanes;
  const bgx::TopoOps* ops = kOps[m->topo];
  a.q_step = (long long)a.n * ops->dims[1], a.qd_step = (long long)a.n * ops->dims[2];
  if ((a.q_traj || a.qd_traj) && a.n_seg > 1)
    return fail(BGX_ERR_ARG, "trajectories of a multi-segment launch are not supported");
  hipStream_t s = static_cast<hipStream_t>(stream);
  return check_hip(init_only ? ops->init(a, s) : ops->rollout(a, s), "kernel launch");
}

int bgx_init(const bgx_model* m, int32_t n, const float* q, const float* qd, float* x_pos,
             float* x_rot, float* xd_ang, float* xd_vel, void* stream) {
  bgx::Launch a{};
  a.n = n;
  a.q = const_cast<float*>(q);
  a.qd = const_cast<float*>(qd);
  a.x_pos = x_pos, a.x_rot = x_rot, a.xd_ang = xd_ang, a.xd_vel = xd_vel;
  return launch(m, true, a, stream);
}

int bgx_step(const bgx_model* m, int32_t n, int32_t n_frames, float* q, float* qd,
             const float* act, float* x_pos, float* x_rot, float* xd_ang, float* xd_vel,
             int32_t* status, void* stream) {
  bgx::Launch a{};
  a.n = n, a.n_steps = 1, a.n_frames = n_frames;
  a.q = q, a.qd = qd, a.act = act, a.act_stride = 0;
  a.x_pos = x_pos, a.x_rot = x_rot, a.xd_ang = xd_ang, a.xd_vel = xd_vel;
  a.status = status;
  return launch(m, false, a, stream);
}

int bgx_rollout(const bgx_model* m, int32_t n, int32_t n_steps, int32_t n_frames, float* q,
                float* qd, const float* act, int64_t act_step_stride, float* q_traj,
                float* qd_traj, float* x_pos, float* x_rot, float* xd_ang, float* xd_vel,
                int32_t* status, void* stream) {
  bgx::Launch a{};
  a.n = n, a.n_steps = n_steps, a.n_frames = n_frames;
  a.q = q, a.qd = qd, a.act = act, a.act_stride = act_step_stride;
  a.q_traj = q_traj, a.qd_traj = qd_traj;
  a.x_pos = x_pos, a.x_rot = x_rot, a.xd_ang = xd_ang, a.xd_vel = xd_vel;
  a.status = status;
  return launch(m, false, a, stream);
}

int bgx_contacts(const bgx_model* m, int32_t n, const float* q, float* dist, float* pos,
                 float* frame, void* stream) {
  if (!m) return fail(BGX_ERR_ARG, "model is null");
  if (!valid_topo(m->topo)) return fail(BGX_ERR_TOPOLOGY, "unknown topology id");
  if (n < 0) return fail(BGX_ERR_ARG, "negative env count");
  const bgx::TopoOps* ops = kOps[m->topo];
  if (n == 0 || !ops->contacts) return BGX_OK;  // no contact candidates
  if (!q || !dist || !pos || !frame) return fail(BGX_ERR_ARG, "null argument");
  if (ops->per_env && n > m->n_tables)
    return fail(BGX_ERR_ARG, "more envs than per-env model tables");
  bgx::Launch a{};
  a.M = m->table, a.n = n, a.epw = m->epw[BGX_KERNELS_STEP];
  a.q = const_cast<float*>(q);
  a.c_dist = dist, a.c_pos = pos, a.c_frame = frame;
  return check_hip(ops->contacts(a, static_cast<hipStream_t>(stream)), "kernel launch");
}

int bgx_frames_vjp(const bgx_model* m, int32_t n, const float* q, const float* qd,
                   const float* gx_pos, const float* gx_rot, const float* gxd_ang,
                   const float* gxd_vel, float* gq, float* gqd, void* stream) {
  if (!m) return fail(BGX_ERR_ARG, "model is null");
  if (!valid_topo(m->topo)) return fail(BGX_ERR_TOPOLOGY, "unknown topology id");
  if (n < 0) return fail(BGX_ERR_ARG, "negative env count");
  const bgx::TopoOps* ops = kOps[m->topo];
  if (!ops->frames_vjp)
    return fail(BGX_ERR_TOPOLOGY, std::string("no frames derivative for topology ") + ops->name);
  if (n == 0) return BGX_OK;
  if (!q || !qd || !gq || !gqd) return fail(BGX_ERR_ARG, "null argument");
  if (ops->per_env && n > m->n_tables)
    return fail(BGX_ERR_ARG, "more envs than per-env model tables");
  bgx::FramesVjp a{m->table, n, q, qd, gx_pos, gx_rot, gxd_ang, gxd_vel, gq, gqd};
  return check_hip(ops->frames_vjp(a, static_cast<hipStream_t>(stream)), "kernel launch");
}

int bgx_step_probe(const bgx_model* m, int32_t n, int32_t n_frames, float* q, float* qd,
                   const float* act, float* probe, void* stream) {
  if (!probe) return fail(BGX_ERR_ARG, "probe is null");
  bgx::Launch a{};
  a.n = n, a.n_steps = 1, a.n_frames = n_frames;
  a.q = q, a.qd = qd, a.act = act, a.act_stride = 0;
  a.probe = probe;
  return launch(m, false, a, stream);
}

// ---------------------------------------------------------------------------
// adjoint of a rollout

namespace {

// the adjoint's GPU path (on: tape-based stage A1/A2 kernels)
constexpr bool kAdjointEnabled = true;

struct AdjWorkspace {
  float* sub_q;   // (segs, k * n_frames, n, NQ)
  float* sub_qd;  // (segs, k * n_frames, n, NV)
  float* q;       // (segs, n, NQ) segment starts, advanced by the recompute
  float* qd;      // (segs, n, NV)
  float* adj;     // (n, NQ + NV)
  float* rec;     // (batch, rec, n): tapes of a batch of substeps
  int64_t segs;   // checkpoint segments recomputed by one launch
  int64_t batch;  // substeps taped by one stage-A1 launch
  int64_t bytes;
};

// The recompute runs up to this many bytes of segments' substep states in
// one launch: segments are independent given their checkpoints, so one
// launch recomputes them side by side.
constexpr int64_t kSegBatchBytes = 1ll << 30;
// Stage A1 tapes up to this many bytes of substeps in one launch: the
// substeps of a segment are independent once their start states are
// recomputed, so one launch fills the chip even for a small batch.
constexpr int64_t kTapeBatchBytes = 6ll << 30;

int64_t align256(int64_t b) { return (b + 255) / 256 * 256; }

// segs / batch 0: as many as fit in kSegBatchBytes / kTapeBatchBytes
AdjWorkspace adj_layout(const bgx::TopoOps* ops, float* base, int64_t n, int64_t k,
                        int64_t n_frames, int64_t segs, int64_t batch) {
  const int64_t nq = ops->dims[1], nv = ops->dims[2];
  AdjWorkspace w{};
  char* p = reinterpret_cast<char*>(base);
  int64_t off = 0;
  auto take = [&](int64_t floats) {
    float* r = base ? reinterpret_cast<float*>(p + off) : nullptr;
    off += align256(floats * (int64_t)sizeof(float));
    return r;
  };
  if (segs <= 0) {
    const int64_t per_seg = (k * n_frames + 1) * n * (nq + nv) * (int64_t)sizeof(float);
    segs = per_seg > 0 ? kSegBatchBytes / per_seg : 1;
  }
  if (segs > 65535) segs = 65535;
  w.segs = segs < 1 ? 1 : segs;
  if (batch <= 0) {
    const int64_t per_sub = (int64_t)ops->adj_rec * n * (int64_t)sizeof(float);
    batch = per_sub > 0 ? kTapeBatchBytes / per_sub : 1;
  }
  if (batch > k * n_frames) batch = k * n_frames;
  if (batch > 65535) batch = 65535;
  w.batch = batch < 1 ? 1 : batch;
  w.sub_q = take(w.segs * k * n_frames * n * nq);
  w.sub_qd = take(w.segs * k * n_frames * n * nv);
  w.q = take(w.segs * n * nq);
  w.qd = take(w.segs * n * nv);
  w.adj = take(n * (nq + nv));
  w.rec = take((int64_t)ops->adj_rec * n * w.batch);
  w.bytes = off;
  return w;
}

}  // namespace

int64_t bgx_adjoint_workspace_bytes(const bgx_model* m, int32_t n, int32_t ckpt_every,
                                    int32_t n_frames) {
  if (!m || !valid_topo(m->topo) || n < 0 || ckpt_every < 1 || n_frames < 1) return -1;
  return adj_layout(kOps[m->topo], nullptr, n, ckpt_every, n_frames, 0, 0).bytes;
}

int64_t bgx_adjoint_workspace_min_bytes(const bgx_model* m, int32_t n, int32_t ckpt_every,
                                        int32_t n_frames) {
  if (!m || !valid_topo(m->topo) || n < 0 || ckpt_every < 1 || n_frames < 1) return -1;
  return adj_layout(kOps[m->topo], nullptr, n, ckpt_every, n_frames, 1, 1).bytes;
}

}  // extern "C"

namespace {

// Envs are independent in the adjoint as in the forward, so the backward
// pass of a batch splits into parts of whole envs, each the same problem at
// a smaller batch.  A part runs on its own stream: the reverse sweep (stage
// A2) of one part keeps a few SIMDs busy for its whole length, and the
// other parts' stage-B and stage-A1 waves fill the rest.  Every env's
// arithmetic is the same in any split (the kernels see the same inputs), so
// the results are bitwise those of one part.  It pays only while the
// streams keep hardware queues of their own: where they share one (more
// streams in the process than queues), the parts run one after another and
// each pays the sweep's full dependent chain, so one part is the default.
constexpr int kMaxAdjDevices = 64;

struct AdjStreams {
  hipStream_t s[kMaxAdjParts - 1];
  hipEvent_t fork, join[kMaxAdjParts - 1];
};
std::mutex g_adj_mu;  // guards the table and orders the event records
AdjStreams* g_adj_streams[kMaxAdjDevices] = {};

// the helper streams and events of the current device, created on first use
// (default priority; kept until the process ends); call with g_adj_mu held
int adj_streams(AdjStreams** out) {
  int dev = -1;
  int rc = check_hip(hipGetDevice(&dev), "hipGetDevice");
  if (rc) return rc;
  if (dev < 0 || dev >= kMaxAdjDevices) return fail(BGX_ERR_ARG, "device index out of range");
  if (!g_adj_streams[dev]) {
    AdjStreams h{};
    for (int i = 0; i < kMaxAdjParts - 1 && !rc; ++i) {
      rc = check_hip(hipStreamCreateWithFlags(&h.s[i], hipStreamNonBlocking), "adjoint stream");
      if (!rc)
        rc = check_hip(hipEventCreateWithFlags(&h.join[i], hipEventDisableTiming),
                       "adjoint event");
    }
    if (!rc)
      rc = check_hip(hipEventCreateWithFlags(&h.fork, hipEventDisableTiming), "adjoint event");
    if (rc) {  // release what was created; the next split call tries again
      for (int i = 0; i < kMaxAdjParts - 1; ++i) {
        if (h.s[i]) (void)hipStreamDestroy(h.s[i]);
        if (h.join[i]) (void)hipEventDestroy(h.join[i]);
      }
      if (h.fork) (void)hipEventDestroy(h.fork);
      return rc;
    }
    g_adj_streams[dev] = new AdjStreams(h);
  }
  *out = g_adj_streams[dev];
  return BGX_OK;
}

// parts of the backward pass for a batch of n envs (0 = one), each a
// multiple of 64 envs but the last
int adj_parts(int setting, int n, int* per) {
  int p = setting > 0 ? setting : 1;
  p = p < 1 ? 1 : (p > kMaxAdjParts ? kMaxAdjParts : p);
  *per = ((n + p - 1) / p + 63) / 64 * 64;
  return (n + *per - 1) / *per;
}

// the part of workspace w that holds envs [env0, env0 + nh) in the layout
// of a batch of nh envs (each buffer is a multiple of n per-env blocks)
AdjWorkspace part_view(const AdjWorkspace& w, const bgx::TopoOps* ops, int64_t k,
                       int64_t n_frames, int64_t env0) {
  const int64_t nq = ops->dims[1], nv = ops->dims[2];
  AdjWorkspace p = w;
  p.sub_q = w.sub_q + w.segs * k * n_frames * env0 * nq;
  p.sub_qd = w.sub_qd + w.segs * k * n_frames * env0 * nv;
  p.q = w.q + w.segs * env0 * nq;
  p.qd = w.qd + w.segs * env0 * nv;
  p.adj = w.adj + env0 * (nq + nv);
  p.rec = w.rec + w.batch * (int64_t)ops->adj_rec * env0;
  return p;
}

// dst (rows, w) <- rows of w floats of src, ss floats apart
__global__ void k_copy_rows(float* dst, const float* src, long long rows, long long w,
                            long long ss) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < rows * w) dst[i] = src[(i / w) * ss + i % w];
}

// The backward pass of envs [env0, env0 + nh) of a batch of n on stream s,
// with that part's workspace w (laid out for nh envs).  The caller's arrays
// keep the batch's layout: per-step blocks of n envs, the part's rows at
// env0.
int adjoint_part(const bgx_model* m, int n, int env0, int nh, int n_steps, int n_frames,
                 int ckpt_every, const float* ckpt_q, const float* ckpt_qd, const float* act,
                 int64_t act_step_stride, const float* gq_traj, const float* gqd_traj,
                 float* gact, float* gq0, float* gqd0, const AdjWorkspace& w,
                 int32_t* status, hipStream_t s) {
  const bgx::TopoOps* ops = kOps[m->topo];
  const int nq = ops->dims[1], nv = ops->dims[2], nu = ops->dims[3];
  const float* act_p = act ? act + (long long)env0 * nu : nullptr;
  int rc;
  const int n_seg = (n_steps + ckpt_every - 1) / ckpt_every;
  for (int seg_hi = n_seg - 1; seg_hi >= 0;) {
    // a group of segments, recomputed side by side by one launch (a short
    // last segment goes alone)
    const bool short_last = seg_hi == n_seg - 1 && n_steps % ckpt_every != 0;
    const int cnt = short_last ? 1 : (seg_hi + 1 < (int)w.segs ? seg_hi + 1 : (int)w.segs);
    const int seg_lo = seg_hi - cnt + 1;
    const int ns_group = short_last ? n_steps - seg_hi * ckpt_every : ckpt_every;
    hipLaunchKernelGGL(k_copy_rows, grid1((long long)cnt * nh * nq), dim3(256), 0, s, w.q,
                       ckpt_q + ((long long)seg_lo * n + env0) * nq, (long long)cnt,
                       (long long)nh * nq, (long long)n * nq);
    hipLaunchKernelGGL(k_copy_rows, grid1((long long)cnt * nh * nv), dim3(256), 0, s, w.qd,
                       ckpt_qd + ((long long)seg_lo * n + env0) * nv, (long long)cnt,
                       (long long)nh * nv, (long long)n * nv);
    bgx::Launch f{};
    f.M = m->table, f.n = nh, f.n_steps = ns_group, f.n_frames = n_frames;
    f.q = w.q, f.qd = w.qd;
    f.act = act_p ? act_p + (long long)seg_lo * ckpt_every * act_step_stride : nullptr;
    f.act_stride = act_step_stride;
    f.n_seg = cnt, f.seg_act_stride = (long long)ckpt_every * act_step_stride;
    f.sub_q = w.sub_q, f.sub_qd = w.sub_qd;
    f.status = status ? status + env0 : nullptr, f.status_or = 1;
    f.epw = m->epw[BGX_KERNELS_ADJOINT];
    f.kc = m->kc;  // the forward's tile and kernel: the recompute repeats its states
    f.lanes = m->lanes;
    if ((rc = check_hip(ops->rollout(f, s), "recompute launch"))) return rc;
    for (int seg = seg_hi; seg >= seg_lo; --seg) {
      const int s0 = seg * ckpt_every;
      // substep states of this segment (the launch's segment stride is its
      // own length ns_group)
      const long long sub_off = (long long)(seg - seg_lo) * ns_group * n_frames * nh;
      float* sq = w.sub_q + sub_off * nq;
      float* sqd = w.sub_qd + sub_off * nv;
      // substeps of the segment, last first, taped in batches of w.batch
      const int n_sub = ns_group * n_frames;
      const long long rec_stride = (long long)ops->adj_rec * nh;
      for (int hi = n_sub - 1; hi >= 0; hi -= (int)w.batch) {
        const int lo = hi + 1 - (int)w.batch > 0 ? hi + 1 - (int)w.batch : 0;
        bgx::AdjLaunch t{};
        t.M = m->table, t.n = nh, t.epw = m->epw[BGX_KERNELS_ADJOINT];
        t.q = sq, t.qd = sqd;
        t.act = act_p ? act_p + (long long)s0 * act_step_stride : nullptr;
        t.act_stride = act_step_stride, t.n_frames = n_frames;
        t.sub_lo = lo, t.n_sub = hi - lo + 1;
        t.rec = w.rec, t.rec_stride = rec_stride;
        if ((rc = check_hip(ops->adj_fwd(t, s), "adjoint stage A1 launch"))) return rc;
        for (int i = hi; i >= lo; --i) {
          const int t_step = s0 + i / n_frames;
          const bool step_end = (i % n_frames) == n_frames - 1;
          const long long row = (long long)t_step * n + env0;
          bgx::AdjLaunch a{};
          a.M = m->table, a.n = nh, a.epw = m->epw[BGX_KERNELS_ADJOINT];
          a.q = sq + (long long)i * nh * nq;
          a.qd = sqd + (long long)i * nh * nv;
          a.act = act_p ? act_p + (long long)t_step * act_step_stride : nullptr;
          a.gq_add = (step_end && gq_traj) ? gq_traj + row * nq : nullptr;
          a.gqd_add = (step_end && gqd_traj) ? gqd_traj + row * nv : nullptr;
          a.adj = w.adj, a.rec = w.rec + (long long)(i - lo) * rec_stride;
          a.gact = gact ? gact + row * nu : nullptr;
          if ((rc = check_hip(ops->adj_rev(a, s), "adjoint stage A2 launch"))) return rc;
          if ((rc = check_hip(ops->adj_b(a, s), "adjoint stage B launch"))) return rc;
        }
      }
    }
    seg_hi = seg_lo - 1;
  }
  hipLaunchKernelGGL(k_cols, grid1((long long)nh * nq), dim3(256), 0, s,
                     gq0 + (long long)env0 * nq, w.adj, (long long)nh, nq, nq + nv, 0);
  hipLaunchKernelGGL(k_cols, grid1((long long)nh * nv), dim3(256), 0, s,
                     gqd0 + (long long)env0 * nv, w.adj, (long long)nh, nv, nq + nv, nq);
  return check_hip(hipGetLastError(), "output copy");
}

// The adjoint of a rollout with its workspace at `workspace` (shared by the
// workspace-argument entry point and the no-allocation one; split_streams
// false keeps every launch on the caller's stream and creates nothing).
int rollout_adjoint_impl(const bgx_model* m, int32_t n, int32_t n_steps, int32_t n_frames,
                         int32_t ckpt_every, const float* ckpt_q, const float* ckpt_qd,
                         const float* act, int64_t act_step_stride, const float* gq_traj,
                         const float* gqd_traj, float* gact, float* gq0, float* gqd0,
                         float* workspace, int64_t workspace_bytes, int32_t* status,
                         void* stream, bool split_streams) {
  if (!m) return fail(BGX_ERR_ARG, "model is null");
  if (!valid_topo(m->topo)) return fail(BGX_ERR_TOPOLOGY, "unknown topology id");
  if (n < 0 || n_steps < 0 || n_frames < 1 || ckpt_every < 1)
    return fail(BGX_ERR_ARG, "bad env, step, frame or checkpoint count");
  if (!kOps[m->topo]->adj_fwd)
    return fail(BGX_ERR_TOPOLOGY, "the adjoint is not compiled for per-env topologies");
  // (a topology without contact candidates or limits never runs the solver)
  const bool has_rows = kOps[m->topo]->dims[4] > 0 || kOps[m->topo]->dims[5] > 0;
  if (has_rows && m->maxiter > bgx::ADJ_MAXITER)
    return fail(BGX_ERR_ARG, "the adjoint unrolls at most " +
                                 std::to_string(bgx::ADJ_MAXITER) + " solver iterations");
  if (n == 0 || n_steps == 0) return BGX_OK;
  if (!ckpt_q || !ckpt_qd || !gq0 || !gqd0 || !workspace)
    return fail(BGX_ERR_ARG, "null checkpoint, output or workspace pointer");
  const bgx::TopoOps* ops = kOps[m->topo];
  const int nq = ops->dims[1], nv = ops->dims[2], nu = ops->dims[3];
  if (nu > 0 && !gact) return fail(BGX_ERR_ARG, "gact is null");
  // segments per recompute launch: the preferred count if the workspace
  // holds it with one tape, else one; then the most substeps per stage-A1
  // launch that the rest holds (the tapes are the layout's last piece)
  const AdjWorkspace w1 = adj_layout(ops, nullptr, n, ckpt_every, n_frames, 1, 1);
  if (workspace_bytes < w1.bytes) return fail(BGX_ERR_ARG, "workspace too small");
  const AdjWorkspace wp = adj_layout(ops, nullptr, n, ckpt_every, n_frames, 0, 1);
  const int64_t segs = workspace_bytes >= wp.bytes ? wp.segs : 1;
  const AdjWorkspace ws = adj_layout(ops, nullptr, n, ckpt_every, n_frames, segs, 1);
  const int64_t per_sub = (int64_t)ops->adj_rec * n * (int64_t)sizeof(float);
  const int64_t fixed = ws.bytes - align256(per_sub);
  int64_t batch = per_sub > 0 ? (workspace_bytes - fixed) / per_sub : 1;
  AdjWorkspace w = adj_layout(ops, workspace, n, ckpt_every, n_frames, segs, batch);
  while (w.bytes > workspace_bytes && batch > 1)
    w = adj_layout(ops, workspace, n, ckpt_every, n_frames, segs, --batch);
  if (!kAdjointEnabled) return fail(BGX_ERR_DISABLED, "bgx_rollout_adjoint is switched off");
  hipStream_t s = static_cast<hipStream_t>(stream);
  int rc;
  const long long nact = (long long)n_steps * n * nu, nadj = (long long)n * (nq + nv);
  if (nu > 0) hipLaunchKernelGGL(k_fill, grid1(nact), dim3(256), 0, s, gact, nact, 0.f);
  hipLaunchKernelGGL(k_fill, grid1(nadj), dim3(256), 0, s, w.adj, nadj, 0.f);
  if (status) hipLaunchKernelGGL(k_fill_i, grid1(n), dim3(256), 0, s, status, (long long)n, 0);
  if ((rc = check_hip(hipGetLastError(), "fill"))) return rc;

  // the parts, on the caller's stream and up to kMaxAdjParts - 1 helper
  // streams forked from it and joined back into it
  int per = n;
  const int parts = split_streams ? adj_parts(m->adj_parts, n, &per) : 1;
  if (parts == 1)
    return adjoint_part(m, n, 0, n, n_steps, n_frames, ckpt_every, ckpt_q, ckpt_qd, act,
                        act_step_stride, gq_traj, gqd_traj, gact, gq0, gqd0, w, status, s);
  std::lock_guard<std::mutex> lock(g_adj_mu);
  AdjStreams* hs = nullptr;
  if ((rc = adj_streams(&hs))) return rc;
  if ((rc = check_hip(hipEventRecord(hs->fork, s), "adjoint fork"))) return rc;
  for (int p = 1; p < parts; ++p)
    if ((rc = check_hip(hipStreamWaitEvent(hs->s[p - 1], hs->fork, 0), "adjoint fork")))
      return rc;
  for (int p = 0; p < parts && !rc; ++p) {
    const int env0 = p * per, nh = n - env0 < per ? n - env0 : per;
    rc = adjoint_part(m, n, env0, nh, n_steps, n_frames, ckpt_every, ckpt_q, ckpt_qd, act,
                      act_step_stride, gq_traj, gqd_traj, gact, gq0, gqd0,
                      part_view(w, ops, ckpt_every, n_frames, env0), status,
                      p ? hs->s[p - 1] : s);
  }
  // join every helper, also after a failed launch, so that the caller's
  // stream never runs ahead of work in flight on the workspace
  for (int p = 1; p < parts; ++p) {
    const int r1 = check_hip(hipEventRecord(hs->join[p - 1], hs->s[p - 1]), "adjoint join");
    const int r2 = check_hip(hipStreamWaitEvent(s, hs->join[p - 1], 0), "adjoint join");
    if (!rc) rc = r1 ? r1 : r2;
  }
  return rc;
}

// Workspace of bgx_rollout_adjoint_noalloc: one array per GPU in this file's
// device code object, held from the library's first use of a GPU until the
// process ends.  In host code `scratch` names a host-side stand-in, never the
// device memory: its device address on each GPU comes from
// hipGetSymbolAddress (scratch_on_current_device) and only that address is
// passed to kernels.
constexpr int64_t kScratchBytes = BGX_ADJOINT_SCRATCH_BYTES;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// the device address of `scratch` on the current device (looked up once per
// device, with that device current)
int scratch_on_current_device(float** out) {
  int dev = -1;
  int rc = check_hip(hipGetDevice(&dev), "hipGetDevice");
  if (rc) return rc;
  if (dev < 0 || dev >= kMaxDevices) return fail(BGX_ERR_ARG, "device index out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    rc = check_hip(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
    if (rc) return rc;
    if (!p) return fail(BGX_ERR_HIP, "hipGetSymbolAddress(scratch) returned null");
    g_scratch_ptr[dev] = p;
  }
  *out = static_cast<float*>(g_scratch_ptr[dev]);
  return BGX_OK;
}

}  // namespace

extern "C" {

int bgx_rollout_adjoint(const bgx_model* m, int32_t n, int32_t n_steps, int32_t n_frames,
                        int32_t ckpt_every, const float* ckpt_q, const float* ckpt_qd,
                        const float* act, int64_t act_step_stride, const float* gq_traj,
                        const float* gqd_traj, float* gact, float* gq0, float* gqd0,
                        float* workspace, int64_t workspace_bytes, int32_t* status,
                        void* stream) {
  return rollout_adjoint_impl(m, n, n_steps, n_frames, ckpt_every, ckpt_q, ckpt_qd, act,
                              act_step_stride, gq_traj, gqd_traj, gact, gq0, gqd0, workspace,
                              workspace_bytes, status, stream, true);
}

int64_t bgx_adjoint_scratch_bytes(void) { return kScratchBytes; }

// ---------------------------------------------------------------------------
// heterogeneous batches: one persistent launch (bgx_batch.h)

int bgx_batch_create(int32_t n_models, const bgx_model* const* models,
                     const int32_t* envs_per_model, bgx_batch** out) {
  if (!out || n_models < 0 || (n_models > 0 && (!models || !envs_per_model)))
    return fail(BGX_ERR_ARG, "null argument or negative model count");
  *out = nullptr;
  int kernel_topo = -1, chain_topo = -1;
  for (int m = 0; m < n_models; ++m) {
    const bgx_model* md = models[m];
    if (!md || !valid_topo(md->topo)) return fail(BGX_ERR_ARG, "null or unknown model");
    if (envs_per_model[m] < 0) return fail(BGX_ERR_ARG, "negative env count");
    const bgx::TopoOps* ops = kOps[md->topo];
    if (ops->per_env && envs_per_model[m] > md->n_tables)
      return fail(BGX_ERR_ARG, "more envs than per-env model tables");
    if (ops->chain) {
      chain_topo = md->topo;
      continue;
    }
    if (!ops->batch)
      return fail(BGX_ERR_TOPOLOGY, std::string("topology ") + ops->name +
                                        " has no batch kernel (codegen.BATCH_SYSTEMS)");
    if (kernel_topo >= 0 && kernel_topo != md->topo)
      return fail(BGX_ERR_TOPOLOGY, "a batch holds chains and envs of at most one other "
                                    "topology");
    kernel_topo = md->topo;
  }
  if (kernel_topo < 0) kernel_topo = chain_topo;  // chains only (or empty)
  // work items: 64 tree envs, or one wavefront of chains (64 / G envs)
  std::vector<bgx::BatchItem> items;
  std::vector<int> rank;  // ordering key: tree items first, then wider chains
  std::vector<int64_t> table_off(n_models), table_len(n_models);
  int64_t env = 0, qo = 0, qdo = 0, ao = 0, tfloats = 0;
  for (int m = 0; m < n_models; ++m) {
    const bgx_model* md = models[m];
    const bgx::TopoOps* ops = kOps[md->topo];
    const int nq = ops->dims[1], nv = ops->dims[2], nu = ops->dims[3];
    const int64_t size = ops->dims[7];
    const int n = envs_per_model[m];
    // the model's tables in the batch's table buffer: the first n of a
    // per-env model, the one shared table otherwise
    table_off[m] = tfloats;
    table_len[m] = ops->per_env ? (int64_t)n * size : size;
    tfloats += table_len[m];
    const int per = ops->chain ? bgx::WAVE / bgx::chain_lanes(ops->dims[0]) : bgx::WAVE;
    for (int i = 0; i < n; i += per) {
      bgx::BatchItem b{};
      b.m_off = table_off[m] + (ops->per_env ? (int64_t)i * size : 0);
      b.M_stride = ops->per_env ? (int)size : 0;
      b.env = env + i;
      b.q = qo + (int64_t)i * nq, b.qd = qdo + (int64_t)i * nv, b.act = ao + (int64_t)i * nu;
      b.kind = ops->chain ? bgx::BATCH_CHAIN : bgx::BATCH_TREE;
      b.n = std::min(per, n - i);
      b.L = ops->chain ? ops->dims[0] : 0;
      items.push_back(b);
      // longest first: a tree wavefront (one env per lane, the whole step
      // with its contact solve) outlasts every chain wavefront; chains by
      // lane count, then length
      rank.push_back(ops->chain ? bgx::chain_lanes(b.L) * 64 + b.L : 1 << 20);
    }
    env += n, qo += (int64_t)n * nq, qdo += (int64_t)n * nv, ao += (int64_t)n * nu;
  }
  std::vector<size_t> order(items.size());
  for (size_t i = 0; i < order.size(); ++i) order[i] = i;
  std::stable_sort(order.begin(), order.end(),
                   [&](size_t a, size_t b) { return rank[a] > rank[b]; });
  std::vector<bgx::BatchItem> sorted;
  for (size_t i : order) sorted.push_back(items[i]);
  bgx_batch* b = new bgx_batch{};
  b->kernel_topo = kernel_topo;
  b->n_items = (int)sorted.size();
  b->n_envs = env, b->q_floats = qo, b->qd_floats = qdo, b->act_floats = ao;
  int rc = BGX_OK;
  if (b->n_items > 0) {
    const bgx::TopoOps* kops = kOps[kernel_topo];
    int dev = 0, cus = 0;
    rc = check_hip(hipGetDevice(&dev), "hipGetDevice");
    if (!rc)
      rc = check_hip(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev),
                     "hipDeviceGetAttribute");
    if (!rc) {
      const int64_t resident = (int64_t)kops->batch_occupancy() * std::max(cus, 1);
      b->grid = (int)std::max<int64_t>(1, std::min<int64_t>(resident, b->n_items));
      rc = check_hip(hipMalloc(&b->items, sorted.size() * sizeof(bgx::BatchItem)), "hipMalloc");
    }
    if (!rc)
      rc = check_hip(hipMemcpy(b->items, sorted.data(), sorted.size() * sizeof(bgx::BatchItem),
                               hipMemcpyHostToDevice),
                     "hipMemcpy");
    if (!rc) rc = check_hip(hipMalloc(&b->tables, std::max<int64_t>(tfloats, 1) * sizeof(float)),
                            "hipMalloc");
    for (int m = 0; m < n_models && !rc; ++m)
      if (table_len[m] > 0)
        rc = check_hip(hipMemcpy(b->tables + table_off[m], models[m]->table,
                                 table_len[m] * sizeof(float), hipMemcpyDeviceToDevice),
                       "hipMemcpy");
    if (!rc) rc = check_hip(hipMalloc(&b->counter, 2 * sizeof(int)), "hipMalloc");
    if (!rc) rc = check_hip(hipMemset(b->counter, 0, 2 * sizeof(int)), "hipMemset");
    if (!rc) rc = check_hip(hipDeviceSynchronize(), "hipDeviceSynchronize");
  }
  if (rc) {
    if (b->items) (void)hipFree(b->items);
    if (b->tables) (void)hipFree(b->tables);
    if (b->counter) (void)hipFree(b->counter);
    delete b;
    return rc;
  }
  *out = b;
  return BGX_OK;
}

int bgx_batch_destroy(bgx_batch* b) {
  if (!b) return BGX_OK;
  int rc = BGX_OK;
  if (b->items) rc = check_hip(hipFree(b->items), "hipFree");
  if (b->tables) {
    const int rc2 = check_hip(hipFree(b->tables), "hipFree");
    if (!rc) rc = rc2;
  }
  if (b->counter) {
    const int rc2 = check_hip(hipFree(b->counter), "hipFree");
    if (!rc) rc = rc2;
  }
  delete b;
  return rc;
}

int bgx_batch_sizes(const bgx_batch* b, int64_t* sizes) {
  if (!b || !sizes) return fail(BGX_ERR_ARG, "null argument");
  sizes[0] = b->n_envs, sizes[1] = b->q_floats, sizes[2] = b->qd_floats;
  sizes[3] = b->act_floats, sizes[4] = b->n_items, sizes[5] = b->grid;
  return BGX_OK;
}

int bgx_batch_rollout(const bgx_batch* b, int32_t n_steps, int32_t n_frames, float* q,
                      float* qd, const float* act, int64_t act_step_stride, float* q_traj,
                      float* qd_traj, int32_t* status, void* stream) {
  if (!b) return fail(BGX_ERR_ARG, "batch is null");
  if (n_steps < 0 || n_frames < 0) return fail(BGX_ERR_ARG, "negative step count");
  if (b->n_items == 0) return BGX_OK;
  if (!q || !qd) return fail(BGX_ERR_ARG, "q/qd must be non-null");
  if (act && act_step_stride < b->act_floats && n_steps > 1)
    return fail(BGX_ERR_ARG, "action step stride shorter than a step's actions");
  bgx::Launch a{};
  a.M = b->tables;
  a.items = b->items, a.n_items = b->n_items, a.counter = b->counter;
  a.n = (int)b->n_envs, a.n_seg = 1;
  a.n_steps = n_steps, a.n_frames = n_frames;
  a.q = q, a.qd = qd, a.act = b->act_floats > 0 ? act : nullptr, a.act_stride = act_step_stride;
  a.q_traj = q_traj, a.qd_traj = qd_traj;
  a.q_step = b->q_floats, a.qd_step = b->qd_floats;
  a.status = status;
  return check_hip(kOps[b->kernel_topo]->batch(a, b->grid, static_cast<hipStream_t>(stream)),
                   "batch kernel launch");
}

int bgx_rollout_adjoint_noalloc(const bgx_model* m, int32_t n, int32_t n_steps,
                                int32_t n_frames, int32_t ckpt_every, const float* ckpt_q,
                                const float* ckpt_qd, const float* act,
                                int64_t act_step_stride, const float* gq_traj,
                                const float* gqd_traj, float* gact, float* gq0, float* gqd0,
                                int32_t* status, void* stream) {
  if (!m) return fail(BGX_ERR_ARG, "model is null");
  if (!valid_topo(m->topo)) return fail(BGX_ERR_TOPOLOGY, "unknown topology id");
  if (n < 0 || n_steps < 0 || n_frames < 1 || ckpt_every < 1)
    return fail(BGX_ERR_ARG, "bad env, step, frame or checkpoint count");
  if (n == 0 || n_steps == 0) return BGX_OK;
  const int64_t need = adj_layout(kOps[m->topo], nullptr, n, ckpt_every, n_frames, 1, 1).bytes;
  if (need > kScratchBytes)
    return fail(BGX_ERR_ARG, "bgx_rollout_adjoint_noalloc: " + std::to_string(n) +
                                 " envs with a checkpoint every " + std::to_string(ckpt_every) +
                                 " env steps need " + std::to_string(need) +
                                 " bytes of workspace; its scratch holds " +
                                 std::to_string(kScratchBytes) +
                                 " (use bgx_rollout_adjoint with a workspace)");
  float* ws = nullptr;
  int rc = scratch_on_current_device(&ws);
  if (rc) return rc;
  return rollout_adjoint_impl(m, n, n_steps, n_frames, ckpt_every, ckpt_q, ckpt_qd, act,
                              act_step_stride, gq_traj, gqd_traj, gact, gq0, gqd0, ws,
                              kScratchBytes, status, stream, false);
}

}  // extern "C"
