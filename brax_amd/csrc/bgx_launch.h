// Arguments of the rollout kernel: one launch over a homogeneous batch (one
// model, env-major arrays), or one persistent launch over a heterogeneous
// batch's work table (bgx_batch.h).
#pragma once

namespace bgx {

constexpr int BATCH_TREE = 0;
constexpr int BATCH_CHAIN = 1;

// One work item: n consecutive envs of one model.  A homogeneous launch
// forms its workgroup's item from blockIdx; a batch launch reads items from
// its table.  Offsets are in floats from the launch's arrays.
struct BatchItem {
  // model table of the item's first env: Launch::M + m_off (an offset, not
  // a pointer: the kernel argument M is known to be global memory, so the
  // table is read with global loads)
  long long m_off;
  long long env;     // first env (index into status, probe, frames)
  long long q, qd;   // offsets of the first env's q and qd rows
  long long act;     // offset of the first env's action row within a step
  int kind;          // BATCH_TREE or BATCH_CHAIN
  int n;             // envs in the item
  int L;             // chain: links (= its q, qd and action width)
  int M_stride;      // floats between consecutive envs' tables (0: shared)
};

struct Launch {
  const float* M;
  int n;
  int n_steps;
  int n_frames;
  float* q;
  float* qd;
  const float* act;
  long long act_stride;
  float* q_traj;   // step s row of an env at q_traj + s * q_step + (its q offset)
  float* qd_traj;
  float* x_pos;
  float* x_rot;
  float* xd_ang;
  float* xd_vel;
  int* status;
  int status_or;  // OR the status bits into status instead of assigning
  float* probe;
  float* sub_q;   // (n_steps * n_frames, n, NQ): state at the start of each substep
  float* sub_qd;  // (n_steps * n_frames, n, NV)
  int epw;        // envs (lanes) per workgroup, 1..WAVE; 0 = WAVE
  // n_seg > 1 (the adjoint's recompute): n_seg independent rollouts of the
  // same length, blockIdx.y = j: state at q + j n NQ (qd alike), actions
  // from act + j seg_act_stride, substep states at sub_q + j n_steps
  // n_frames n NQ (sub_qd alike), status bits ORed atomically
  int n_seg;
  long long seg_act_stride;
  // contacts of the state (bgx_contacts): dist (n, NCAND), pos (n, NCAND, 3),
  // frame (n, NCAND, 3, 3) rows n, t1, t2
  float* c_dist;
  float* c_pos;
  float* c_frame;
  // floats per step row of q_traj and qd_traj (n NQ, n NV for a homogeneous
  // launch; the packed widths for a batch)
  long long q_step, qd_step;
  // a homogeneous launch of a chain topology: its link count
  int chain_L;
  // a homogeneous launch of a tree topology: its fast contact slots (the
  // LDS tile; 0 = the topology's default)
  int kc;
  // lanes per env of a homogeneous tree launch: 1 = one thread per env,
  // 0 = the multi-lane kernel where the topology has it (bgx_mtree.h)
  int lanes;
  // batch mode (items non-null): a persistent launch over the work table;
  // the arrays are packed and the items give each env's offsets; probe,
  // frames and substep states are not written
  const BatchItem* items;
  int n_items;
  int* counter;  // [2]: next item, workgroups done (zero between launches)
};

}  // namespace bgx
