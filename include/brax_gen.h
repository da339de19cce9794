/* brax_gen.h -- C ABI of the MI355X generalized-coordinates physics step.
 *
 * The drop-in boundary for brax's generalized pipeline.  The reference's
 * interface is a Python module with two functions,
 *   init(sys, q, qd, unused_act=None, unused_ctrl=None, debug=False) -> State
 *     (brax/generalized/pipeline.py:32-61)
 *   step(sys, state, act, debug=False) -> State
 *     (brax/generalized/pipeline.py:64-94)
 * driven n_frames times per env step by PipelineEnv.pipeline_step
 * (brax/envs/base.py:128-137) and batched by jax.vmap
 * (brax/envs/wrappers/training.py:59-72).  These entry points are the batched
 * equivalents; brax_amd/generalized/pipeline.py binds them with ctypes.
 *
 * Conventions: every array is a device pointer owned by the caller, float32,
 * env-major ("(N, NQ)" means env e's q at q + e * NQ).  Calls are enqueued on
 * the given HIP stream (NULL = default stream) and return at once; a nonzero
 * return is an error code and bgx_last_error() describes it.  The library
 * holds no per-call workspace and no mutable global state beyond the
 * per-thread error string, so calls on different streams may overlap; the
 * one exception is bgx_rollout_adjoint_noalloc, whose calls share one
 * scratch array per GPU.
 */
#ifndef BRAX_GEN_H_
#define BRAX_GEN_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  BGX_OK = 0,
  BGX_ERR_ARG = 1,      /* null pointer, bad size or count */
  BGX_ERR_TOPOLOGY = 2, /* topology id not compiled into this library */
  BGX_ERR_HIP = 3,      /* HIP runtime error (allocation, launch) */
  BGX_ERR_DISABLED = 4, /* entry point switched off at compile time */
};

/* per-env status bits written by bgx_step / bgx_rollout */
enum {
  BGX_STATUS_SLOW_PATH = 1, /* more active contacts than the fast slots */
  BGX_STATUS_NONFINITE = 2, /* q or qd not finite after the call */
};

typedef struct bgx_model bgx_model;

/* Library version (major * 10000 + minor * 100 + patch). */
int bgx_version(void);

/* Message of the last error on this thread ("" if none). */
const char* bgx_last_error(void);

/* Topologies compiled into the library.  A topology is the static structure
 * of a System (link types and parents, actuator and limit dof maps, contact
 * candidates): the reference bakes the same fields into its compiled program
 * (brax/base.py:471-480).  dims receives 8 ints:
 * {num_links, q_size, qd_size, act_size, limit_rows, contact_candidates,
 *  fast_contact_slots, model_table_floats}. */
int bgx_num_topologies(void);
const char* bgx_topology_name(int topo_id);
const char* bgx_topology_signature(int topo_id);
int bgx_topology_dims(int topo_id, int32_t* dims);
/* 1 if the topology takes one model table per env (bgx_model_create with
 * k tables for up to k envs), 0 if one table is shared, -1 if unknown. */
int bgx_topology_per_env(int topo_id);
/* The fast-contact-slot counts (LDS tiles) the topology's step kernels are
 * compiled for, default first, into tiles (up to 3 entries); returns the
 * count, -1 if the id is unknown. */
int bgx_topology_tiles(int topo_id, int32_t* tiles);

/* Uploads a model table (float32, layout of brax_amd/codegen.py pack_model)
 * for a topology; for a per-env topology, k tables back to back (env e uses
 * table e).  Replaces brax.io.mjcf.load's output feeding the pipeline
 * (brax/io/mjcf.py:317-489): the System arrays the step reads. */
int bgx_model_create(int topo_id, const float* host_table, int64_t n_floats,
                     bgx_model** out);
int bgx_model_destroy(bgx_model* model);
int bgx_model_topology(const bgx_model* model);

/* Launch shape of a model's kernels: envs_per_workgroup lanes (1, 2, 4, ...,
 * 64; 0 = 64) per workgroup for the step kernels (bgx_init, bgx_step,
 * bgx_rollout, bgx_step_probe) or the adjoint kernels (bgx_rollout_adjoint,
 * its recompute included).  Fewer lanes per workgroup spread a small batch
 * over more compute units; results are bitwise the same for every choice.
 * No counterpart in the reference, where XLA picks the launch shape; the
 * autotuner (brax_amd/generalized/autotune.py) sets it. */
#define BGX_KERNELS_STEP 0
#define BGX_KERNELS_ADJOINT 1
/* kernels = BGX_KERNELS_TILE sets the step kernels' LDS tile instead: the
 * fast contact slots per env (one of bgx_topology_tiles; 0 = the default,
 * dims[6]); the autotuner times the compiled tiles.  Results agree within
 * rounding across tiles (an env with more active contacts than slots takes
 * the slot-per-candidate path). */
#define BGX_KERNELS_TILE 2
/* kernels = BGX_KERNELS_LANES sets the step kernels' lanes per env: 1 runs
 * one thread per env; the topology's multi-lane width (the default, 0,
 * where the topology is compiled with the multi-lane tree kernel) runs an
 * env on a group of lanes (a lane per dof).  The two agree within rounding;
 * the getter returns the width in use. */
#define BGX_KERNELS_LANES 3
/* kernels = BGX_KERNELS_ADJOINT_PARTS splits bgx_rollout_adjoint's backward
 * pass into that many parts of whole envs (1-4; 0 = one), each on its own
 * stream forked from and joined back into the caller's (streams and events
 * created on a device's first split call, default priority, kept until the
 * process ends).  Results are bitwise the same for every split; the getter
 * returns the setting.  The split is faster only while its streams get
 * hardware queues of their own (few other streams in the process); where
 * they share queues it is slower.  bgx_rollout_adjoint_noalloc always runs
 * one part on the caller's stream. */
#define BGX_KERNELS_ADJOINT_PARTS 4
int bgx_model_set_launch(bgx_model* model, int kernels, int32_t envs_per_workgroup);
/* the current setting (64 for 0), or -1 on bad arguments */
int bgx_model_get_launch(const bgx_model* model, int kernels);

/* pipeline.init (brax/generalized/pipeline.py:32-61): world link frames x
 * and velocities xd of N envs from q (N, NQ) and qd (N, NV).  Outputs:
 * x_pos (N, L, 3), x_rot (N, L, 4), xd_ang (N, L, 3), xd_vel (N, L, 3). */
int bgx_init(const bgx_model* model, int32_t n_envs, const float* q,
             const float* qd, float* x_pos, float* x_rot, float* xd_ang,
             float* xd_vel, void* stream);

/* pipeline.step (brax/generalized/pipeline.py:64-94) applied n_frames times
 * with one action, as PipelineEnv.pipeline_step does (brax/envs/base.py:
 * 128-137).  q (N, NQ) and qd (N, NV) are updated in place; act (N, NU) may
 * be NULL (zero action).  x_* / xd_* (as in bgx_init) receive the frames at
 * the new state when non-NULL; status (N) receives BGX_STATUS_* bits when
 * non-NULL. */
int bgx_step(const bgx_model* model, int32_t n_envs, int32_t n_frames,
             float* q, float* qd, const float* act, float* x_pos,
             float* x_rot, float* xd_ang, float* xd_vel, int32_t* status,
             void* stream);

/* n_steps env steps in one launch: step t applies act + t * act_step_stride
 * (each (N, NU)) for n_frames substeps.  q_traj (n_steps, N, NQ) and qd_traj
 * (n_steps, N, NV) receive the state after every env step when non-NULL.
 * Other arguments as bgx_step.  The launch holds no host synchronization, so
 * it can be captured in a HIP graph. */
int bgx_rollout(const bgx_model* model, int32_t n_envs, int32_t n_steps,
                int32_t n_frames, float* q, float* qd, const float* act,
                int64_t act_step_stride, float* q_traj, float* qd_traj,
                float* x_pos, float* x_rot, float* xd_ang, float* xd_vel,
                int32_t* status, void* stream);

/* Every contact candidate of the states q (N, NQ), as the reference's
 * debug contact output (brax/generalized/pipeline.py:58-59, 91-92; MJX plane
 * collisions, brax/contact.py:28-67): dist (N, C), pos (N, C, 3) and frame
 * (N, C, 3, 3) with rows normal, t1, t2, for the C = dims[5] candidates in
 * the topology's order.  A topology without candidates writes nothing. */
int bgx_contacts(const bgx_model* model, int32_t n_envs, const float* q, float* dist,
                 float* pos, float* frame, void* stream);

/* Reverse-mode derivative of pipeline.init's frames (kinematics.forward,
 * brax/kinematics.py:31-108), as jax.grad differentiates the reference's
 * State.x / State.xd: gq (N, NQ) and gqd (N, NV) receive the cotangents of
 * q and qd given those of x_pos (N, L, 3), x_rot (N, L, 4), xd_ang (N, L, 3)
 * and xd_vel (N, L, 3) (each NULL = zero) at the states (q, qd).  Hinge-chain
 * topologies return BGX_ERR_TOPOLOGY. */
int bgx_frames_vjp(const bgx_model* model, int32_t n_envs, const float* q, const float* qd,
                   const float* gx_pos, const float* gx_rot, const float* gxd_ang,
                   const float* gxd_vel, float* gq, float* gqd, void* stream);

/* Diagnostic variant of bgx_step: probe (N, 2 * NV + 3) receives, for the
 * last substep, qf_smooth (dynamics.forward), qdd (integrator), and the
 * active contact count, solver iterations and line-search trials. */
int bgx_step_probe(const bgx_model* model, int32_t n_envs, int32_t n_frames,
                   float* q, float* qd, const float* act, float* probe,
                   void* stream);

/* Workspace bytes for bgx_rollout_adjoint with n_envs, a checkpoint every
 * ckpt_every env steps and n_frames substeps per env step: the preferred
 * size (the tapes of a whole checkpoint segment, up to 6 GiB of them, so
 * one launch tapes all its substeps), and the least size (one substep's
 * tape per launch).  Any size from the least up works; -1 on bad
 * arguments. */
int64_t bgx_adjoint_workspace_bytes(const bgx_model* model, int32_t n_envs,
                                    int32_t ckpt_every, int32_t n_frames);
int64_t bgx_adjoint_workspace_min_bytes(const bgx_model* model, int32_t n_envs,
                                        int32_t ckpt_every, int32_t n_frames);

/* Reverse-mode gradient of a rollout (bgx_rollout with n_steps env steps of
 * n_frames substeps), replacing jax.grad through the reference's scan of
 * pipeline.step (brax/envs/base.py:128-137, differentiated as in
 * brax/training/agents/apg/train.py:188-198; the solver is differentiated
 * unrolled with its decisions fixed, as jaxopt implicit_diff=False does at
 * brax/generalized/constraint.py:229-238).
 *   ckpt_q (n_ckpt, N, NQ), ckpt_qd (n_ckpt, N, NV): the state before env
 *     steps 0, ckpt_every, 2 ckpt_every, ... (n_ckpt = ceil(n_steps /
 *     ckpt_every)); each segment is recomputed from its checkpoint.
 *   act: as bgx_rollout (NULL = zero action).
 *   gq_traj (n_steps, N, NQ), gqd_traj (n_steps, N, NV): cotangents of the
 *     state after each env step (NULL = zero).
 *   Outputs: gact (n_steps, N, NU), gq0 (N, NQ), gqd0 (N, NV): cotangents of
 *     the actions and of the initial state.  status (N, may be NULL)
 *     receives BGX_STATUS_* bits of the recomputed substeps.
 * workspace: device memory of workspace_bytes bytes, at least
 * bgx_adjoint_workspace_min_bytes().  The
 * call enqueues kernels only (no allocation, no synchronization).
 * Topologies with per-env model tables return BGX_ERR_TOPOLOGY. */
int bgx_rollout_adjoint(const bgx_model* model, int32_t n_envs, int32_t n_steps,
                        int32_t n_frames, int32_t ckpt_every,
                        const float* ckpt_q, const float* ckpt_qd,
                        const float* act, int64_t act_step_stride,
                        const float* gq_traj, const float* gqd_traj,
                        float* gact, float* gq0, float* gqd0, float* workspace,
                        int64_t workspace_bytes, int32_t* status,
                        void* stream);

/* A heterogeneous batch: envs of several models stepped by one persistent
 * kernel launch, replacing the reference's vmap over per-env System arrays
 * (brax/envs/wrappers/training.py:223-260; one compiled program per
 * structure there).  models[m] contributes envs_per_model[m] envs (a per-env
 * model takes its first tables); the models may be hinge chains of any
 * length (per-env tables) and models of at most one other topology, which
 * must have a batch kernel.  The batch's arrays are packed model after
 * model, env after env: q holds sizes[1] floats (model m's envs as
 * (n_m, NQ_m) rows), qd sizes[2], a step's actions sizes[3].  Creation
 * allocates and uploads the device work table (synchronously).  sizes
 * receives {envs, q floats, qd floats, action floats per step, work items,
 * workgroups}. */
typedef struct bgx_batch bgx_batch;
int bgx_batch_create(int32_t n_models, const bgx_model* const* models,
                     const int32_t* envs_per_model, bgx_batch** out);
int bgx_batch_destroy(bgx_batch* batch);
int bgx_batch_sizes(const bgx_batch* batch, int64_t* sizes);

/* n_steps env steps of n_frames substeps of every env of the batch in one
 * launch (bgx_rollout for a batch): q, qd (packed) updated in place; step t
 * applies act + t * act_step_stride (packed, NULL = zero action);
 * q_traj (n_steps, q floats) and qd_traj receive the state after every env
 * step when non-NULL; status (envs) the BGX_STATUS_NONFINITE bit per env.
 * Each env's results are bitwise those of bgx_rollout on its model alone.
 * The batch's work counter is shared by its launches: launches of one batch
 * must not overlap (one stream).  Graph-capturable. */
int bgx_batch_rollout(const bgx_batch* batch, int32_t n_steps, int32_t n_frames, float* q,
                      float* qd, const float* act, int64_t act_step_stride, float* q_traj,
                      float* qd_traj, int32_t* status, void* stream);

/* Bytes of the per-GPU scratch array of bgx_rollout_adjoint_noalloc. */
#define BGX_ADJOINT_SCRATCH_BYTES (128ll << 20)
int64_t bgx_adjoint_scratch_bytes(void);

/* bgx_rollout_adjoint without a workspace argument, for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid allocation).  No counterpart in the reference, whose
 * jax.grad allocates through XLA.  Its workspace is a static device array of
 * BGX_ADJOINT_SCRATCH_BYTES on each GPU the library runs on (held until the
 * process ends); the call makes no allocation.  Arguments, outputs and
 * results as bgx_rollout_adjoint (bitwise the same gradients); a problem
 * whose least workspace (bgx_adjoint_workspace_min_bytes) exceeds the array
 * returns BGX_ERR_ARG.  Runs on the current device.  Every call on a GPU
 * shares that GPU's array: calls that overlap on one GPU (other streams or
 * threads) race and must be ordered by the caller. */
int bgx_rollout_adjoint_noalloc(const bgx_model* model, int32_t n_envs, int32_t n_steps,
                                int32_t n_frames, int32_t ckpt_every,
                                const float* ckpt_q, const float* ckpt_qd,
                                const float* act, int64_t act_step_stride,
                                const float* gq_traj, const float* gqd_traj,
                                float* gact, float* gq0, float* gqd0, int32_t* status,
                                void* stream);

#ifdef __cplusplus
}
#endif

#endif /* BRAX_GEN_H_ */
