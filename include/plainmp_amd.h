/*
 * plainmp_amd: C ABI of the MI355X (gfx950) FK + sphere-SDF collision path.
 *
 * The ABI is the drop-in boundary under plainmp's C++ classes: each batched
 * entry point replaces a per-configuration reference method (cited next to
 * it, paths relative to the plainmp source tree) with one call over a batch.
 *
 * Conventions
 *   - Every function returns a pmp_status; PMP_OK is 0.  On failure a
 *     thread-local message is available from pmp_last_error().
 *     PMP_ERR_INVALID maps to std::invalid_argument / ValueError,
 *     PMP_ERR_RUNTIME and PMP_ERR_HIP to std::runtime_error / RuntimeError
 *     (the reference's error types, constraints/primitive_sphere_collision.cpp:171-174,
 *     kinematics/kinematics.cpp:244,271).
 *   - Descriptor structs (pmp_robot_desc, pmp_scene_desc) hold HOST pointers;
 *     the tables are copied to device memory owned by the context.
 *   - Batch inputs and outputs are DEVICE pointers owned by the caller.
 *     Configurations are FP64 structure-of-arrays: q[col * ldq + b] for
 *     column col of configuration b (ldq >= B).  Outputs use the same
 *     "batch index fastest" layout with leading dimension B.
 *   - Work is enqueued on the context's stream (pmp_set_stream); calls do
 *     not synchronise the host unless documented.  Calls on one context
 *     share its device workspaces: a call made on another stream than the
 *     previous call waits (on the device, through an event) for that call's
 *     work, so consecutive calls never overlap.  Calls recorded into a
 *     caller's own stream capture are not ordered this way.
 *   - A context is not re-entrant: use one context per host thread.
 */
#ifndef PLAINMP_AMD_H_
#define PLAINMP_AMD_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int32_t pmp_status;
#define PMP_OK 0
#define PMP_ERR_INVALID 1
#define PMP_ERR_RUNTIME 2
#define PMP_ERR_HIP 3

/* Kinematic node kinds (kinematics/algorithm.cpp:109-158, 160-179). */
#define PMP_NODE_ROOT 0        /* the base link; pose = base pose            */
#define PMP_NODE_FIXED 1       /* static local transform                     */
#define PMP_NODE_REVOLUTE 2    /* revolute or continuous joint driven by q    */
#define PMP_NODE_PRISMATIC 3   /* prismatic joint driven by q                 */
#define PMP_NODE_TRANSLATION 4 /* consider_rotation = false link (position only) */

/* Base types (kinematics/kinematics.hpp:49). */
#define PMP_BASE_FIXED 0
#define PMP_BASE_PLANAR 1
#define PMP_BASE_FLOATING 2

/* Primitive types (collision/primitive_sdf.hpp:77). */
#define PMP_PRIM_BOX 1
#define PMP_PRIM_CYLINDER 2
#define PMP_PRIM_SPHERE 3
#define PMP_PRIM_GROUND 4
#define PMP_PRIM_CLOUD 5

/* Motion validator types (ompl/ompl_thin_wrap.hpp:50-59). */
#define PMP_VALIDATOR_BOX 0
#define PMP_VALIDATOR_EUCLIDEAN 1

/*
 * Kinematic tree + sphere tables of a SphereCollisionCst.
 * Nodes are in evaluation order: every node's parent precedes it.  Node 0
 * must be the root.  Replaces the state that KinematicModel
 * (kinematics/kinematics.hpp:51-169) and SphereCollisionCst
 * (constraints/primitive_sphere_collision.hpp:55-116) build in their ctors.
 */
typedef struct pmp_robot_desc {
  int32_t n_nodes;
  const int32_t* node_parent;        /* [n_nodes] parent node, -1 for the root          */
  const int32_t* node_kind;          /* [n_nodes] PMP_NODE_*                           */
  const int32_t* node_q_col;         /* [n_nodes] column of q for joints, else -1       */
  const double* node_quat;           /* [n_nodes*4] local rotation x,y,z,w              */
  const double* node_trans;          /* [n_nodes*3] local translation                  */
  const double* node_axis;           /* [n_nodes*3] joint axis (not normalised)        */
  const uint8_t* node_identity;      /* [n_nodes] local rotation skipped when composing */
  int32_t n_q_cols;                  /* joint columns of q (= controlled joints)       */
  int32_t base_type;                 /* PMP_BASE_*; PLANAR appends (x, y, yaw) and
                                        FLOATING (x, y, z, roll, pitch, yaw) columns   */
  double base_pose[7];               /* x, y, z, qx, qy, qz, qw when base_type is FIXED */
  int32_t n_spheres;                 /* in SphereCollisionCst group order              */
  const int32_t* sphere_node;        /* [n_spheres] node the sphere is attached to      */
  const double* sphere_local;        /* [n_spheres*3] position in the node frame        */
  const double* sphere_radius;       /* [n_spheres]                                    */
  const int32_t* sphere_group;       /* [n_spheres] group index                         */
  int32_t n_groups;
  const uint8_t* group_only_self;    /* [n_groups] 1 = skipped by external checks      */
  int32_t n_self_pairs;              /* group pairs checked for self collision         */
  const int32_t* self_pairs;         /* [n_self_pairs*2] group indices                 */
} pmp_robot_desc;

/*
 * One primitive of the flattened SDF list (all_sdfs_cache_,
 * constraints/primitive_sphere_collision.cpp:426-448).  Pose, flags and AABB
 * are given as the reference's Pose / update_aabb compute them
 * (collision/primitive_sdf.hpp:30-75, 250-255, 400-421, 500-521, 558-561).
 */
typedef struct pmp_prim_desc {
  int32_t type;             /* PMP_PRIM_*                                            */
  uint8_t axis_aligned;
  uint8_t z_axis_aligned;
  double position[3];
  double rotation[9];       /* row-major, world <- local                            */
  double rotation_inv[9];   /* row-major general inverse of rotation                 */
  double size[3];           /* box: width; cylinder: radius, height; sphere: radius;
                               ground: height; cloud: point radius                   */
  double lb[3], ub[3];      /* axis-aligned bounds                                   */
  const double* points;     /* cloud only: [n_points*3] xyz                          */
  int64_t n_points;
} pmp_prim_desc;

typedef struct pmp_scene_desc {
  int32_t n_prims;
  const pmp_prim_desc* prims;
} pmp_scene_desc;

/* Box / euclidean motion validator (ompl/motion_validator.hpp:42-63). */
typedef struct pmp_validator_desc {
  int32_t type;             /* PMP_VALIDATOR_*                                      */
  double resolution;        /* EUCLIDEAN                                            */
  const double* box_width;  /* BOX: [n_q_cols] host pointer                         */
} pmp_validator_desc;

/* Launch schedule of the persistent kernels, chosen by the autotuner. */
typedef struct pmp_schedule {
  int32_t block;            /* threads per workgroup (multiple of 64), 0 = default  */
  int32_t waves_per_cu;     /* persistent waves per compute unit, 0 = default       */
  int32_t eval_whole_chunks;  /* evaluate / trajectory batches of at least this many
                                 64-configuration chunks walk each configuration as one
                                 unit, smaller ones split it per sphere block and self
                                 pair; 0 = default (512)                               */
  int32_t pair_shares;      /* pair distances: the primitives are split into this many
                               shares, each walked by its own waves (more stores in
                               flight for a small batch); 0 = default (about four
                               waves per SIMD), clamped to half the primitive count */
} pmp_schedule;

typedef struct pmp_context pmp_context;
/* A captured HIP graph of one of the *_graph entry points below, replayed by
 * pmp_graph_launch and freed by pmp_graph_destroy. */
typedef struct pmp_graph pmp_graph;

const char* pmp_last_error(void);
/* 2: pmp_schedule gained pair_shares (a caller built against version 1
   passes a shorter struct to pmp_set_schedule). */
int32_t pmp_abi_version(void);

pmp_status pmp_create(int32_t device, pmp_context** out);
pmp_status pmp_destroy(pmp_context* ctx);
/* hip_stream is a hipStream_t (NULL = default stream). */
pmp_status pmp_set_stream(pmp_context* ctx, void* hip_stream);
pmp_status pmp_synchronize(pmp_context* ctx);

/* Tables.  Replace KinematicModel(xml) + SphereCollisionCst ctor + set_sdf
 * (kinematics/kinematics.cpp:183-223, constraints/primitive_sphere_collision.cpp:117-179,
 *  constraints/primitive_sphere_collision.hpp:73-76). */
pmp_status pmp_upload_robot(pmp_context* ctx, const pmp_robot_desc* robot);
pmp_status pmp_upload_scene(pmp_context* ctx, const pmp_scene_desc* scene);
pmp_status pmp_set_band(pmp_context* ctx, double band_m);      /* FP64 re-decision band, default 1e-5 m */
pmp_status pmp_set_schedule(pmp_context* ctx, const pmp_schedule* schedule);

/* World centres of all spheres, centers[(s*3 + k) * B + b] (FP32).
 * Replaces update_kintree + get_all_spheres
 * (constraints/primitive.hpp:38-62, constraints/primitive_sphere_collision.cpp:412-424).
 * accurate = 0 uses the Taylor sin/cos of is_valid, 1 exact sin/cos. */
pmp_status pmp_fk_spheres_batch(pmp_context* ctx, const double* q, int64_t ldq, int64_t B, int32_t accurate,
                                float* centers);

/* Jacobian of a point fixed to one link with respect to q, rotation
 * ignored, with exact FK: jac[(k * q_dim + j) * B + b] for k < 3 (FP32).
 * link: the link's node index (get_link_ids numbering, as pmp_link_poses_batch);
 * point: HOST [3] in the link frame, NULL for the link origin.  Replaces
 * KinematicModel::get_jacobian / get_attached_point_jacobian with
 * RotationType::IGNORE (kinematics/algorithm.cpp:212-433), the Jacobian of
 * LinkPoseCst for a position target (constraints/primitive_link_pose.cpp:32-44). */
pmp_status pmp_link_jacobian_batch(pmp_context* ctx, const double* q, int64_t ldq, int64_t B, int32_t link,
                                   const double* point, float* jac);

/* One damped Gauss-Newton update of R restarts of a multi-start IK, a lane
 * per restart, in place on q[j * R + r] (FP64, device).  No counterpart in
 * the reference, whose solve_ik (src/plainmp/ik.py:42-156) runs SLSQP one
 * seed after another on the LinkPoseCst residual
 * (constraints/primitive_link_pose.cpp:32-57): here the residual rows are
 * that constraint's (position; + getRPY() - rpy or quaternion coefficients
 * for 6- or 7-vector targets, with rpy_derivative / q_derivative rows,
 * kinematics/algorithm.cpp:21-47) plus collision values below `margin` as
 * penalty rows; the step solves (A^T A + damping I) s = -A^T r (Cholesky,
 * FP64), is scaled to at most max_step in the max norm and projected on
 * [lo, hi].  Inputs are device arrays of this iteration: poses from
 * pmp_link_poses_batch, jac0 from pmp_link_jacobian_batch at the link origin
 * and, for kind 6 / 7, jac_u[k] at the link's unit axes e_k; vals / cjac from
 * pmp_evaluate_batch (n_cst rows, or NULL).  q_dim <= 16. */
typedef struct pmp_ik_step_desc {
  int32_t kind;          /* 3, 6 or 7 */
  int32_t link;          /* node index of the target link */
  const double* target;  /* device [kind] */
  const float* jac0;
  const float* jac_u[3];
  const float* vals;
  const float* cjac;
  int32_t n_cst;
  const double* lo;      /* device [q_dim] */
  const double* hi;
  double damping, max_step, margin, penalty;
} pmp_ik_step_desc;
pmp_status pmp_ik_step_batch(pmp_context* ctx, double* q, int64_t R, const double* poses,
                             const pmp_ik_step_desc* step);

/* n_iter whole iterations of that IK, captured once into a HIP graph and
 * replayed by pmp_graph_launch: per iteration the link Jacobians into
 * step->jac0 and, for kind 6 / 7, step->jac_u[k] (as pmp_link_jacobian_batch
 * at the link origin and at its unit axes e_k, ldq = R), the collision rows
 * into step->vals / step->cjac when n_cst > 0 (as pmp_evaluate_batch), then
 * the step on q.  The step reads the link pose from the FP32 walk of the
 * first Jacobian launch (into a buffer of the graph's own) instead of FP64
 * link poses.  Those buffers are written (device memory the caller owns,
 * valid while the graph exists); the context rules of pmp_check_edges_graph
 * apply.  Replaces the host loop over iterations of the batched IK (the
 * reference's counterpart is the SLSQP loop of solve_ik,
 * src/plainmp/ik.py:42-156). */
pmp_status pmp_ik_graph(pmp_context* ctx, double* q, int64_t R, const pmp_ik_step_desc* step, int32_t n_iter,
                        pmp_graph** out);

/* Link poses of all nodes, poses[(n*7 + k) * B + b] as x,y,z,qx,qy,qz,qw
 * (FP64).  Replaces set_joint_angles + get_link_pose
 * (kinematics/algorithm.cpp:88-107, kinematics/kinematics.hpp:131-136). */
pmp_status pmp_link_poses_batch(pmp_context* ctx, const double* q, int64_t ldq, int64_t B, int32_t accurate,
                                double* poses);

/* Per-sphere minimum signed clearance over all primitives,
 * clearance[s * B + b] = min_p (sdf_p(c_s) - r_s) (FP32), FK mode as above. */
pmp_status pmp_sphere_clearance_batch(pmp_context* ctx, const double* q, int64_t ldq, int64_t B,
                                      int32_t accurate, float* clearance);

/* Validity of each configuration: valid[b] in {0, 1}.
 * Replaces IneqConstraintBase::is_valid -> SphereCollisionCst::is_valid_dirty
 * (constraints/primitive.hpp:113-117, constraints/primitive_sphere_collision.cpp:181-254).
 * FP32 with every decision within the band re-decided in FP64 on the GPU.
 * n_recheck (device, optional) accumulates the number of re-decided configurations. */
pmp_status pmp_is_valid_batch(pmp_context* ctx, const double* q, int64_t ldq, int64_t B, uint8_t* valid,
                              int32_t* n_recheck);

/* pmp_is_valid_batch without any allocation: no hipMalloc, hipFree or
 * allocating library call, for callers that cannot allocate (code under HIP
 * graph capture, tight serving loops).  No counterpart in the reference.
 * Its workspaces come from a static device array of 32 MiB that every GPU
 * the library runs on holds until the process ends; a batch that needs more
 * (scenes with a point cloud: 28 B per primitive-checked sphere and
 * configuration plus 4 B per configuration) is refused with PMP_ERR_INVALID.
 * All calls on one GPU share that array: calls that overlap on one GPU (from
 * several contexts, or captured into graphs that replay concurrently) race. */
pmp_status pmp_is_valid_batch_noalloc(pmp_context* ctx, const double* q, int64_t ldq, int64_t B, uint8_t* valid,
                                      int32_t* n_recheck);

/* Reference-semantics value and Jacobian rows, vals[k * B + b],
 * jac[(k * q_dim + j) * B + b] for k < cst_dim (ext first, then self).
 * Replaces ConstraintBase::evaluate -> SphereCollisionCst::evaluate_dirty
 * (constraints/primitive.hpp:74-79, constraints/primitive_sphere_collision.cpp:256-399). */
pmp_status pmp_evaluate_batch(pmp_context* ctx, const double* q, int64_t ldq, int64_t B, float* vals,
                              float* jac);

/* Trajectories: X holds the waypoints of n_traj trajectories back to back
 * (columns as q, leading dimension ldx, N waypoints in all); trajectory i
 * spans waypoints [offsets[i], offsets[i+1]) (HOST array of n_traj + 1,
 * offsets[0] = 0, non-decreasing, offsets[n_traj] = N).  One launch writes
 *   vals / jac  the evaluate rows of every waypoint, laid out as
 *               pmp_evaluate_batch with B = N;
 *   cost[i]     sum over interior t of sum_j w_j^2 (x_{t-1,j} - 2 x_{t,j} + x_{t+1,j})^2
 *               (FP32, device, n_traj);
 *   grad        grad[j * N + n] = d cost / d x_{n,j} (FP32, device).
 * weights: HOST [q_dim], NULL for ones.  Replaces SequentialCst::evaluate for
 * the collision constraint (constraints/sequential_constraint.cpp:63-91, without
 * its 1e-12 sparsity marker) and the smoothness term x^T P x with
 * P = smoothcost_fullmat (src/plainmp/nlp_solver/sqp_based_solver.py:74-96),
 * gradient 2 P x. */
pmp_status pmp_traj_cost_grad_batch(pmp_context* ctx, const double* X, int64_t ldx, int64_t N, const int64_t* offsets,
                                    int32_t n_traj, const double* weights, float* vals, float* jac, float* cost,
                                    float* grad);

/* The SQP solver's inequality sequence in its sparse form: SequentialCst::evaluate
 * (constraints/sequential_constraint.cpp:63-127) of the collision constraint added
 * at every waypoint plus, when msbox is given, the motion-step box constraint
 * (the sequence src/plainmp/nlp_solver/sqp_based_solver.py:48-71 builds), for
 * n_traj trajectories at once (X, ldx, N, offsets as pmp_traj_cost_grad_batch).
 * Trajectory i with T_i waypoints has rows_i = cst_dim * T_i + 2 q_dim (T_i - 1)
 * (no box rows without msbox; 0 rows when T_i = 0) and q_dim * T_i columns:
 * the collision rows of every waypoint (t-major), then q_t - q_{t+1} + w and
 * q_{t+1} - q_t + w per segment.  Its values go to c[row_off[i] ...], its
 * Jacobian as CSC to colptr[colptr_off[i] ...] (q_dim T_i + 1 int32 column
 * pointers counted from the trajectory's first entry), rowind / values
 * [nnz_off[i] ...] (int32 row indices ascending within a column, FP64 values:
 * collision entries J + 1e-12 as the reference writes them, box entries +-1).
 * pmp_traj_sqp_sizes fills the HOST arrays row_off, nnz_off, colptr_off
 * [n_traj + 1] (the last entry is the total); X and the outputs are device
 * pointers, offsets and msbox ([q_dim] or NULL) host arrays. */
pmp_status pmp_traj_sqp_sizes(pmp_context* ctx, const int64_t* offsets, int32_t n_traj, int32_t with_msbox,
                              int64_t* row_off, int64_t* nnz_off, int64_t* colptr_off);
pmp_status pmp_traj_sqp_ineq(pmp_context* ctx, const double* X, int64_t ldx, int64_t N, const int64_t* offsets,
                             int32_t n_traj, const double* msbox, double* c, int32_t* colptr, int32_t* rowind,
                             double* values);

/* Signed distance of every (sphere, primitive) pair with exact FK and its
 * derivative with respect to q: dist[(s*P + p) * B + b],
 * grad[((s*P + p) * q_dim + j) * B + b].  grad may be NULL.  No reference
 * counterpart as one call: it batches the per-pair terms that
 * evaluate_ext_collision reduces (constraints/primitive_sphere_collision.cpp:256-322). */
pmp_status pmp_pair_clearance_batch(pmp_context* ctx, const double* q, int64_t ldq, int64_t B, float* dist,
                                    float* grad);

/* Adjoint of pmp_pair_clearance_batch: grad_q[j * B + b] =
 * sum_{s,p} grad_dist[(s*P + p) * B + b] * d dist[s,p,b] / d q_j. */
pmp_status pmp_pair_clearance_vjp(pmp_context* ctx, const double* q, int64_t ldq, int64_t B,
                                  const float* grad_dist, float* grad_q);

/* Edge validity: valid[e] for the motion s1[e] -> s2[e] (columns as q,
 * leading dimension lds).  Replaces CustomValidatorBase::checkMotion with the
 * box or euclidean step ratio (ompl/motion_validator.cpp:17-126).
 * n_checked (device, optional): states checkMotion evaluates in its visiting
 * order, i.e. up to and including the first invalid one (0 for an infinite
 * step ratio, -1 for an edge needing more than 2^24 states, which is reported
 * invalid). */
pmp_status pmp_check_edges_batch(pmp_context* ctx, const double* s1, const double* s2, int64_t lds, int64_t E,
                                 const pmp_validator_desc* validator, uint8_t* valid, int32_t* n_checked);

/* Same as pmp_check_edges_batch, captured once into a HIP graph and replayed
 * by pmp_graph_launch on the context stream (with none set, on a stream of
 * the context's own that the null stream synchronizes with).  The pointers must stay valid
 * while the graph exists, and the graph must be destroyed before its context.
 * The graph reads the context's robot, scene and cull tables as they were at
 * capture: after pmp_upload_robot, pmp_upload_scene or pmp_set_band (e.g. a
 * set_sdf of the host mirror) pmp_graph_launch returns PMP_ERR_INVALID and
 * the graph has to be captured again. */
pmp_status pmp_check_edges_graph(pmp_context* ctx, const double* s1, const double* s2, int64_t lds, int64_t E,
                                 const pmp_validator_desc* validator, uint8_t* valid, int32_t* n_checked,
                                 pmp_graph** out);
pmp_status pmp_graph_launch(pmp_graph* graph);
pmp_status pmp_graph_destroy(pmp_graph* graph);
/* Same as pmp_traj_cost_grad_batch, captured once into a HIP graph (its
 * offsets and weights copied into the graph's own device buffers) and
 * replayed by pmp_graph_launch: trajectories of uneven length in one launch
 * per replay, the loop of SequentialCst::evaluate over waypoints
 * (constraints/sequential_constraint.cpp:63-127) for every trajectory.  X and
 * the outputs must stay valid while the graph exists; the context rules of
 * pmp_check_edges_graph apply (a table change makes launch refuse). */
pmp_status pmp_traj_cost_grad_graph(pmp_context* ctx, const double* X, int64_t ldx, int64_t N, const int64_t* offsets,
                                    int32_t n_traj, const double* weights, float* vals, float* jac, float* cost,
                                    float* grad, pmp_graph** out);
/* Duration of the FP32 edge kernel of the most recent pmp_check_edges_batch
 * on this context, from HIP events recorded around it on the context stream
 * (waits for that kernel).  No reference counterpart: it serves throughput
 * measurement. */
pmp_status pmp_edge_kernel_ms(pmp_context* ctx, float* ms);
/* Edges the most recent edge batch on this context re-decided in FP64
 * (waits for it).  No reference counterpart: a diagnostic of the band. */
pmp_status pmp_edge_recheck_count(pmp_context* ctx, int32_t* n);

/* Primitive SDF of points (FP64): out[i] = min over the given primitives of
 * the signed distance at points[k * N + i] (mode 0), or is_outside(p, radius)
 * of the union as 0/1 (mode 1).  Replaces SDFBase::evaluate / evaluate_batch /
 * is_outside (collision/primitive_sdf.hpp:83-97, 147-170). */
pmp_status pmp_prims_evaluate(pmp_context* ctx, const pmp_scene_desc* prims, const double* points, int64_t N,
                              int32_t mode, double radius, double* out);

/* Exact nearest neighbour by brute force (FP64): for each query
 * queries[k * M + i], sqdist[i] and index[i] of the nearest of the N points
 * points[k * N + n].  Replaces KDTree::query / sqdist (collision/kdtree.cpp:26-38). */
pmp_status pmp_nearest_batch(pmp_context* ctx, const double* points, int64_t N, const double* queries, int64_t M,
                             double* sqdist, int32_t* index);

/* Sizes of the tile tables of a cloud of n_points points (1 <= n_points <
 * 2^30): tiles of 64 points, super-tiles of 16 tiles, and the scratch bytes
 * pmp_cloud_build_tiles needs.  Host values. */
pmp_status pmp_cloud_tile_sizes(int64_t n_points, int64_t* n_tiles, int64_t* n_supers, int64_t* scratch_bytes);

/* The spatial structure of a point cloud built on the device from
 * device-resident points (FP64, points[3 * i + k]): the points in Morton order
 * in tiles of 64 with their bounds and the matrix-core operands the cloud
 * distance kernels read, and super-tiles of 16 tiles (the tables
 * pmp_upload_scene builds for each cloud primitive).  Replaces the KD-tree
 * build of CloudSDF (collision/kdtree.cpp:40-65, collision/primitive_sdf.hpp:568-583).
 * Device outputs, per tile: 48 B at tiles (FP32 lb[3], ub[3], centre[3],
 * int32 count), 48 B at tiles64 (FP64 lb[3], ub[3]), 256 floats at tile_b
 * (for point 16 m + j: tile_b[64 m + 16 k + j] = -2 p'_k (k < 3) and |p'|^2
 * (k = 3), p' = p - centre rounded to FP32; padding 0, 0, 0, 1e30) and 256
 * doubles at points64 (x, y, z, 0 per point; padding 1e30); per super-tile
 * 32 B at supers (FP32 lb[3], ub[3], int32 first tile, tile count).
 * scratch: scratch_bytes of device memory.  Asynchronous on the context's
 * stream. */
pmp_status pmp_cloud_build_tiles(pmp_context* ctx, const double* points, int64_t n_points, void* tiles,
                                 void* tiles64, float* tile_b, double* points64, void* supers, void* scratch);

/* Number of spheres, primitives, q columns and constraint rows of the
 * uploaded tables (host values). */
pmp_status pmp_query_sizes(pmp_context* ctx, int32_t* n_spheres, int32_t* n_prims, int32_t* q_dim,
                           int32_t* cst_dim);

#ifdef __cplusplus
}
#endif

#endif /* PLAINMP_AMD_H_ */
