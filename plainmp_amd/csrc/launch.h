// Host-side launchers of the kernels in kernels.hip (and cloud.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "geom.hpp"
#include "tables.h"

namespace pmp {

constexpr int kPendStates = 16;  // in-band states an edge may queue for the FP64 pass
// Slots of EdgeArgs::counter.  The work counter, the pending-list length and
// the count of edges re-decided in FP64 inside edges_kernel are raised by
// edges_kernel; the last workgroup of edges_recheck_kernel publishes the
// FP64 re-decisions of the launch (listed + inline) as the recheck count and
// sets the others back to zero, so a launch leaves them as it found them
// (zero when the buffer is allocated) and no clearing node is needed in a
// captured graph.
enum EdgeCounter { kCntWork = 0, kCntRecheck = 1, kCntPending = 2, kCntDone = 3, kCntInline = 4, kNumEdgeCounters = 5 };

struct EdgeArgs {
  const double* s1;
  const double* s2;
  int64_t lds, E;
  int32_t vtype;
  double resolution;
  const double* width;      // device [n_cols]
  const double* inv_width;  // device [n_cols], 1 / width as the reference precomputes
  const uint8_t* seq;       // device [128 * 128] farthest-first sequence table
  int32_t* counter;         // device [kNumEdgeCounters], see EdgeCounter
  int32_t* pend;            // [E][kPendStates] in-band visiting indices of pending edges
  int32_t* pend_info;       // [E][2] number queued (-1: more than kPendStates), first certain collision
  int32_t* plist;           // [E] the pending edges, in the order edges_kernel finished them
  int32_t* pend_pair;       // [E][kPendStates][2] the queued state's one in-band candidate (block, primitive),
                            //   block -1 when the uncertainty is not one candidate
  uint8_t* valid;
  int32_t* n_checked;
};

// The SQP inequality sequence as CSC per trajectory from evaluate rows
// (vals [C][N], jac [C][D][N]); offsets, row_off, nnz_off, msbox on the device.
struct SqpArgs {
  const double* X;
  int64_t ldx, N;
  int D, C;
  const int64_t* offsets;  // [n_traj + 1]
  int n_traj;
  const double* msbox;     // [D] or nullptr
  const float* vals;
  const float* jac;
  const int64_t* row_off;  // [n_traj + 1]
  const int64_t* nnz_off;  // [n_traj + 1]
  double* c;
  int32_t* colptr;         // trajectory i from D * offsets[i] + i
  int32_t* rowind;
  double* values;
};
hipError_t launch_sqp_ineq(const SqpArgs& a, hipStream_t st);
// Waves of the edge kernel variant (cloud or not) one CU holds at once with
// workgroups of `block` threads (0 when the runtime cannot tell).
int edges_resident_waves_per_cu(bool cloud, int block, size_t lds);
// pose (optional): [k * B + b] the link's FP32 pose (x, y, z, qx, qy, qz, qw)
// from the same walk.
hipError_t launch_link_jacobian(const KinArgs<float>& ka, const double* q, int64_t ldq, int64_t B, int link,
                                const float (&off)[3], float* jac, float* pose, hipStream_t st);
// p[0, n) = v by a kernel (safe under stream capture, unlike a memset node).
hipError_t launch_fill_i32(int32_t* p, int32_t v, int64_t n, hipStream_t st);
hipError_t launch_centers(const KinArgs<float>& ka, const double* q, int64_t ldq, int64_t B, int taylor, float* out,
                          hipStream_t st);
hipError_t launch_poses(const KinArgs<double>& ka, const double* q, int64_t ldq, int64_t B, int taylor, double* out,
                        hipStream_t st);
struct CloudRange {
  int32_t tile_begin, n_tiles;
  float radius;
  int32_t super_begin, n_super;
};

hipError_t launch_clearance(const KinArgs<float>& ka, const SceneArgs<float>& sc, const double* q, int64_t ldq,
                            int64_t B, int taylor, float* out, float* cen, hipStream_t st);
hipError_t launch_valid(const KinArgs<float>& k32, const KinArgs<double>& k64, const SceneArgs<float>& s32,
                        const SceneArgs<double>& s64, const double* q, int64_t ldq, int64_t B, float band,
                        uint8_t* valid, int32_t* n_recheck, int block, const CloudWork& cw,
                        const CloudRange* clouds, int n_clouds, int32_t* flags, int32_t* unit_flags,
                        hipStream_t st);
// Batches up to this size take the pass with a lane per (configuration,
// evaluate unit); unit_flags: [B] int32, zero on entry and left zero.
constexpr int64_t kValidUnitsMaxBatch = 65536;
// noalloc.hip: the scratch of pmp_is_valid_batch_noalloc, a static device
// array of kNoAllocScratchBytes per GPU.  noalloc_scratch_init looks up this
// device's copy (call with `device` current); noalloc_scratch returns it
// (null before the lookup).
constexpr size_t kNoAllocScratchBytes = (size_t)32 << 20;
hipError_t noalloc_scratch_init(int device);
void* noalloc_scratch(int device);
// cloud.hip
// unit_flags (optional): valid[] is formed from the unit pass's words by
// this pass (and the words cleared), instead of by valid_units_finish_kernel.
hipError_t launch_cloud_valid(const CloudTile* tiles, const float* tileB, const CloudSuper* supers,
                              const CloudRange& cr, const float* cen, const float* cap2, const float* rr, int64_t B,
                              int64_t Q, float band, uint8_t* valid, int32_t* flags, int32_t* unit_flags,
                              hipStream_t st);
hipError_t launch_cloud_min(const CloudTile* tiles, const float* tileB, const CloudSuper* supers,
                            const CloudRange& cr, const float* cen, const float* cap2, int64_t B, int64_t Q,
                            float* out_d2, hipStream_t st);
hipError_t launch_cloud_clearance(float* clearance, const float* d2, const float* radius_out, float rc, int64_t B,
                                  int64_t S, hipStream_t st);
// ev_before / ev_after (optional) are recorded around the FP32 edge kernel.
hipError_t launch_edges(const KinArgs<float>& k32, const KinArgs<double>& k64, const SceneArgs<float>& s32,
                        const SceneArgs<double>& s64, const EdgeArgs& ea, float band, int grid, int block,
                        hipStream_t st, hipEvent_t ev_before = nullptr, hipEvent_t ev_after = nullptr);
// Evaluate rows (and, with offsets, trajectory smoothness) by units.
struct EvalArgs {
  float* vals;
  float* jac;
  const int64_t* offsets;  // device, n_traj + 1; nullptr for plain evaluate
  int n_traj;
  const double* w2;
  float* cost;
  float* grad;
  int whole_chunks;  // 0: default
};
hipError_t launch_eval_units(const KinArgs<float>& ka, const SceneArgs<float>& sc, const double* q, int64_t ldq,
                             int64_t B, const EvalArgs& e, hipStream_t st);
// shares: primitive shares (grid z), 0 = about four waves per SIMD
hipError_t launch_pairs(const KinArgs<float>& ka, const SceneArgs<float>& sc, const double* q, int64_t ldq,
                        int64_t B, float* dist, float* grad, int shares, hipStream_t st);
// Primitive shares (grid z) of the adjoint launch: the schedule's (shares >
// 0) or about three waves per SIMD.
int vjp_shares(const KinArgs<float>& ka, int n_prims, int64_t B, int shares);
// part: workspace of n_items * vjp_shares(...) * n_cols * B floats
hipError_t launch_vjp(const KinArgs<float>& ka, const SceneArgs<float>& sc, const double* q, int64_t ldq, int64_t B,
                      const float* gd, float* part, float* grad_q, int shares, hipStream_t st);
hipError_t launch_prims_eval(const DevPrim<double>* prims, int n, const double* cloud_xyz, const double* pts,
                             int64_t N, int mode, double radius, double* out, hipStream_t st);
hipError_t launch_nearest(const double* pts, int64_t N, const double* qry, int64_t M, double* sqd, int32_t* idx,
                          hipStream_t st);

// Device build of one cloud's tile tables (cloud_build.hip).
constexpr int kSortChunk = 512;  // keys sorted per workgroup in LDS
struct CloudBuildSizes {
  int64_t n_tiles, n_supers;
  int64_t n_keys;         // sort length: a power of two >= max(n, kSortChunk)
  int64_t scratch_bytes;  // keys and bounds
};
CloudBuildSizes cloud_build_sizes(int64_t n);
// pts: device [n][3] FP64.  Writes n_tiles records to t32 / t64, n_tiles *
// 4 * kCloudTile floats to tb, n_tiles * kCloudTile * 4 doubles to p64 and
// n_supers records to sup, whose tile ranges start at tile_base.
hipError_t launch_cloud_build(const double* pts, int64_t n, int32_t tile_base, CloudTile* t32, CloudTile64* t64,
                              float* tb, double* p64, CloudSuper* sup, void* scratch, hipStream_t st);

}  // namespace pmp
