// C ABI of plainmp_amd (declared in include/plainmp_amd.h): context,
// table packing and launches.  Host code compiled by hipcc.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "ik.h"
#include "launch.h"

using namespace pmp;

namespace {

thread_local std::string g_last_error;

struct Error : std::runtime_error {
  pmp_status code;
  Error(pmp_status c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define PMP_HIP(expr)                                                                              \
  do {                                                                                             \
    hipError_t e_ = (expr);                                                                        \
    if (e_ != hipSuccess) throw Error(PMP_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)
#define PMP_REQUIRE(cond, msg)                          \
  do {                                                  \
    if (!(cond)) throw Error(PMP_ERR_INVALID, (msg));   \
  } while (0)

// The context whose work the current entry point enqueues (set by
// set_device); its call is recorded for cross-stream ordering when the entry
// point returns.
thread_local pmp_context* g_call_ctx = nullptr;
void record_call(pmp_context* ctx) noexcept;

template <class F>
pmp_status guarded(F&& f) {
  struct Finish {
    ~Finish() {
      if (g_call_ctx) record_call(g_call_ctx);
      g_call_ctx = nullptr;
    }
  } finish;
  try {
    f();
    return PMP_OK;
  } catch (const Error& e) {
    g_last_error = e.what();
    return e.code;
  } catch (const std::exception& e) {
    g_last_error = e.what();
    return PMP_ERR_RUNTIME;
  }
}

// Device buffer that grows on demand and is freed with the context.
struct DevBuf {
  void* ptr = nullptr;
  size_t bytes = 0;
  void* get(size_t n) {
    if (n > bytes) {
      if (ptr) PMP_HIP(hipFree(ptr));
      ptr = nullptr;
      PMP_HIP(hipMalloc(&ptr, n));
      bytes = n;
    }
    return ptr;
  }
  // Synchronous: the host vectors are usually temporaries, and an
  // asynchronous copy from pageable memory may read them after they are freed.
  template <class T>
  T* upload(const std::vector<T>& v, hipStream_t st) {
    T* p = (T*)get(std::max<size_t>(v.size() * sizeof(T), 16));
    if (!v.empty()) {
      PMP_HIP(hipMemcpyAsync(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, st));
      PMP_HIP(hipStreamSynchronize(st));
    }
    return p;
  }
  void release() {
    if (ptr) (void)hipFree(ptr);
    ptr = nullptr;
    bytes = 0;
  }
};

// Device buffers of one edge launch setup: the direct calls of a context
// share one set; every captured graph owns its own, so that later calls
// cannot change what a replay reads.
struct EdgeBufs {
  DevBuf width, inv_width, pend, counter;
  std::vector<double> h_width;  // what width / inv_width hold, so a repeated call uploads nothing
  void release() {
    for (DevBuf* b : {&width, &inv_width, &pend, &counter}) b->release();
    h_width.clear();
  }
};

// Farthest-first visiting order of the motion validators
// (ompl/generate_sequence.py:11-63): element i+1 maximises the minimum
// distance to the elements already chosen, ties to the larger value.
std::vector<uint8_t> sequence_table() {
  std::vector<uint8_t> t(128 * 128, 0);
  for (int n = 1; n <= 128; ++n) {
    std::vector<int> seq{0};
    std::vector<bool> used(n, false);
    used[0] = true;
    for (int i = 1; i < n; ++i) {
      int best = -1, best_d = -1;
      for (int v = 0; v < n; ++v) {
        if (used[v]) continue;
        int d = n;
        for (int s : seq) d = std::min(d, std::abs(v - s));
        if (d > best_d || (d == best_d && v > best)) {
          best_d = d;
          best = v;
        }
      }
      seq.push_back(best);
      used[best] = true;
    }
    for (int i = 0; i < n; ++i) t[(n - 1) * 128 + i] = (uint8_t)seq[i];
  }
  return t;
}

}  // namespace

namespace {
// Where a node's frame can be, for the static part of the cull table.
// World rotation: |q|^2 in [n2_lo, n2_hi], so |R v| <= grow |v|.  Origin:
// within rad of anchor (rad infinite when it depends on a prismatic joint
// or a moving base).
struct NodeReach {
  double n2_lo = 1.0, n2_hi = 1.0, grow = 1.0;
  bool rot_static = false;
  double q[4] = {0, 0, 0, 1}, t[3] = {0, 0, 0};  // the exact pose when rot_static
  double anchor[3] = {0, 0, 0};
  double rad = INFINITY;
};
}  // namespace

struct pmp_context {
  int device = 0;
  hipStream_t stream = nullptr;
  hipStream_t capture_stream = nullptr;  // used for graph capture when no stream is set
  hipEvent_t edge_ev0 = nullptr, edge_ev1 = nullptr;  // around the last direct edge kernel
  float band = 1e-5f;
  pmp_schedule sched{0, 0, 0};
  int n_cu = 256;
  int max_lds = 65536;  // LDS bytes per workgroup on this device
  // robot
  bool has_robot = false;
  DevBuf nodes32, nodes64, sph32, sph64, meta, pairs, blocks, items, units, unit_all, ws_vjp;
  KinArgs<float> k32{};
  KinArgs<double> k64{};
  int n_spheres = 0, n_groups = 0, n_cols = 0, n_nodes = 0;
  DevBuf rad_out;  // sphere radius by group-order index (FP32)
  // host copies for the cull table (rebuilt whenever robot, scene or band change)
  std::vector<DevSphereBlock> h_blocks;
  std::vector<int> h_block_node;
  std::vector<NodeReach> h_reach;
  // scene
  DevBuf prims32, prims64, vprims32, vprims64, supers, tiles32, tiles64, tileB, pts64;
  DevBuf cloud_stage, cloud_scratch;  // points of one cloud and the sort keys of its tile build
  SceneArgs<float> s32{};
  SceneArgs<double> s64{};
  int n_clouds = 0;
  std::vector<CloudRange> clouds;
  std::vector<DevPrim<double>> h_vprims;  // kind-sorted, FP64 bounds
  int h_vrange[kNumKinds + 1] = {0};
  DevBuf cull, cull_range, ecull, ecull_range, vorig;
  std::vector<int32_t> h_vorig;
  // misc
  DevBuf traj_off, traj_w2;
  DevBuf sqp_rows, sqp_tables;  // SQP assembly: evaluate rows of every waypoint; offsets / box widths
  std::vector<int64_t> h_sqp_tables;  // what sqp_tables holds
  std::vector<int64_t> h_traj_off;
  std::vector<double> h_traj_w2;
  EdgeBufs eb;  // edge buffers of direct calls
  DevBuf seq, tmp_prims, tmp_cloud, ws_cen, ws_cloud, ws_flags;
  DevBuf ws_vunit;  // per-configuration flags of the unit validity pass (kValidUnitsMaxBatch)
  // Cross-stream order of calls (they share the workspaces above): an event
  // recorded on the stream of the last call, waited on by a call on another.
  hipEvent_t call_ev = nullptr;
  hipStream_t call_stream = nullptr;
  bool call_recorded = false;
  // Bumped whenever the device tables a captured graph reads are replaced.
  uint64_t tables_gen = 0;
};

struct pmp_graph {
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  hipStream_t stream = nullptr;
  EdgeBufs eb;
  DevBuf traj_off, traj_w2;  // a trajectory graph's offsets and squared weights
  DevBuf ik_pose;            // an IK graph's FP32 link poses
  const pmp_context* ctx = nullptr;
  uint64_t tables_gen = 0;  // the context's tables_gen at capture
};

namespace {

bool capturing(hipStream_t s) {
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  return hipStreamIsCapturing(s, &cs) != hipSuccess || cs != hipStreamCaptureStatusNone;
}

// Every entry point that enqueues work calls this first.  Calls on one
// context share its workspaces, so a call on another stream than the last
// call's waits for that call's work first (nothing is ordered while the
// stream is being captured).
void set_device(pmp_context* ctx) {
  PMP_HIP(hipSetDevice(ctx->device));
  g_call_ctx = ctx;
  if (ctx->call_recorded && ctx->call_stream != ctx->stream && !capturing(ctx->stream))
    PMP_HIP(hipStreamWaitEvent(ctx->stream, ctx->call_ev, 0));
}
hipStream_t st(pmp_context* ctx) { return ctx->stream; }

// The stream a graph is captured on and replayed on: the context stream, or
// with none set (the null stream) a stream of the context's own created
// without hipStreamNonBlocking, which the null stream synchronizes with, so
// that replays are ordered with the caller's null-stream work as direct calls
// are (the null stream itself cannot be captured).
hipStream_t graph_stream(pmp_context* ctx) {
  if (ctx->stream) return ctx->stream;
  if (!ctx->capture_stream) PMP_HIP(hipStreamCreate(&ctx->capture_stream));
  return ctx->capture_stream;
}

// FP64 quaternion product (Eigen order) and q * v (Eigen's formula, which for
// a quaternion s q' with unit q' gives (1 - s^2) v + s^2 R(q') v).
void hq_mul(const double* a, const double* b, double* o) {
  o[0] = a[3] * b[0] + a[0] * b[3] + a[1] * b[2] - a[2] * b[1];
  o[1] = a[3] * b[1] - a[0] * b[2] + a[1] * b[3] + a[2] * b[0];
  o[2] = a[3] * b[2] + a[0] * b[1] - a[1] * b[0] + a[2] * b[3];
  o[3] = a[3] * b[3] - a[0] * b[0] - a[1] * b[1] - a[2] * b[2];
}
void hq_rot(const double* q, const double* v, double* o) {
  const double t[3] = {2 * (q[1] * v[2] - q[2] * v[1]), 2 * (q[2] * v[0] - q[0] * v[2]),
                       2 * (q[0] * v[1] - q[1] * v[0])};
  o[0] = v[0] + q[3] * t[0] + (q[1] * t[2] - q[2] * t[1]);
  o[1] = v[1] + q[3] * t[1] + (q[2] * t[0] - q[0] * t[2]);
  o[2] = v[2] + q[3] * t[2] + (q[0] * t[1] - q[1] * t[0]);
}

// Bounds of every node frame for the cull table, walking the tree from the
// base: the quaternion norm^2 multiplies along the chain (a unit-length
// joint axis keeps the joint quaternion's norm in [min(1,|a|^2), max(1,|a|^2)]
// up to the sin/cos error), and a node origin lies within the parent's
// reach plus grow * |translation| of it.
std::vector<NodeReach> node_reach(const pmp_robot_desc* r) {
  constexpr double kSinCos = 1e-4;  // |s^2 + c^2 - 1| of the FP32 half-angle sin/cos (Taylor or exact)
  constexpr double kRound = 1e-6;   // relative FP32 rounding of |q|^2 per composition
  constexpr double kPad = 1e-5;     // margin between the bounds and what a lane may compute
  const int N = r->n_nodes;
  std::vector<NodeReach> R(N);
  auto finish = [&](NodeReach& c) {
    c.n2_lo *= 1.0 - kPad;
    c.n2_hi *= 1.0 + kPad;
    c.grow = 1.0 + 2.0 * std::max(0.0, std::max(c.n2_hi - 1.0, 1.0 - c.n2_lo));
  };
  NodeReach& b = R[0];
  if (r->base_type == PMP_BASE_FIXED) {
    double n2 = 0.0;
    for (int k = 0; k < 4; ++k) {
      b.q[k] = r->base_pose[3 + k];
      n2 += b.q[k] * b.q[k];
    }
    for (int k = 0; k < 3; ++k) b.t[k] = b.anchor[k] = r->base_pose[k];
    b.n2_lo = n2 * (1.0 - kRound);
    b.n2_hi = n2 * (1.0 + kRound);
    b.rot_static = true;
    b.rad = 0.0;
  } else {
    b.n2_lo = 1.0 - 10 * kRound;  // FP32 quaternion of the base roll / pitch / yaw
    b.n2_hi = 1.0 + 10 * kRound;
  }
  finish(b);
  for (int i = 1; i < N; ++i) {
    const NodeReach& P = R[r->node_parent[i]];
    NodeReach& C = R[i];
    const int kind = r->node_kind[i];
    const double* oq = r->node_quat + 4 * i;
    const double* lt = r->node_trans + 3 * i;
    const double* ax = r->node_axis + 3 * i;
    const bool ident = r->node_identity[i] != 0;
    const double oq2 = ident ? 1.0 : oq[0] * oq[0] + oq[1] * oq[1] + oq[2] * oq[2] + oq[3] * oq[3];
    double lo = P.n2_lo, hi = P.n2_hi;
    if (kind == PMP_NODE_FIXED) {
      lo *= oq2;
      hi *= oq2;
    } else if (kind == PMP_NODE_REVOLUTE) {
      const double a2 = ax[0] * ax[0] + ax[1] * ax[1] + ax[2] * ax[2];
      lo *= oq2 * std::min(1.0, a2) * (1.0 - kSinCos);
      hi *= oq2 * std::max(1.0, a2) * (1.0 + kSinCos);
    }
    C.n2_lo = lo * (1.0 - kRound);
    C.n2_hi = hi * (1.0 + kRound);
    const double ltn = std::sqrt(lt[0] * lt[0] + lt[1] * lt[1] + lt[2] * lt[2]);
    if (kind == PMP_NODE_ROOT) {
      C = R[0];
      continue;
    } else if (kind == PMP_NODE_PRISMATIC || !std::isfinite(P.rad)) {
      C.rad = INFINITY;
    } else if (P.rot_static) {
      double tw[3];
      hq_rot(P.q, lt, tw);
      for (int k = 0; k < 3; ++k) C.t[k] = C.anchor[k] = P.t[k] + tw[k];
      C.rad = 0.0;
      if (kind == PMP_NODE_FIXED) {
        C.rot_static = true;
        if (ident)
          for (int k = 0; k < 4; ++k) C.q[k] = P.q[k];
        else
          hq_mul(P.q, oq, C.q);
      } else if (kind == PMP_NODE_TRANSLATION) {
        C.rot_static = true;
        for (int k = 0; k < 4; ++k) C.q[k] = P.q[k];
      }
    } else {
      for (int k = 0; k < 3; ++k) C.anchor[k] = P.anchor[k];
      C.rad = P.rad + P.grow * ltn;
    }
    finish(C);
  }
  return R;
}

float round_down(double x) {
  float f = (float)x;
  if ((double)f > x) f = std::nextafter(f, -INFINITY);
  return f;
}
float round_up(double x) {
  float f = (float)x;
  if ((double)f < x) f = std::nextafter(f, INFINITY);
  return f;
}

// Candidates of every sphere block: each non-cloud primitive whose AABB,
// grown by the block's reach, the block's bounding centre can enter; with a
// fixed base a primitive is dropped when the grown box lies outside the ball
// that the centre can ever reach.
void build_cull(pmp_context* ctx, double margin, std::vector<CullEntry>& ent, std::vector<CullRange>& rng) {
  const int nb = (int)ctx->h_blocks.size();
  ent.clear();
  rng.assign(std::max(nb, 1), CullRange{});
  for (int b = 0; b < nb; ++b) {
    const NodeReach& nr = ctx->h_reach[ctx->h_block_node[b]];
    const DevSphereBlock& bk = ctx->h_blocks[b];
    const double reach = (double)bk.G * nr.grow + margin + 1e-4;
    const double gn = std::sqrt((double)bk.g[0] * bk.g[0] + (double)bk.g[1] * bk.g[1] + (double)bk.g[2] * bk.g[2]);
    const double ball = nr.rad + nr.grow * gn + 1e-4;
    for (int k = 0; k < kNumKinds; ++k) {
      rng[b].at[k] = (int)ent.size();
      if (k == kKindCloud) continue;
      for (int p = ctx->h_vrange[k]; p < ctx->h_vrange[k + 1]; ++p) {
        const DevPrim<double>& P = ctx->h_vprims[p];
        double lo[3], hi[3], d2 = 0.0;
        for (int d = 0; d < 3; ++d) {
          lo[d] = P.lb[d] - reach;
          hi[d] = P.ub[d] + reach;
          const double e = std::max(std::max(lo[d] - nr.anchor[d], nr.anchor[d] - hi[d]), 0.0);
          d2 += e * e;
        }
        if (std::isfinite(ball) && std::sqrt(d2) > ball) continue;  // NaN keeps the candidate
        CullEntry ce{};
        for (int d = 0; d < 3; ++d) {
          ce.lb[d] = round_down(lo[d]);
          ce.ub[d] = round_up(hi[d]);
        }
        ce.prim = p;
        ce.orig = ctx->h_vorig[p];
        ent.push_back(ce);
      }
    }
    rng[b].at[kNumKinds] = (int)ent.size();
  }
}

// The validity table (boxes grown by the band) and the evaluate table (by
// the 0.1 m cutoff of evaluate_ext_collision's prefilter).
void rebuild_cull(pmp_context* ctx) {
  ++ctx->tables_gen;
  std::vector<CullEntry> ent;
  std::vector<CullRange> rng;
  hipStream_t s0 = st(ctx);
  build_cull(ctx, (double)ctx->band, ent, rng);
  const CullEntry* c = ctx->cull.upload(ent, s0);
  const CullRange* cr = ctx->cull_range.upload(rng, s0);
  ctx->s32.cull = ctx->s64.cull = c;
  ctx->s32.cull_range = ctx->s64.cull_range = cr;
  build_cull(ctx, 0.1, ent, rng);
  c = ctx->ecull.upload(ent, s0);
  cr = ctx->ecull_range.upload(rng, s0);
  ctx->s32.ecull = ctx->s64.ecull = c;
  ctx->s32.ecull_range = ctx->s64.ecull_range = cr;
}

template <class T>
DevPrim<T> make_prim(const pmp_prim_desc& d, int cloud_begin) {
  DevPrim<T> p{};
  p.type = d.type;
  p.flags = (d.axis_aligned ? 1 : 0) | (d.z_axis_aligned ? 2 : 0);
  p.cloud_begin = cloud_begin;
  p.cloud_count = d.type == PMP_PRIM_CLOUD ? (int32_t)d.n_points : 0;
  for (int k = 0; k < 3; ++k) {
    p.pos[k] = (T)d.position[k];
    p.lb[k] = (T)d.lb[k];
    p.ub[k] = (T)d.ub[k];
  }
  for (int k = 0; k < 9; ++k) {
    p.R[k] = (T)d.rotation[k];
    p.Rinv[k] = (T)d.rotation_inv[k];
  }
  switch (d.type) {
    case PMP_PRIM_BOX:  // half_width_ = 0.5 * width (primitive_sdf.hpp:283)
      p.p0 = (T)(0.5 * d.size[0]);
      p.p1 = (T)(0.5 * d.size[1]);
      p.p2 = (T)(0.5 * d.size[2]);
      break;
    case PMP_PRIM_CYLINDER:  // r, half height, r^2 (:431-437)
      p.p0 = (T)d.size[0];
      p.p1 = (T)(0.5 * d.size[1]);
      p.p2 = (T)(d.size[0] * d.size[0]);
      break;
    case PMP_PRIM_SPHERE:  // r, r^2 (:533-537)
      p.p0 = (T)d.size[0];
      p.p1 = (T)(d.size[0] * d.size[0]);
      break;
    case PMP_PRIM_GROUND:
    case PMP_PRIM_CLOUD:
      p.p0 = (T)d.size[0];
      break;
  }
  return p;
}

void check_prim(const pmp_prim_desc& d) {
  PMP_REQUIRE(d.type >= PMP_PRIM_BOX && d.type <= PMP_PRIM_CLOUD, "unknown primitive type");
  if (d.type == PMP_PRIM_CLOUD) PMP_REQUIRE(d.points != nullptr && d.n_points > 0, "cloud primitive without points");
}

// Pack a primitive list into FP32 / FP64 device records and cloud tables.
void pack_prims(const pmp_scene_desc& sc, std::vector<DevPrim<float>>& p32, std::vector<DevPrim<double>>& p64,
                std::vector<float>& c32, std::vector<double>& c64, int& n_clouds) {
  n_clouds = 0;
  int64_t cloud_at = 0;
  for (int i = 0; i < sc.n_prims; ++i) {
    const pmp_prim_desc& d = sc.prims[i];
    check_prim(d);
    p32.push_back(make_prim<float>(d, (int)cloud_at));
    p64.push_back(make_prim<double>(d, (int)cloud_at));
    if (d.type == PMP_PRIM_CLOUD) {
      ++n_clouds;
      PMP_REQUIRE(cloud_at + d.n_points < (int64_t)1 << 30, "cloud too large");
      for (int64_t j = 0; j < d.n_points; ++j) {
        const double* y = d.points + 3 * j;
        double n2 = y[0] * y[0] + y[1] * y[1] + y[2] * y[2];
        for (int k = 0; k < 3; ++k) {
          c32.push_back((float)y[k]);
          c64.push_back(y[k]);
        }
        c32.push_back((float)n2);
        c64.push_back(n2);
      }
      cloud_at += d.n_points;
    }
  }
}

void require_robot(pmp_context* ctx) {
  PMP_REQUIRE(ctx != nullptr, "null context");
  PMP_REQUIRE(ctx->has_robot, "no robot tables uploaded (call pmp_upload_robot first)");
}

}  // namespace

namespace {
// End of an entry point that enqueued work: remember it for set_device.
void record_call(pmp_context* ctx) noexcept {
  if (capturing(ctx->stream)) return;
  if (!ctx->call_ev) return;  // created with the context
  if (hipEventRecord(ctx->call_ev, ctx->stream) != hipSuccess) return;
  ctx->call_recorded = true;
  ctx->call_stream = ctx->stream;
}
}  // namespace

extern "C" {

const char* pmp_last_error(void) { return g_last_error.c_str(); }
int32_t pmp_abi_version(void) { return 2; }

pmp_status pmp_create(int32_t device, pmp_context** out) {
  return guarded([&] {
    PMP_REQUIRE(out != nullptr, "null output pointer");
    int n = 0;
    PMP_HIP(hipGetDeviceCount(&n));
    PMP_REQUIRE(device >= 0 && device < n, "device index out of range");
    auto* ctx = new pmp_context();
    ctx->device = device;
    try {
      PMP_HIP(hipSetDevice(device));
      hipDeviceProp_t prop;
      PMP_HIP(hipGetDeviceProperties(&prop, device));
      ctx->n_cu = prop.multiProcessorCount;
      ctx->max_lds = (int)prop.sharedMemPerBlock;
      PMP_HIP(noalloc_scratch_init(device));
      PMP_HIP(hipEventCreateWithFlags(&ctx->call_ev, hipEventDisableTiming));
      ctx->seq.upload(sequence_table(), nullptr);
      PMP_HIP(hipDeviceSynchronize());
    } catch (...) {
      delete ctx;
      throw;
    }
    *out = ctx;
  });
}

pmp_status pmp_destroy(pmp_context* ctx) {
  return guarded([&] {
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    (void)hipDeviceSynchronize();
    for (DevBuf* b : {&ctx->nodes32, &ctx->nodes64, &ctx->sph32, &ctx->sph64, &ctx->meta, &ctx->pairs, &ctx->blocks, &ctx->items, &ctx->units, &ctx->unit_all, &ctx->ws_vjp,
                      &ctx->vprims32, &ctx->vprims64, &ctx->supers, &ctx->cull, &ctx->ecull, &ctx->ecull_range, &ctx->vorig, &ctx->cull_range, &ctx->rad_out,
                      &ctx->prims32, &ctx->prims64, &ctx->tiles32, &ctx->tiles64, &ctx->tileB, &ctx->pts64, &ctx->seq,
                      &ctx->tmp_prims, &ctx->tmp_cloud, &ctx->ws_cen, &ctx->cloud_stage, &ctx->cloud_scratch, &ctx->ws_vunit,
                      &ctx->ws_cloud, &ctx->ws_flags, &ctx->traj_off, &ctx->traj_w2, &ctx->sqp_rows, &ctx->sqp_tables})
      b->release();
    ctx->eb.release();
    if (ctx->capture_stream) (void)hipStreamDestroy(ctx->capture_stream);
    if (ctx->edge_ev0) (void)hipEventDestroy(ctx->edge_ev0);
    if (ctx->edge_ev1) (void)hipEventDestroy(ctx->edge_ev1);
    if (ctx->call_ev) (void)hipEventDestroy(ctx->call_ev);
    delete ctx;
  });
}

pmp_status pmp_set_stream(pmp_context* ctx, void* stream) {
  return guarded([&] {
    PMP_REQUIRE(ctx != nullptr, "null context");
    ctx->stream = (hipStream_t)stream;
  });
}

pmp_status pmp_synchronize(pmp_context* ctx) {
  return guarded([&] {
    PMP_REQUIRE(ctx != nullptr, "null context");
    set_device(ctx);
    PMP_HIP(hipStreamSynchronize(ctx->stream));
  });
}

pmp_status pmp_set_band(pmp_context* ctx, double band) {
  return guarded([&] {
    PMP_REQUIRE(ctx != nullptr, "null context");
    PMP_REQUIRE(band >= 0.0 && band < 1.0, "band must be in [0, 1) m");
    ctx->band = (float)band;
    if (ctx->has_robot) {
      set_device(ctx);
      rebuild_cull(ctx);
    }
  });
}

pmp_status pmp_set_schedule(pmp_context* ctx, const pmp_schedule* s) {
  return guarded([&] {
    PMP_REQUIRE(ctx != nullptr && s != nullptr, "null argument");
    PMP_REQUIRE(s->block == 0 || (s->block % 64 == 0 && s->block <= 256), "block must be a multiple of 64, <= 256");
    PMP_REQUIRE(s->waves_per_cu >= 0 && s->waves_per_cu <= 32, "waves_per_cu must be in [0, 32]");
    PMP_REQUIRE(s->eval_whole_chunks >= 0, "eval_whole_chunks must be >= 0");
    PMP_REQUIRE(s->pair_shares >= 0 && s->pair_shares <= 256, "pair_shares must be in [0, 256]");
    ctx->sched = *s;
  });
}

pmp_status pmp_upload_robot(pmp_context* ctx, const pmp_robot_desc* r) {
  return guarded([&] {
    PMP_REQUIRE(ctx != nullptr && r != nullptr, "null argument");
    ++ctx->tables_gen;  // captured graphs of the old tables must not replay
    const int N = r->n_nodes;
    PMP_REQUIRE(N >= 1 && N <= kMaxNodes, "n_nodes out of range");
    PMP_REQUIRE(r->node_parent && r->node_kind && r->node_q_col && r->node_quat && r->node_trans && r->node_axis &&
                    r->node_identity,
                "missing node table");
    PMP_REQUIRE(r->node_parent[0] == -1 && r->node_kind[0] == PMP_NODE_ROOT, "node 0 must be the root");
    PMP_REQUIRE(r->base_type >= PMP_BASE_FIXED && r->base_type <= PMP_BASE_FLOATING, "unknown base type");
    const int n_q = r->n_q_cols;
    const int n_cols = n_q + (r->base_type == PMP_BASE_PLANAR ? 3 : 0) + (r->base_type == PMP_BASE_FLOATING ? 6 : 0);
    PMP_REQUIRE(n_q >= 0 && n_cols <= kMaxCols, "too many q columns");
    const int S = r->n_spheres, G = r->n_groups;
    PMP_REQUIRE(S >= 0 && S <= kMaxSpheres, "n_spheres out of range");
    PMP_REQUIRE(G >= 0 && (S == 0 || G >= 1), "n_groups out of range");
    PMP_REQUIRE(r->n_self_pairs >= 0 && r->n_self_pairs <= kMaxSelfPairs, "n_self_pairs out of range");
    std::vector<std::vector<int>> children(N);
    for (int i = 1; i < N; ++i) {
      const int p = r->node_parent[i];
      PMP_REQUIRE(p >= 0 && p < i, "node parents must precede their children");
      children[p].push_back(i);
      const int k = r->node_kind[i];
      PMP_REQUIRE(k >= PMP_NODE_FIXED && k <= PMP_NODE_TRANSLATION, "unknown node kind");
      if (k == PMP_NODE_REVOLUTE || k == PMP_NODE_PRISMATIC)
        PMP_REQUIRE(r->node_q_col[i] >= 0 && r->node_q_col[i] < n_q, "joint node without a valid q column");
    }
    for (int s = 0; s < S; ++s) {
      PMP_REQUIRE(r->sphere_node[s] >= 0 && r->sphere_node[s] < N, "sphere attached to an unknown node");
      PMP_REQUIRE(r->sphere_group[s] >= 0 && r->sphere_group[s] < G, "sphere group out of range");
      PMP_REQUIRE(s == 0 || r->sphere_group[s] >= r->sphere_group[s - 1], "spheres must be in group order");
    }
    // nodes carrying a self-pair group keep their pose to the end
    std::vector<int> group_node(G, -1);
    for (int s = 0; s < S; ++s) {
      const int g = r->sphere_group[s];
      PMP_REQUIRE(group_node[g] < 0 || group_node[g] == r->sphere_node[s], "a group spans several nodes");
      group_node[g] = r->sphere_node[s];
    }
    std::vector<bool> keep_to_end(N, false);
    for (int k = 0; k < r->n_self_pairs; ++k)
      for (int e = 0; e < 2; ++e) {
        const int g = r->self_pairs[2 * k + e];
        PMP_REQUIRE(g >= 0 && g < G && group_node[g] >= 0, "self pair names an empty or unknown group");
        keep_to_end[group_node[g]] = true;
      }
    // register slots: a node keeps its pose if a later child is not the next
    // node or if the self pass needs it
    std::vector<int> slot(N, -1), last_child(N, -1);
    for (int i = 0; i < N; ++i)
      for (int c : children[i]) last_child[i] = std::max(last_child[i], c);
    std::vector<bool> used(kMaxSlots, false);
    std::vector<int> parent_slot(N, -1);
    for (int i = 0; i < N; ++i) {
      const int p = i == 0 ? -1 : r->node_parent[i];
      if (p <= 0)
        parent_slot[i] = -1;  // the root's pose is the base pose
      else if (p == i - 1)
        parent_slot[i] = -2;
      else {
        PMP_REQUIRE(slot[p] >= 0, "internal: parent pose not kept");
        parent_slot[i] = slot[p];
      }
      // release parents whose last child is this node
      if (p > 0 && last_child[p] == i && slot[p] >= 0 && !keep_to_end[p]) used[slot[p]] = false;
      bool need = keep_to_end[i];
      for (int c : children[i])
        if (c != i + 1) need = true;
      if (i == 0) need = false;
      if (need) {
        int f = -1;
        for (int k = 0; k < kMaxSlots; ++k)
          if (!used[k]) {
            f = k;
            break;
          }
        PMP_REQUIRE(f >= 0, "kinematic tree too branched: more than 32 live poses");
        used[f] = true;
        slot[i] = f;
      }
    }
    // q columns that move each node (reference rptable_), base columns move all
    std::vector<uint64_t> mask(N, 0);
    uint64_t base_bits = 0;
    for (int c = n_q; c < n_cols; ++c) base_bits |= 1ull << c;
    uint64_t rev = 0;
    if (r->base_type == PMP_BASE_PLANAR) rev |= 1ull << (n_q + 2);
    if (r->base_type == PMP_BASE_FLOATING) rev |= (1ull << (n_q + 3)) | (1ull << (n_q + 4)) | (1ull << (n_q + 5));
    for (int i = 0; i < N; ++i) {
      uint64_t m = base_bits;
      for (int a = i; a > 0; a = r->node_parent[a]) {
        const int k = r->node_kind[a];
        if (k == PMP_NODE_REVOLUTE || k == PMP_NODE_PRISMATIC) m |= 1ull << r->node_q_col[a];
      }
      mask[i] = m;
      if (r->node_kind[i] == PMP_NODE_REVOLUTE) rev |= 1ull << r->node_q_col[i];
    }
    // spheres sorted by node (stable: groups stay contiguous and in order)
    std::vector<int> order(S);
    for (int s = 0; s < S; ++s) order[s] = s;
    std::stable_sort(order.begin(), order.end(),
                     [&](int a, int b) { return r->sphere_node[a] < r->sphere_node[b]; });
    std::vector<int> pos_of(S);
    for (int k = 0; k < S; ++k) pos_of[order[k]] = k;
    std::vector<DevSphere<float>> sp32(S);
    std::vector<DevSphere<double>> sp64(S);
    std::vector<DevSphereMeta> meta(S);
    int n_ext = 0;
    for (int k = 0; k < S; ++k) {
      const int s = order[k];
      const double* l = r->sphere_local + 3 * s;
      sp32[k] = {(float)l[0], (float)l[1], (float)l[2], (float)r->sphere_radius[s]};
      sp64[k] = {l[0], l[1], l[2], r->sphere_radius[s]};
      const int g = r->sphere_group[s];
      const bool ext = !(r->group_only_self && r->group_only_self[g]);
      meta[k] = {s, ext ? n_ext : -1, r->sphere_node[s], g};
      n_ext += ext;
    }
    // Sphere blocks: per node, its primitive-checked spheres sorted along
    // their longest extent and cut into runs of at most kBlockSpheres, each
    // with a bounding sphere (rounded up to FP32) in the link frame.
    std::vector<DevSphereBlock> blocks;
    std::vector<int> blk_begin(N, 0), blk_end(N, 0), block_node;
    const std::vector<NodeReach> reach = node_reach(r);
    for (int i = 0; i < N; ++i) {
      blk_begin[i] = (int)blocks.size();
      std::vector<int> mem;
      for (int k = 0; k < S; ++k)
        if (meta[k].node == i && meta[k].ext >= 0) mem.push_back(k);
      if (!mem.empty()) {
        auto coord = [&](int k, int d) { return d == 0 ? sp64[k].x : (d == 1 ? sp64[k].y : sp64[k].z); };
        double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
        for (int k : mem)
          for (int d = 0; d < 3; ++d) {
            lo[d] = std::min(lo[d], coord(k, d));
            hi[d] = std::max(hi[d], coord(k, d));
          }
        int axis = 0;
        for (int d = 1; d < 3; ++d)
          if (hi[d] - lo[d] > hi[axis] - lo[axis]) axis = d;
        std::stable_sort(mem.begin(), mem.end(), [&](int a, int b) { return coord(a, axis) < coord(b, axis); });
        const int m = (int)mem.size(), nblk = (m + kBlockSpheres - 1) / kBlockSpheres;
        for (int q = 0; q < nblk; ++q) {
          const int a = m * q / nblk, z = m * (q + 1) / nblk;
          DevSphereBlock bk{};
          bk.n = z - a;
          double blo[3] = {INFINITY, INFINITY, INFINITY}, bhi[3] = {-INFINITY, -INFINITY, -INFINITY};
          for (int t = a; t < z; ++t)
            for (int d = 0; d < 3; ++d) {
              blo[d] = std::min(blo[d], coord(mem[t], d) - sp64[mem[t]].r);
              bhi[d] = std::max(bhi[d], coord(mem[t], d) + sp64[mem[t]].r);
            }
          for (int d = 0; d < 3; ++d) bk.g[d] = (float)(0.5 * (blo[d] + bhi[d]));
          double G = 0.0;
          for (int t = a; t < z; ++t) {
            const int k = mem[t];
            const double ex = coord(k, 0) - bk.g[0], ey = coord(k, 1) - bk.g[1], ez = coord(k, 2) - bk.g[2];
            G = std::max(G, std::sqrt(ex * ex + ey * ey + ez * ez) + sp64[k].r);
            bk.ext[t - a] = meta[k].ext;
            bk.sph[t - a] = k;
            bk.s[t - a][0] = sp32[k].x;
            bk.s[t - a][1] = sp32[k].y;
            bk.s[t - a][2] = sp32[k].z;
            bk.s[t - a][3] = sp32[k].r;
          }
          bk.G = (float)G;
          if ((double)bk.G < G) bk.G = std::nextafter(bk.G, INFINITY);
          if (!(G == G)) bk.G = INFINITY;  // never cull around a NaN bound
          blocks.push_back(bk);
          block_node.push_back(i);
        }
      }
      blk_end[i] = (int)blocks.size();
    }
    std::vector<DevNode<float>> n32(N);
    std::vector<DevNode<double>> n64(N);
    for (int i = 0; i < N; ++i) {
      int b = S, e = S;
      for (int k = 0; k < S; ++k)
        if (meta[k].node == i) {
          b = std::min(b, k);
          e = k + 1;
        }
      if (b == S) b = e = 0;
      auto fill = [&](auto& nd) {
        using T = std::remove_reference_t<decltype(nd.quat[0])>;
        nd.kind = r->node_kind[i];
        nd.parent_slot = parent_slot[i];
        nd.out_slot = slot[i];
        nd.q_col = (nd.kind == PMP_NODE_REVOLUTE || nd.kind == PMP_NODE_PRISMATIC) ? r->node_q_col[i] : -1;
        nd.sphere_begin = b;
        nd.sphere_end = e;
        nd.identity = r->node_identity[i] ? 1 : 0;
        nd.col_mask = mask[i];
        nd.block_begin = blk_begin[i];
        nd.block_end = blk_end[i];
        nd.n2_lo = (T)reach[i].n2_lo;
        nd.n2_hi = (T)reach[i].n2_hi;
        for (int k = 0; k < 4; ++k) nd.quat[k] = (T)r->node_quat[4 * i + k];
        for (int k = 0; k < 3; ++k) {
          nd.trans[k] = (T)r->node_trans[3 * i + k];
          nd.axis[k] = (T)r->node_axis[3 * i + k];
        }
      };
      fill(n32[i]);
      fill(n64[i]);
    }
    std::vector<DevSelfPair> pairs;
    for (int k = 0; k < r->n_self_pairs; ++k) {
      int g1 = r->self_pairs[2 * k], g2 = r->self_pairs[2 * k + 1];
      auto range = [&](int g, int& b, int& e) {
        b = S;
        e = 0;
        for (int s = 0; s < S; ++s)
          if (r->sphere_group[s] == g) {
            b = std::min(b, pos_of[s]);
            e = std::max(e, pos_of[s] + 1);
          }
      };
      DevSelfPair pr{};
      range(g1, pr.begin1, pr.end1);
      range(g2, pr.begin2, pr.end2);
      // bounding sphere of a group's spheres in its link frame (the centre
      // of their box, the radius covering every sphere, rounded up)
      auto bound = [&](int b, int e, float* out) {
        auto coord = [&](int k, int d) { return d == 0 ? sp64[k].x : (d == 1 ? sp64[k].y : sp64[k].z); };
        double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
        for (int t = b; t < e; ++t)
          for (int d = 0; d < 3; ++d) {
            lo[d] = std::min(lo[d], coord(t, d) - sp64[t].r);
            hi[d] = std::max(hi[d], coord(t, d) + sp64[t].r);
          }
        for (int d = 0; d < 3; ++d) out[d] = (float)(0.5 * (lo[d] + hi[d]));
        double R = 0.0;
        for (int t = b; t < e; ++t) {
          const double ex = coord(t, 0) - out[0], ey = coord(t, 1) - out[1], ez = coord(t, 2) - out[2];
          R = std::max(R, std::sqrt(ex * ex + ey * ey + ez * ez) + sp64[t].r);
        }
        out[3] = round_up(R);
        if (!(R == R) || b >= e) out[3] = INFINITY;  // never cull around a NaN or empty bound
      };
      bound(pr.begin1, pr.end1, pr.bound1);
      bound(pr.begin2, pr.end2, pr.bound2);
      pr.node1 = group_node[g1];
      pr.node2 = group_node[g2];
      pr.slot1 = slot[pr.node1];
      pr.slot2 = slot[pr.node2];
      PMP_REQUIRE(pr.slot1 >= 0 && pr.slot2 >= 0, "internal: self-pair pose not kept");
      pairs.push_back(pr);
    }
    set_device(ctx);
    hipStream_t s0 = st(ctx);
    KinArgs<float> a{};
    KinArgs<double> d{};
    a.nodes = ctx->nodes32.upload(n32, s0);
    d.nodes = ctx->nodes64.upload(n64, s0);
    a.spheres = ctx->sph32.upload(sp32, s0);
    d.spheres = ctx->sph64.upload(sp64, s0);
    a.meta = d.meta = ctx->meta.upload(meta, s0);
    a.pairs = d.pairs = ctx->pairs.upload(pairs, s0);
    a.blocks = d.blocks = ctx->blocks.upload(blocks, s0);
    std::vector<DevSphereItem> items;
    for (int i = 0; i < N; ++i)
      for (int b0 = n32[i].sphere_begin; b0 < n32[i].sphere_end; b0 += kItemSpheres)
        items.push_back({i, b0, std::min(b0 + kItemSpheres, n32[i].sphere_end), 0});
    a.items = d.items = ctx->items.upload(items, s0);
    a.n_items = d.n_items = (int32_t)items.size();
    // evaluate units: one per sphere item, then one per self pair; and the whole configuration
    std::vector<DevEvalUnit> units;
    for (int b = 0; b < (int)blocks.size(); ++b) units.push_back({b, b + 1, 0, 0, block_node[b], {0, 0, 0}});
    for (int k = 0; k < (int)pairs.size(); ++k)
      units.push_back({0, 0, k, k + 1, std::max(pairs[k].node1, pairs[k].node2), {0, 0, 0}});
    std::vector<DevEvalUnit> whole{{0, (int32_t)blocks.size(), 0, (int32_t)pairs.size(), N - 1, {0, 0, 0}}};
    a.units = d.units = ctx->units.upload(units, s0);
    a.n_units = d.n_units = (int32_t)units.size();
    a.unit_all = d.unit_all = ctx->unit_all.upload(whole, s0);
    std::vector<float> rad(std::max(S, 1), 0.f);
    for (int s2 = 0; s2 < S; ++s2) rad[s2] = (float)r->sphere_radius[s2];
    ctx->rad_out.upload(rad, s0);
    a.n_nodes = d.n_nodes = N;
    a.n_spheres = d.n_spheres = S;
    a.n_pairs = d.n_pairs = (int)pairs.size();
    a.n_cols = d.n_cols = n_cols;
    a.n_q_cols = d.n_q_cols = n_q;
    a.base_type = d.base_type = r->base_type;
    a.n_ext_spheres = d.n_ext_spheres = n_ext;
    a.rev_mask = d.rev_mask = rev;
    int n_slots = 0;
    for (int i = 0; i < N; ++i) n_slots = std::max(n_slots, slot[i] + 1);
    a.n_slots = d.n_slots = n_slots;
    a.slot_off = d.slot_off = 0;
    a.slot_wave_bytes = n_slots * 7 * 64 * (int)sizeof(float);
    d.slot_wave_bytes = n_slots * 7 * 64 * (int)sizeof(double);
    a.lds_limit = d.lds_limit = ctx->max_lds;
    PMP_REQUIRE(slot_bytes(d, 64) <= (size_t)ctx->max_lds,
                "kinematic tree too branched for this device: " + std::to_string(n_slots) +
                    " live poses need more LDS than a workgroup has");
    for (int k = 0; k < 3; ++k) {
      a.base_t[k] = (float)r->base_pose[k];
      d.base_t[k] = r->base_pose[k];
    }
    for (int k = 0; k < 4; ++k) {
      a.base_q[k] = (float)r->base_pose[3 + k];
      d.base_q[k] = r->base_pose[3 + k];
    }
    PMP_HIP(hipStreamSynchronize(s0));
    ctx->k32 = a;
    ctx->k64 = d;
    ctx->n_spheres = S;
    ctx->n_groups = G;
    ctx->n_cols = n_cols;
    ctx->n_nodes = N;
    ctx->h_blocks = blocks;
    ctx->h_block_node = block_node;
    ctx->h_reach = reach;
    ctx->has_robot = true;
    rebuild_cull(ctx);
  });
}

pmp_status pmp_upload_scene(pmp_context* ctx, const pmp_scene_desc* sc) {
  return guarded([&] {
    PMP_REQUIRE(ctx != nullptr && sc != nullptr, "null argument");
    ++ctx->tables_gen;  // captured graphs of the old tables must not replay
    PMP_REQUIRE(sc->n_prims >= 0 && sc->n_prims <= kMaxPrims, "n_prims out of range");
    PMP_REQUIRE(sc->n_prims == 0 || sc->prims != nullptr, "missing primitive list");
    std::vector<DevPrim<float>> p32;
    std::vector<DevPrim<double>> p64;
    std::vector<float> c32;
    std::vector<double> c64;
    int n_clouds = 0;
    pack_prims(*sc, p32, p64, c32, c64, n_clouds);
    PMP_REQUIRE(n_clouds <= kMaxClouds, "at most 4 point clouds per scene");
    // cloud tile tables: sizes here, contents built on the device below
    std::vector<CloudRange> clouds;
    std::vector<int> cloud_prim;
    int64_t n_tiles = 0, n_sup = 0, scratch = 16, max_pts = 0;
    for (int i = 0; i < sc->n_prims; ++i) {
      if (sc->prims[i].type != PMP_PRIM_CLOUD) continue;
      const CloudBuildSizes z = cloud_build_sizes(sc->prims[i].n_points);
      const int begin = (int)n_tiles, oc = (int)clouds.size();
      p32[i].cloud_begin = p64[i].cloud_begin = begin;
      p32[i].cloud_count = p64[i].cloud_count = (int)z.n_tiles;
      p32[i].flags |= oc << 4;
      p64[i].flags |= oc << 4;
      clouds.push_back({begin, (int)z.n_tiles, (float)sc->prims[i].size[0], (int)n_sup, (int)z.n_supers});
      cloud_prim.push_back(i);
      n_tiles += z.n_tiles;
      n_sup += z.n_supers;
      scratch = std::max(scratch, z.scratch_bytes);
      max_pts = std::max(max_pts, sc->prims[i].n_points);
    }
    // kind-sorted copy for the validity pass (its answer does not depend on
    // the primitive order)
    std::vector<DevPrim<float>> v32;
    std::vector<DevPrim<double>> v64;
    std::vector<int32_t> vorig;
    int vrange[kNumKinds + 1] = {0};
    for (int k = 0; k < kNumKinds; ++k) {
      vrange[k] = (int)v32.size();
      for (size_t i = 0; i < p32.size(); ++i)
        if (prim_kind(p32[i]) == k) {
          v32.push_back(p32[i]);
          v64.push_back(p64[i]);
          vorig.push_back((int32_t)i);
        }
    }
    vrange[kNumKinds] = (int)v32.size();
    set_device(ctx);
    hipStream_t s0 = st(ctx);
    SceneArgs<float> a{};
    SceneArgs<double> d{};
    a.prims = ctx->prims32.upload(p32, s0);
    a.vprims = ctx->vprims32.upload(v32, s0);
    d.vprims = ctx->vprims64.upload(v64, s0);
    a.vorig = d.vorig = ctx->vorig.upload(vorig, s0);
    for (int k = 0; k <= kNumKinds; ++k) a.vrange[k] = d.vrange[k] = vrange[k];
    d.prims = ctx->prims64.upload(p64, s0);
    CloudTile* t32 = (CloudTile*)ctx->tiles32.get(std::max<size_t>(n_tiles * sizeof(CloudTile), 16));
    CloudTile64* t64 = (CloudTile64*)ctx->tiles64.get(std::max<size_t>(n_tiles * sizeof(CloudTile64), 16));
    float* tb = (float*)ctx->tileB.get(std::max<size_t>(n_tiles * 4 * kCloudTile * sizeof(float), 16));
    double* p64c = (double*)ctx->pts64.get(std::max<size_t>(n_tiles * kCloudTile * 4 * sizeof(double), 16));
    CloudSuper* sup = (CloudSuper*)ctx->supers.get(std::max<size_t>(n_sup * sizeof(CloudSuper), 16));
    if (!clouds.empty()) {
      double* stage = (double*)ctx->cloud_stage.get(max_pts * 3 * sizeof(double));
      void* scr = ctx->cloud_scratch.get(scratch);
      for (size_t c = 0; c < clouds.size(); ++c) {
        const pmp_prim_desc& pd = sc->prims[cloud_prim[c]];
        const CloudRange& cr = clouds[c];
        PMP_HIP(hipMemcpyAsync(stage, pd.points, pd.n_points * 3 * sizeof(double), hipMemcpyHostToDevice, s0));
        PMP_HIP(launch_cloud_build(stage, pd.n_points, cr.tile_begin, t32 + cr.tile_begin, t64 + cr.tile_begin,
                                   tb + (int64_t)cr.tile_begin * 4 * kCloudTile, p64c + (int64_t)cr.tile_begin * kCloudTile * 4,
                                   sup + cr.super_begin, scr, s0));
        PMP_HIP(hipStreamSynchronize(s0));  // the next cloud reuses the staging buffer
      }
    }
    a.tiles = d.tiles = t32;
    a.supers = d.supers = sup;
    a.tiles64 = d.tiles64 = t64;
    a.tileB = d.tileB = tb;
    a.pts64 = d.pts64 = p64c;
    a.n_prims = d.n_prims = sc->n_prims;
    a.n_cloud = d.n_cloud = n_clouds;
    for (int i = 0; i <= kMaxClouds; ++i)
      a.super_at[i] = d.super_at[i] = i < n_clouds ? clouds[i].super_begin : (int32_t)n_sup;
    ctx->s32 = a;
    ctx->s64 = d;
    ctx->n_clouds = n_clouds;
    ctx->clouds = clouds;
    ctx->h_vprims = v64;
    ctx->h_vorig = vorig;
    for (int k = 0; k <= kNumKinds; ++k) ctx->h_vrange[k] = vrange[k];
    if (ctx->has_robot) rebuild_cull(ctx);
  });
}

pmp_status pmp_query_sizes(pmp_context* ctx, int32_t* n_spheres, int32_t* n_prims, int32_t* q_dim, int32_t* cst_dim) {
  return guarded([&] {
    require_robot(ctx);
    if (n_spheres) *n_spheres = ctx->n_spheres;
    if (n_prims) *n_prims = ctx->s32.n_prims;
    if (q_dim) *q_dim = ctx->n_cols;
    if (cst_dim) *cst_dim = (ctx->s32.n_prims > 0 ? 1 : 0) + (ctx->k32.n_pairs > 0 ? 1 : 0);
  });
}

#define PMP_BATCH_CHECKS(q, ldq, B)                                       \
  require_robot(ctx);                                                     \
  PMP_REQUIRE((B) >= 0, "negative batch size");                           \
  PMP_REQUIRE((B) == 0 || (q) != nullptr, "null configuration pointer");  \
  PMP_REQUIRE((ldq) >= (B), "leading dimension smaller than the batch")

pmp_status pmp_fk_spheres_batch(pmp_context* ctx, const double* q, int64_t ldq, int64_t B, int32_t accurate,
                                float* centers) {
  return guarded([&] {
    PMP_BATCH_CHECKS(q, ldq, B);
    PMP_REQUIRE(B == 0 || centers, "null output");
    if (B == 0 || ctx->n_spheres == 0) return;
    set_device(ctx);
    PMP_HIP(launch_centers(ctx->k32, q, ldq, B, accurate ? 0 : 1, centers, st(ctx)));
  });
}

pmp_status pmp_link_poses_batch(pmp_context* ctx, const double* q, int64_t ldq, int64_t B, int32_t accurate,
                                double* poses) {
  return guarded([&] {
    PMP_BATCH_CHECKS(q, ldq, B);
    PMP_REQUIRE(B == 0 || poses, "null output");
    if (B == 0) return;
    set_device(ctx);
    PMP_HIP(launch_poses(ctx->k64, q, ldq, B, accurate ? 0 : 1, poses, st(ctx)));
  });
}

pmp_status pmp_link_jacobian_batch(pmp_context* ctx, const double* q, int64_t ldq, int64_t B, int32_t link,
                                   const double* point, float* jac) {
  return guarded([&] {
    PMP_BATCH_CHECKS(q, ldq, B);
    PMP_REQUIRE(B == 0 || jac, "null output");
    PMP_REQUIRE(link >= 0 && link < ctx->n_nodes, "link index out of range");
    if (B == 0) return;
    set_device(ctx);
    const float off[3] = {point ? (float)point[0] : 0.f, point ? (float)point[1] : 0.f, point ? (float)point[2] : 0.f};
    PMP_HIP(launch_link_jacobian(ctx->k32, q, ldq, B, link, off, jac, nullptr, st(ctx)));
  });
}

}  // extern "C"

namespace {
// Argument checks of an IK step (R > 0) and its kernel arguments.
IkArgs ik_args(pmp_context* ctx, double* q, int64_t R, const double* poses, const pmp_ik_step_desc* d) {
  PMP_REQUIRE(q && d->target && d->jac0 && d->lo && d->hi, "null argument");
  PMP_REQUIRE(d->kind == 3 || (d->jac_u[0] && d->jac_u[1] && d->jac_u[2]), "rotation targets need jac_u");
  PMP_REQUIRE(d->n_cst == 0 || (d->vals && d->cjac), "null collision rows");
  IkArgs a{};
  a.R = R;
  a.D = ctx->n_cols;
  a.C = d->n_cst;
  a.kind = d->kind;
  a.link = d->link;
  a.poses = poses;
  a.tgt = d->target;
  a.J0 = d->jac0;
  for (int k = 0; k < 3; ++k) a.Ju[k] = d->jac_u[k];
  a.vals = d->vals;
  a.cjac = d->cjac;
  a.lo = d->lo;
  a.hi = d->hi;
  a.damping = d->damping;
  a.max_step = d->max_step;
  a.margin = d->margin;
  a.wpen = std::sqrt(d->penalty);
  a.q = q;
  return a;
}
void ik_checks(pmp_context* ctx, int64_t R, const pmp_ik_step_desc* d) {
  require_robot(ctx);
  PMP_REQUIRE(d != nullptr && R >= 0, "null step descriptor or negative batch");
  PMP_REQUIRE(d->kind == 3 || d->kind == 6 || d->kind == 7, "All poses must be 3 or 6 or 7 dimensional");
  PMP_REQUIRE(ctx->n_cols <= 16, "the batched IK step handles at most 16 q columns");
  PMP_REQUIRE(d->link >= 0 && d->link < ctx->n_nodes, "link index out of range");
  PMP_REQUIRE(d->n_cst >= 0 && d->n_cst <= 2, "n_cst out of range");
}
}  // namespace

extern "C" {

pmp_status pmp_ik_step_batch(pmp_context* ctx, double* q, int64_t R, const double* poses,
                             const pmp_ik_step_desc* d) {
  return guarded([&] {
    ik_checks(ctx, R, d);
    if (R == 0) return;
    PMP_REQUIRE(poses != nullptr, "null argument");
    const IkArgs a = ik_args(ctx, q, R, poses, d);
    set_device(ctx);
    PMP_HIP(launch_ik_step(a, st(ctx)));
  });
}

pmp_status pmp_ik_graph(pmp_context* ctx, double* q, int64_t R, const pmp_ik_step_desc* d, int32_t n_iter,
                        pmp_graph** out) {
  return guarded([&] {
    PMP_REQUIRE(out != nullptr, "null output pointer");
    ik_checks(ctx, R, d);
    PMP_REQUIRE(R > 0 && n_iter > 0, "graph of an empty IK batch");
    IkArgs a = ik_args(ctx, q, R, nullptr, d);
    set_device(ctx);
    auto* g = new pmp_graph();
    try {
      // the link pose comes from the Jacobian launch's FP32 walk, into the
      // graph's own buffer
      float* pose = (float*)g->ik_pose.get(sizeof(float) * 7 * (size_t)R);
      a.pose32 = pose;
      hipStream_t cs = graph_stream(ctx);
      g->stream = cs;
      EvalArgs e{};
      e.vals = const_cast<float*>(d->vals);
      e.jac = const_cast<float*>(d->cjac);
      e.whole_chunks = ctx->sched.eval_whole_chunks;
      PMP_HIP(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal));
      hipError_t le = hipSuccess;
      for (int it = 0; it < n_iter && le == hipSuccess; ++it) {
        const float origin[3] = {0.f, 0.f, 0.f};
        le = launch_link_jacobian(ctx->k32, q, R, R, d->link, origin, const_cast<float*>(d->jac0), pose, cs);
        for (int k = 0; k < 3 && d->kind > 3 && le == hipSuccess; ++k) {
          const float unit[3] = {k == 0 ? 1.f : 0.f, k == 1 ? 1.f : 0.f, k == 2 ? 1.f : 0.f};
          le = launch_link_jacobian(ctx->k32, q, R, R, d->link, unit, const_cast<float*>(d->jac_u[k]), nullptr, cs);
        }
        if (d->n_cst > 0 && le == hipSuccess) le = launch_eval_units(ctx->k32, ctx->s32, q, R, R, e, cs);
        if (le == hipSuccess) le = launch_ik_step(a, cs);
      }
      hipError_t ce = hipStreamEndCapture(cs, &g->graph);
      PMP_HIP(le);
      PMP_HIP(ce);
      PMP_HIP(hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0));
      g->ctx = ctx;
      g->tables_gen = ctx->tables_gen;
    } catch (...) {
      if (g->exec) (void)hipGraphExecDestroy(g->exec);
      if (g->graph) (void)hipGraphDestroy(g->graph);
      g->ik_pose.release();
      delete g;
      throw;
    }
    *out = g;
  });
}

pmp_status pmp_sphere_clearance_batch(pmp_context* ctx, const double* q, int64_t ldq, int64_t B, int32_t accurate,
                                      float* clearance) {
  return guarded([&] {
    PMP_BATCH_CHECKS(q, ldq, B);
    PMP_REQUIRE(B == 0 || clearance, "null output");
    if (B == 0 || ctx->n_spheres == 0) return;
    set_device(ctx);
    const int64_t S = ctx->n_spheres, Q = S * B;
    float* cen = ctx->n_clouds ? (float*)ctx->ws_cen.get(sizeof(float) * 3 * Q) : nullptr;
    PMP_HIP(launch_clearance(ctx->k32, ctx->s32, q, ldq, B, accurate ? 0 : 1, clearance, cen, st(ctx)));
    for (const CloudRange& c : ctx->clouds) {  // nearest point of every cloud, no cap
      float* d2 = (float*)ctx->ws_cloud.get(sizeof(float) * Q);
      PMP_HIP(launch_cloud_min(ctx->s32.tiles, ctx->s32.tileB, ctx->s32.supers, c, cen, nullptr, B, Q, d2, st(ctx)));
      PMP_HIP(launch_cloud_clearance(clearance, d2, (const float*)ctx->rad_out.ptr, c.radius, B, S, st(ctx)));
    }
  });
}

pmp_status pmp_is_valid_batch(pmp_context* ctx, const double* q, int64_t ldq, int64_t B, uint8_t* valid,
                              int32_t* n_recheck) {
  return guarded([&] {
    PMP_BATCH_CHECKS(q, ldq, B);
    PMP_REQUIRE(B == 0 || valid, "null output");
    if (B == 0) return;
    set_device(ctx);
    CloudWork cw{};
    int32_t* flags = nullptr;
    if (ctx->n_clouds) {
      const int64_t Q = (int64_t)ctx->k32.n_ext_spheres * B;
      cw.B = B;
      cw.cen = (float*)ctx->ws_cen.get(sizeof(float) * std::max<int64_t>(3 * Q, 1));
      float* w = (float*)ctx->ws_cloud.get(sizeof(float) * std::max<int64_t>(4 * Q * ctx->n_clouds, 1));
      for (int oc = 0; oc < ctx->n_clouds; ++oc) {  // per cloud: cap2, rr, then the d2 bracket (upper, lower)
        cw.cap2[oc] = w + 4 * Q * oc;
        cw.rr[oc] = cw.cap2[oc] + Q;
      }
      // the cloud pass's flags are zero on entry and left zero: cleared when
      // the buffer is new
      const bool fresh_flags = ctx->ws_flags.bytes < sizeof(int32_t) * (size_t)B;
      flags = (int32_t*)ctx->ws_flags.get(sizeof(int32_t) * B);
      if (fresh_flags) PMP_HIP(launch_fill_i32(flags, 0, (int64_t)(ctx->ws_flags.bytes / sizeof(int32_t)), st(ctx)));
    }
    // batches up to kValidUnitsMaxBatch: a lane per (configuration, unit);
    // the flags start at zero (cleared when the buffer is new) and the pass
    // leaves them so
    int32_t* unit_flags = nullptr;
    if (B <= kValidUnitsMaxBatch && ctx->k32.n_units > 0) {
      const bool fresh = ctx->ws_vunit.bytes < sizeof(int32_t) * (size_t)kValidUnitsMaxBatch;
      unit_flags = (int32_t*)ctx->ws_vunit.get(sizeof(int32_t) * kValidUnitsMaxBatch);
      if (fresh) PMP_HIP(launch_fill_i32(unit_flags, 0, kValidUnitsMaxBatch, st(ctx)));
    }
    PMP_HIP(launch_valid(ctx->k32, ctx->k64, ctx->s32, ctx->s64, q, ldq, B, ctx->band, valid, n_recheck,
                         ctx->sched.block, cw, ctx->clouds.data(), ctx->n_clouds, flags, unit_flags, st(ctx)));
  });
}

pmp_status pmp_is_valid_batch_noalloc(pmp_context* ctx, const double* q, int64_t ldq, int64_t B, uint8_t* valid,
                                      int32_t* n_recheck) {
  return guarded([&] {
    PMP_BATCH_CHECKS(q, ldq, B);
    PMP_REQUIRE(B == 0 || valid, "null output");
    if (B == 0) return;
    set_device(ctx);
    unsigned char* ws = (unsigned char*)noalloc_scratch(ctx->device);
    PMP_REQUIRE(ws != nullptr, "internal: no-allocation scratch not set up on this device");
    // The same workspaces as pmp_is_valid_batch, carved from the static
    // scratch (256-byte aligned pieces).
    auto carve = [&](size_t& at, size_t bytes) {
      unsigned char* p = ws + at;
      at += (bytes + 255) & ~(size_t)255;
      return p;
    };
    size_t need = 0;
    CloudWork cw{};
    int32_t* flags = nullptr;
    int32_t* unit_flags = nullptr;
    if (ctx->n_clouds) {
      const int64_t Q = (int64_t)ctx->k32.n_ext_spheres * B;
      cw.B = B;
      cw.cen = (float*)carve(need, sizeof(float) * 3 * (size_t)Q);
      float* w = (float*)carve(need, sizeof(float) * 4 * (size_t)Q * ctx->n_clouds);
      for (int oc = 0; oc < ctx->n_clouds; ++oc) {
        cw.cap2[oc] = w + 4 * Q * oc;
        cw.rr[oc] = cw.cap2[oc] + Q;
      }
      flags = (int32_t*)carve(need, sizeof(int32_t) * (size_t)B);
    }
    if (B <= kValidUnitsMaxBatch && ctx->k32.n_units > 0) unit_flags = (int32_t*)carve(need, sizeof(int32_t) * (size_t)B);
    PMP_REQUIRE(need <= kNoAllocScratchBytes,
                "batch needs " + std::to_string(need) + " bytes of scratch, more than the " +
                    std::to_string(kNoAllocScratchBytes) +
                    " bytes pmp_is_valid_batch_noalloc has; split the batch or call pmp_is_valid_batch");
    // the scratch keeps the previous call's contents: the unit pass and the
    // cloud pass need their flags at zero (fill kernels, not memset nodes)
    if (unit_flags) PMP_HIP(launch_fill_i32(unit_flags, 0, B, st(ctx)));
    if (flags) PMP_HIP(launch_fill_i32(flags, 0, B, st(ctx)));
    PMP_HIP(launch_valid(ctx->k32, ctx->k64, ctx->s32, ctx->s64, q, ldq, B, ctx->band, valid, n_recheck,
                         ctx->sched.block, cw, ctx->clouds.data(), ctx->n_clouds, flags, unit_flags, st(ctx)));
  });
}

pmp_status pmp_evaluate_batch(pmp_context* ctx, const double* q, int64_t ldq, int64_t B, float* vals, float* jac) {
  return guarded([&] {
    PMP_BATCH_CHECKS(q, ldq, B);
    PMP_REQUIRE(B == 0 || (vals && jac), "null output");
    if (B == 0) return;
    set_device(ctx);
    EvalArgs e{};
    e.vals = vals;
    e.jac = jac;
    e.whole_chunks = ctx->sched.eval_whole_chunks;
    PMP_HIP(launch_eval_units(ctx->k32, ctx->s32, q, ldq, B, e, st(ctx)));
  });
}

namespace {
// Argument checks shared by the direct and the captured trajectory call.
void check_traj_args(pmp_context* ctx, const double* X, int64_t ldx, int64_t N, const int64_t* offsets,
                     int32_t n_traj, float* vals, float* jac, float* cost, float* grad) {
  PMP_BATCH_CHECKS(X, ldx, N);
  PMP_REQUIRE(n_traj >= 0 && offsets != nullptr, "missing trajectory offsets");
  PMP_REQUIRE(offsets[0] == 0 && offsets[n_traj] == N, "offsets must start at 0 and end at N");
  for (int32_t i = 0; i < n_traj; ++i) PMP_REQUIRE(offsets[i + 1] >= offsets[i], "offsets must not decrease");
  PMP_REQUIRE(N == 0 || (vals && jac && grad), "null output");
  PMP_REQUIRE(n_traj == 0 || cost, "null output");
}
}  // namespace

pmp_status pmp_traj_cost_grad_graph(pmp_context* ctx, const double* X, int64_t ldx, int64_t N, const int64_t* offsets,
                                    int32_t n_traj, const double* weights, float* vals, float* jac, float* cost,
                                    float* grad, pmp_graph** out) {
  return guarded([&] {
    PMP_REQUIRE(out != nullptr, "null output pointer");
    check_traj_args(ctx, X, ldx, N, offsets, n_traj, vals, jac, cost, grad);
    PMP_REQUIRE(n_traj > 0, "graph of an empty trajectory batch");
    set_device(ctx);
    auto* g = new pmp_graph();
    try {
      // the graph's own copies of the offsets and weights, in place before capture
      std::vector<int64_t> off(offsets, offsets + n_traj + 1);
      std::vector<double> w2(std::max(ctx->n_cols, 1), 1.0);
      for (int j = 0; j < ctx->n_cols; ++j) w2[j] = weights ? weights[j] * weights[j] : 1.0;
      g->traj_off.upload(off, st(ctx));
      g->traj_w2.upload(w2, st(ctx));
      PMP_HIP(hipStreamSynchronize(st(ctx)));
      hipStream_t cs = graph_stream(ctx);
      g->stream = cs;
      EvalArgs e{};
      e.vals = vals;
      e.jac = jac;
      e.offsets = (const int64_t*)g->traj_off.ptr;
      e.n_traj = n_traj;
      e.w2 = (const double*)g->traj_w2.ptr;
      e.cost = cost;
      e.grad = grad;
      e.whole_chunks = ctx->sched.eval_whole_chunks;
      PMP_HIP(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal));
      hipError_t le = launch_eval_units(ctx->k32, ctx->s32, X, ldx, N, e, cs);
      hipError_t ce = hipStreamEndCapture(cs, &g->graph);
      PMP_HIP(le);
      PMP_HIP(ce);
      PMP_HIP(hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0));
      g->ctx = ctx;
      g->tables_gen = ctx->tables_gen;
    } catch (...) {
      if (g->exec) (void)hipGraphExecDestroy(g->exec);
      if (g->graph) (void)hipGraphDestroy(g->graph);
      g->traj_off.release();
      g->traj_w2.release();
      delete g;
      throw;
    }
    *out = g;
  });
}

pmp_status pmp_traj_cost_grad_batch(pmp_context* ctx, const double* X, int64_t ldx, int64_t N, const int64_t* offsets,
                                    int32_t n_traj, const double* weights, float* vals, float* jac, float* cost,
                                    float* grad) {
  return guarded([&] {
    check_traj_args(ctx, X, ldx, N, offsets, n_traj, vals, jac, cost, grad);
    if (n_traj == 0) return;
    set_device(ctx);
    std::vector<int64_t> off(offsets, offsets + n_traj + 1);
    std::vector<double> w2(std::max(ctx->n_cols, 1), 1.0);
    for (int j = 0; j < ctx->n_cols; ++j) w2[j] = weights ? weights[j] * weights[j] : 1.0;
    // the tables are re-sent only when they change (repeated solver iterations reuse them)
    if (off != ctx->h_traj_off) {
      ctx->traj_off.upload(off, st(ctx));
      ctx->h_traj_off = off;
    }
    if (w2 != ctx->h_traj_w2) {
      ctx->traj_w2.upload(w2, st(ctx));
      ctx->h_traj_w2 = w2;
    }
    const int64_t* doff = (const int64_t*)ctx->traj_off.ptr;
    const double* dw = (const double*)ctx->traj_w2.ptr;
    EvalArgs e{};
    e.vals = vals;
    e.jac = jac;
    e.offsets = doff;
    e.n_traj = n_traj;
    e.w2 = dw;
    e.cost = cost;
    e.grad = grad;
    e.whole_chunks = ctx->sched.eval_whole_chunks;
    PMP_HIP(launch_eval_units(ctx->k32, ctx->s32, X, ldx, N, e, st(ctx)));
  });
}

}  // extern "C"

namespace {
int cst_dim_of(const pmp_context* ctx) { return (ctx->s32.n_prims > 0 ? 1 : 0) + (ctx->k32.n_pairs > 0 ? 1 : 0); }

// Row, nonzero and column-pointer offsets of every trajectory's piece of the
// SQP inequality sequence (the last entry is the total).
void sqp_sizes(const pmp_context* ctx, const int64_t* offsets, int32_t n_traj, bool with_msbox, int64_t* row_off,
               int64_t* nnz_off, int64_t* colptr_off) {
  PMP_REQUIRE(ctx != nullptr && offsets != nullptr && row_off && nnz_off && colptr_off, "null argument");
  PMP_REQUIRE(n_traj >= 0 && offsets[0] == 0, "offsets must start at 0");
  const int64_t D = ctx->n_cols, C = cst_dim_of(ctx);
  row_off[0] = nnz_off[0] = colptr_off[0] = 0;
  for (int32_t i = 0; i < n_traj; ++i) {
    const int64_t T = offsets[i + 1] - offsets[i];
    PMP_REQUIRE(T >= 0, "offsets must not decrease");
    const int64_t segs = T > 0 && with_msbox ? T - 1 : 0;
    row_off[i + 1] = row_off[i] + C * T + 2 * D * segs;
    nnz_off[i + 1] = nnz_off[i] + D * T * C + 4 * D * segs;
    colptr_off[i + 1] = colptr_off[i] + D * T + 1;
    PMP_REQUIRE(nnz_off[i + 1] - nnz_off[i] < ((int64_t)1 << 31), "trajectory too long for int32 indices");
  }
}
}  // namespace

extern "C" {

pmp_status pmp_traj_sqp_sizes(pmp_context* ctx, const int64_t* offsets, int32_t n_traj, int32_t with_msbox,
                              int64_t* row_off, int64_t* nnz_off, int64_t* colptr_off) {
  return guarded([&] { sqp_sizes(ctx, offsets, n_traj, with_msbox != 0, row_off, nnz_off, colptr_off); });
}

pmp_status pmp_traj_sqp_ineq(pmp_context* ctx, const double* X, int64_t ldx, int64_t N, const int64_t* offsets,
                             int32_t n_traj, const double* msbox, double* c, int32_t* colptr, int32_t* rowind,
                             double* values) {
  return guarded([&] {
    PMP_BATCH_CHECKS(X, ldx, N);
    PMP_REQUIRE(n_traj >= 0 && offsets != nullptr, "missing trajectory offsets");
    PMP_REQUIRE(offsets[0] == 0 && offsets[n_traj] == N, "offsets must start at 0 and end at N");
    for (int32_t i = 0; i < n_traj; ++i) PMP_REQUIRE(offsets[i + 1] >= offsets[i], "offsets must not decrease");
    PMP_REQUIRE(n_traj == 0 || colptr, "null output");
    PMP_REQUIRE(N == 0 || (c && rowind && values), "null output");
    if (n_traj == 0) return;
    set_device(ctx);
    const int D = ctx->n_cols, C = cst_dim_of(ctx);
    std::vector<int64_t> row_off(n_traj + 1), nnz_off(n_traj + 1), cp_off(n_traj + 1);
    sqp_sizes(ctx, offsets, n_traj, msbox != nullptr, row_off.data(), nnz_off.data(), cp_off.data());
    // offsets, row / nnz offsets and box widths in one small table, uploaded
    // (a synchronous copy) only when it changes, and before the evaluate
    // launch so that the copy never waits for it
    std::vector<int64_t> tab;
    tab.insert(tab.end(), offsets, offsets + n_traj + 1);
    tab.insert(tab.end(), row_off.begin(), row_off.end());
    tab.insert(tab.end(), nnz_off.begin(), nnz_off.end());
    const size_t box_at = tab.size();
    tab.resize(box_at + D);
    if (msbox) std::memcpy(&tab[box_at], msbox, sizeof(double) * D);
    if (tab != ctx->h_sqp_tables || ctx->sqp_tables.ptr == nullptr) {
      ctx->sqp_tables.upload(tab, st(ctx));
      ctx->h_sqp_tables = tab;
    }
    const int64_t* dt = (const int64_t*)ctx->sqp_tables.ptr;
    // evaluate rows of every waypoint (the plain evaluate launch over N)
    float* rows = (float*)ctx->sqp_rows.get(sizeof(float) * std::max<int64_t>(1, (int64_t)C * (D + 1) * N));
    if (N > 0 && C > 0) {
      EvalArgs e{};
      e.vals = rows;
      e.jac = rows + (int64_t)C * N;
      e.whole_chunks = ctx->sched.eval_whole_chunks;
      PMP_HIP(launch_eval_units(ctx->k32, ctx->s32, X, ldx, N, e, st(ctx)));
    }
    SqpArgs a{};
    a.X = X;
    a.ldx = ldx;
    a.N = N;
    a.D = D;
    a.C = C;
    a.offsets = dt;
    a.n_traj = n_traj;
    a.msbox = msbox ? (const double*)(dt + box_at) : nullptr;
    a.vals = rows;
    a.jac = rows + (int64_t)C * N;
    a.row_off = dt + (n_traj + 1);
    a.nnz_off = dt + 2 * (n_traj + 1);
    a.c = c;
    a.colptr = colptr;
    a.rowind = rowind;
    a.values = values;
    PMP_HIP(launch_sqp_ineq(a, st(ctx)));
  });
}

pmp_status pmp_pair_clearance_batch(pmp_context* ctx, const double* q, int64_t ldq, int64_t B, float* dist,
                                    float* grad) {
  return guarded([&] {
    PMP_BATCH_CHECKS(q, ldq, B);
    PMP_REQUIRE(B == 0 || dist, "null output");
    if (B == 0 || ctx->n_spheres == 0 || ctx->s32.n_prims == 0) return;
    set_device(ctx);
    PMP_HIP(launch_pairs(ctx->k32, ctx->s32, q, ldq, B, dist, grad, ctx->sched.pair_shares, st(ctx)));
  });
}

pmp_status pmp_pair_clearance_vjp(pmp_context* ctx, const double* q, int64_t ldq, int64_t B, const float* grad_dist,
                                  float* grad_q) {
  return guarded([&] {
    PMP_BATCH_CHECKS(q, ldq, B);
    PMP_REQUIRE(B == 0 || (grad_dist && grad_q), "null argument");
    if (B == 0) return;
    set_device(ctx);
    const int shares = vjp_shares(ctx->k32, ctx->s32.n_prims, B, ctx->sched.pair_shares);
    float* part = (float*)ctx->ws_vjp.get(
        sizeof(float) * std::max<int64_t>(1, (int64_t)ctx->k32.n_items * shares * ctx->n_cols * B));
    PMP_HIP(launch_vjp(ctx->k32, ctx->s32, q, ldq, B, grad_dist, part, grad_q, ctx->sched.pair_shares, st(ctx)));
  });
}

}  // extern "C"

namespace {

void prepare_edges(pmp_context* ctx, const double* s1, const double* s2, int64_t lds, int64_t E,
                   const pmp_validator_desc* v, uint8_t* valid, int32_t* n_checked, EdgeBufs& eb, EdgeArgs& ea) {
  require_robot(ctx);
  PMP_REQUIRE(E >= 0, "negative edge count");
  PMP_REQUIRE(E == 0 || (s1 && s2 && valid), "null argument");
  PMP_REQUIRE(lds >= E, "leading dimension smaller than the edge count");
  PMP_REQUIRE(E < (int64_t)1 << 31, "too many edges");
  PMP_REQUIRE(v != nullptr, "null validator");
  const int D = ctx->n_cols;
  std::vector<double> w(std::max(D, 1), 1.0), iw(std::max(D, 1), 1.0);
  if (v->type == PMP_VALIDATOR_BOX) {
    PMP_REQUIRE(v->box_width != nullptr, "box dimension and space dimension mismatch");
    for (int i = 0; i < D; ++i) {
      w[i] = v->box_width[i];
      iw[i] = 1.0 / v->box_width[i];  // BoxMotionValidator ctor (motion_validator.cpp:81-87)
    }
  } else {
    PMP_REQUIRE(v->type == PMP_VALIDATOR_EUCLIDEAN, "unknown validator type");
  }
  set_device(ctx);
  ea.s1 = s1;
  ea.s2 = s2;
  ea.lds = lds;
  ea.E = E;
  ea.vtype = v->type;
  ea.resolution = v->resolution;
  // The width tables change only with the validator: a call that repeats the
  // last one's widths enqueues no copy (and, the copies being synchronous,
  // does not wait for the stream).
  if (eb.h_width != w || eb.width.ptr == nullptr) {
    eb.width.upload(w, st(ctx));
    eb.inv_width.upload(iw, st(ctx));
    eb.h_width = w;
  }
  ea.width = (const double*)eb.width.ptr;
  ea.inv_width = (const double*)eb.inv_width.ptr;
  ea.seq = (const uint8_t*)ctx->seq.ptr;
  // The kernels leave the counters at zero after every launch; a new buffer
  // starts there.
  const bool fresh = eb.counter.bytes < sizeof(int32_t) * kNumEdgeCounters;  // new (or grown) buffer
  ea.counter = (int32_t*)eb.counter.get(sizeof(int32_t) * kNumEdgeCounters);
  if (fresh) PMP_HIP(launch_fill_i32(ea.counter, 0, kNumEdgeCounters, st(ctx)));
  const int64_t Ea = std::max<int64_t>(E, 1);
  int32_t* pend = (int32_t*)eb.pend.get(sizeof(int32_t) * Ea * (3 * kPendStates + 3));
  ea.pend = pend;
  ea.pend_info = pend + Ea * kPendStates;
  ea.plist = pend + Ea * (kPendStates + 2);
  ea.pend_pair = pend + Ea * (kPendStates + 3);
  ea.valid = valid;
  ea.n_checked = n_checked;
}

// Persistent grid: waves_per_cu waves on every CU (default 16, four per
// SIMD), never more waves than edges.
// The schedule's waves per CU are capped at what the edge kernel variant can
// hold at once, so a grid with a wave for every edge (whose waves take their
// edge by index) is launched only when all of it is resident; otherwise the
// resident grid pulls edges from the counter.
int edge_grid(pmp_context* ctx, int block, int64_t E) {
  const int occ = edges_resident_waves_per_cu(ctx->s32.n_cloud > 0, block, slot_bytes(ctx->k64, block));
  int wpc = ctx->sched.waves_per_cu > 0 ? ctx->sched.waves_per_cu : 16;
  if (occ > 0) wpc = std::min(wpc, occ);
  const int64_t by_edges = (E * 64 + block - 1) / block;
  return (int)std::max<int64_t>(1, std::min<int64_t>((int64_t)ctx->n_cu * wpc * 64 / block, by_edges));
}

}  // namespace

extern "C" {

pmp_status pmp_check_edges_batch(pmp_context* ctx, const double* s1, const double* s2, int64_t lds, int64_t E,
                                 const pmp_validator_desc* v, uint8_t* valid, int32_t* n_checked) {
  return guarded([&] {
    EdgeArgs ea{};
    prepare_edges(ctx, s1, s2, lds, E, v, valid, n_checked, ctx->eb, ea);
    if (E == 0) return;
    const int block = slot_block(ctx->k64, ctx->sched.block > 0 ? ctx->sched.block : 256);
    if (!ctx->edge_ev0) {
      PMP_HIP(hipEventCreate(&ctx->edge_ev0));
      PMP_HIP(hipEventCreate(&ctx->edge_ev1));
    }
    PMP_HIP(launch_edges(ctx->k32, ctx->k64, ctx->s32, ctx->s64, ea, ctx->band, edge_grid(ctx, block, E), block,
                         st(ctx), ctx->edge_ev0, ctx->edge_ev1));
  });
}

pmp_status pmp_check_edges_graph(pmp_context* ctx, const double* s1, const double* s2, int64_t lds, int64_t E,
                                 const pmp_validator_desc* v, uint8_t* valid, int32_t* n_checked, pmp_graph** out) {
  return guarded([&] {
    PMP_REQUIRE(out != nullptr, "null output pointer");
    PMP_REQUIRE(E > 0, "graph of an empty edge batch");
    auto* g = new pmp_graph();
    try {
      EdgeArgs ea{};
      prepare_edges(ctx, s1, s2, lds, E, v, valid, n_checked, g->eb, ea);
      PMP_HIP(hipStreamSynchronize(st(ctx)));  // width tables are in place before capture
      hipStream_t cs = graph_stream(ctx);
      const int block = slot_block(ctx->k64, ctx->sched.block > 0 ? ctx->sched.block : 256);
      g->stream = cs;
      PMP_HIP(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal));
      hipError_t le = launch_edges(ctx->k32, ctx->k64, ctx->s32, ctx->s64, ea, ctx->band, edge_grid(ctx, block, E),
                                   block, cs);
      hipError_t ce = hipStreamEndCapture(cs, &g->graph);
      PMP_HIP(le);
      PMP_HIP(ce);
      PMP_HIP(hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0));
      g->ctx = ctx;
      g->tables_gen = ctx->tables_gen;
    } catch (...) {
      if (g->exec) (void)hipGraphExecDestroy(g->exec);
      if (g->graph) (void)hipGraphDestroy(g->graph);
      g->eb.release();
      delete g;
      throw;
    }
    *out = g;
  });
}

pmp_status pmp_graph_launch(pmp_graph* g) {
  return guarded([&] {
    PMP_REQUIRE(g != nullptr, "null graph");
    PMP_REQUIRE(g->ctx->tables_gen == g->tables_gen,
                "the context's robot, scene or band changed after this graph was captured; capture it again");
    PMP_HIP(hipGraphLaunch(g->exec, g->stream));
  });
}

pmp_status pmp_graph_destroy(pmp_graph* g) {
  return guarded([&] {
    if (!g) return;
    (void)hipStreamSynchronize(g->stream);
    if (g->exec) PMP_HIP(hipGraphExecDestroy(g->exec));
    if (g->graph) PMP_HIP(hipGraphDestroy(g->graph));
    g->eb.release();
    g->traj_off.release();
    g->traj_w2.release();
    g->ik_pose.release();
    delete g;
  });
}

pmp_status pmp_edge_recheck_count(pmp_context* ctx, int32_t* n) {
  return guarded([&] {
    PMP_REQUIRE(ctx != nullptr && n != nullptr, "null argument");
    PMP_REQUIRE(ctx->eb.counter.ptr != nullptr, "no edge batch has run on this context");
    set_device(ctx);
    PMP_HIP(hipMemcpyAsync(n, (const int32_t*)ctx->eb.counter.ptr + kCntRecheck, sizeof(int32_t), hipMemcpyDeviceToHost,
                           st(ctx)));
    PMP_HIP(hipStreamSynchronize(st(ctx)));
  });
}

pmp_status pmp_edge_kernel_ms(pmp_context* ctx, float* ms) {
  return guarded([&] {
    PMP_REQUIRE(ctx != nullptr && ms != nullptr, "null argument");
    PMP_REQUIRE(ctx->edge_ev0 != nullptr, "no edge batch has run on this context");
    set_device(ctx);
    PMP_HIP(hipEventSynchronize(ctx->edge_ev1));
    PMP_HIP(hipEventElapsedTime(ms, ctx->edge_ev0, ctx->edge_ev1));
  });
}

pmp_status pmp_prims_evaluate(pmp_context* ctx, const pmp_scene_desc* prims, const double* points, int64_t N,
                              int32_t mode, double radius, double* out) {
  return guarded([&] {
    PMP_REQUIRE(ctx != nullptr && prims != nullptr, "null argument");
    PMP_REQUIRE(N >= 0, "negative point count");
    PMP_REQUIRE(N == 0 || (points && out), "null argument");
    PMP_REQUIRE(mode == 0 || mode == 1, "mode must be 0 (evaluate) or 1 (is_outside)");
    PMP_REQUIRE(prims->n_prims >= 1 && prims->n_prims <= kMaxPrims, "primitive count out of range");
    std::vector<DevPrim<float>> p32;
    std::vector<DevPrim<double>> p64;
    std::vector<float> c32;
    std::vector<double> c64;
    int nc = 0;
    pack_prims(*prims, p32, p64, c32, c64, nc);
    if (N == 0) return;
    set_device(ctx);
    auto* dp = ctx->tmp_prims.upload(p64, st(ctx));
    auto* dc = ctx->tmp_cloud.upload(c64, st(ctx));
    PMP_HIP(launch_prims_eval(dp, prims->n_prims, dc, points, N, mode, radius, out, st(ctx)));
  });
}

pmp_status pmp_nearest_batch(pmp_context* ctx, const double* points, int64_t N, const double* queries, int64_t M,
                             double* sqdist, int32_t* index) {
  return guarded([&] {
    PMP_REQUIRE(ctx != nullptr, "null context");
    PMP_REQUIRE(N >= 1 && N < (int64_t)1 << 31, "point count out of range");
    PMP_REQUIRE(M >= 0, "negative query count");
    PMP_REQUIRE(M == 0 || (points && queries && sqdist && index), "null argument");
    if (M == 0) return;
    set_device(ctx);
    PMP_HIP(launch_nearest(points, N, queries, M, sqdist, index, st(ctx)));
  });
}

pmp_status pmp_cloud_tile_sizes(int64_t n_points, int64_t* n_tiles, int64_t* n_supers, int64_t* scratch_bytes) {
  return guarded([&] {
    PMP_REQUIRE(n_points >= 1 && n_points < (int64_t)1 << 30, "point count out of range");
    const CloudBuildSizes z = cloud_build_sizes(n_points);
    if (n_tiles) *n_tiles = z.n_tiles;
    if (n_supers) *n_supers = z.n_supers;
    if (scratch_bytes) *scratch_bytes = z.scratch_bytes;
  });
}

pmp_status pmp_cloud_build_tiles(pmp_context* ctx, const double* points, int64_t n_points, void* tiles,
                                 void* tiles64, float* tile_b, double* points64, void* supers, void* scratch) {
  return guarded([&] {
    PMP_REQUIRE(ctx != nullptr, "null context");
    PMP_REQUIRE(n_points >= 1 && n_points < (int64_t)1 << 30, "point count out of range");
    PMP_REQUIRE(points && tiles && tiles64 && tile_b && points64 && supers && scratch, "null argument");
    set_device(ctx);
    PMP_HIP(launch_cloud_build(points, n_points, 0, (CloudTile*)tiles, (CloudTile64*)tiles64, tile_b, points64,
                               (CloudSuper*)supers, scratch, st(ctx)));
  });
}

}  // extern "C"
