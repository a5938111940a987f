// Lane-per-configuration forward kinematics.
//
// One lane owns one configuration.  The node loop is wave-uniform: node
// records, sphere tables and primitives are read with uniform indices
// (scalar loads into SGPRs), while the pose chain and the per-configuration
// results stay in VGPRs.  Poses that later nodes (or the self-collision pass)
// need are kept in slots whose indices the host assigns (wave-uniform); the
// slots live in LDS (SlotFile).
//
// The chain follows KinematicModel::set_joint_angles_impl and
// build_cache_until(_inner) (kinematics/algorithm.cpp:109-210):
//   revolute   local q = origin_identity ? J : origin_q * J, J = (s axis, c)
//              of the half angle; composed without the identity shortcut
//   prismatic  local t = origin_t + axis q, local rotation identity (shortcut)
//   fixed      static local transform with its identity flag
//   translation-only links (consider_rotation = false) take the parent
//   rotation matrix times the local offset and keep the parent rotation.
#pragma once

#include "geom.hpp"

namespace pmp {

__device__ __forceinline__ bool wave_any(bool p) { return __any(p); }

// Error of the tile-local FP32 form of a squared distance to a cloud point,
//   d2 = |c'|^2 + (|p'|^2 - 2 c'.p'),   c' = c - o, p' = p - o,
// against the exact |c - p|^2 (c the FP32 query, p the FP64 point, o the
// tile's FP32 centre).  With u = 2^-24, a = |c'| and rho >= |p - o| for every
// point of the tile, the rounding of c', p' and |p'|^2 (u each), the four
// products and sums of the K = 4 dot product (gamma_4 = 4u of
// 2 a rho + rho^2), |c'|^2 (3u a^2) and the final sum (u d2), using
// |c - p| <= a + rho, add up to at most 8u (a + rho)^2 <= 16u (a^2 + rho^2).
// The band takes twice that, 2e-6 (a^2 + rho^2), in case the matrix cores
// round less tightly than IEEE single precision.  For the compact tiles of a
// dense cloud (a, rho of a few cm) it is ~1e-8 m^2; for tiles spread over
// metres it exceeds the 1e-5 m validity band, so the decisions that use it
// carry it as a per-tile term.
constexpr float kCloudGemmErr = 2e-6f;

// rho^2 of a tile: the squared distance from its centre o to the farthest
// corner of its (outward-rounded) bounds.
template <class Tile>
__device__ __forceinline__ float tile_rho2(const Tile& T) {
  const float ex = fmaxf(T.ub[0] - T.o[0], T.o[0] - T.lb[0]);
  const float ey = fmaxf(T.ub[1] - T.o[1], T.o[1] - T.lb[1]);
  const float ez = fmaxf(T.ub[2] - T.o[2], T.o[2] - T.lb[2]);
  return ex * ex + ey * ey + ez * ez;
}

// Visitors that declare `static constexpr bool kBlocks = true` receive each
// node's sphere blocks through blocks(node, pose) instead of one sphere() call
// per sphere.
template <class V, class = void>
struct uses_blocks {
  static constexpr bool value = false;
};
template <class V>
struct uses_blocks<V, decltype((void)V::kBlocks)> {
  static constexpr bool value = V::kBlocks;
};

// Visitors that declare `static constexpr bool kRanged = true` walk nodes
// only up to vis.last_node() and visit only the self pairs
// [vis.pair_lo(), vis.pair_hi()).
template <class V, class = void>
struct is_ranged {
  static constexpr bool value = false;
};
template <class V>
struct is_ranged<V, decltype((void)V::kRanged)> {
  static constexpr bool value = V::kRanged;
};

// Visitors that declare `static constexpr bool kPairCull = true` provide
// bool pair_near(V3<T> g1, V3<T> g2, T reach): whether, in this lane, the
// sphere pairs of a group pair whose bounding spheres (world centres g1, g2,
// summed radii `reach`) are this far apart can still matter; a group pair
// that matters to no live lane is skipped.
template <class V, class = void>
struct is_pair_cull {
  static constexpr bool value = false;
};
template <class V>
struct is_pair_cull<V, decltype((void)V::kPairCull)> {
  static constexpr bool value = V::kPairCull;
};

// Kept poses (ancestors of later branches, links of self pairs) live in LDS:
// the launch reserves slot_bytes(ka, block) bytes of dynamic shared memory
// from byte offset ka.slot_off, ka.slot_wave_bytes per wave, laid out
// [wave][slot][component][lane] so that the 64 lanes of a wave touch 64
// consecutive words (no bank conflicts) and nothing goes to scratch.  A wave
// owns its chunk: a kernel that runs FK in FP32 and then in FP64 in the same
// wave (the edge pass) sizes the chunk for FP64 and reuses it.  Slot indices
// are wave-uniform; each lane reads only what it wrote, so no barrier is
// needed.
template <class T>
struct SlotFile {
  T* base;
  __device__ __forceinline__ explicit SlotFile(const KinArgs<T>& ka) {
    HIP_DYNAMIC_SHARED(unsigned char, pmp_dyn_lds)
    base = reinterpret_cast<T*>(pmp_dyn_lds + ka.slot_off + (threadIdx.x >> 6) * ka.slot_wave_bytes) +
           (threadIdx.x & 63);
  }
  __device__ __forceinline__ Pose<T> get(int i) const {
    const T* p = base + i * 7 * 64;
    Pose<T> r;
    r.q = {p[0], p[64], p[128], p[192]};
    r.t = {p[256], p[320], p[384]};
    return r;
  }
  __device__ __forceinline__ void set(int i, const Pose<T>& P) {
    T* p = base + i * 7 * 64;
    p[0] = P.q.x;
    p[64] = P.q.y;
    p[128] = P.q.z;
    p[192] = P.q.w;
    p[256] = P.t.x;
    p[320] = P.t.y;
    p[384] = P.t.z;
  }
};

// Configuration sources.  FK reads each column once, in FP64, and converts.
struct BatchQ {  // q[col * ldq + b]
  const double* __restrict__ q;
  int64_t ldq, b;
  __device__ __forceinline__ double get(int col) const { return q[(int64_t)col * ldq + b]; }
};
// Waypoint of an edge: OMPL RealVectorStateSpace::interpolate, from + (to - from) * t,
// or the end state itself (CustomValidatorBase::checkMotion checks s2 directly,
// ompl/motion_validator.cpp:52).
struct EdgeQ {
  const double* __restrict__ s1;
  const double* __restrict__ s2;
  int64_t lds, e;
  double t;
  bool at_end;
  __device__ __forceinline__ double get(int col) const {
    double a = cload(s1, (int64_t)col * lds + e), b = cload(s2, (int64_t)col * lds + e);
    return at_end ? b : a + (b - a) * t;
  }
};

// Base pose: the constant base pose for a fixed base, else built from the
// trailing columns of q as ConstraintBase::update_kintree does
// (constraints/primitive.hpp:38-62).
template <class T, class QS>
__device__ __forceinline__ Pose<T> base_pose(const KinArgs<T>& ka, const QS& qs) {
  Pose<T> base;
  if (ka.base_type == PMP_BASE_FIXED) {
    base.q = {ka.base_q[0], ka.base_q[1], ka.base_q[2], ka.base_q[3]};
    base.t = {ka.base_t[0], ka.base_t[1], ka.base_t[2]};
  } else if (ka.base_type == PMP_BASE_PLANAR) {
    int n = ka.n_q_cols;
    base.t = {(T)qs.get(n), (T)qs.get(n + 1), T(0)};
    base.q = quat_from_rpy<T>(T(0), T(0), (T)qs.get(n + 2));
  } else {
    int n = ka.n_q_cols;
    base.t = {(T)qs.get(n), (T)qs.get(n + 1), (T)qs.get(n + 2)};
    base.q = quat_from_rpy<T>((T)qs.get(n + 3), (T)qs.get(n + 4),
                              (T)qs.get(n + 5));
  }
  return base;
}

// Visitor interface:
//   bool alive() const                          lane still has work
//   void begin(const Pose<T>& base, const QS&)  base pose of this configuration
//   void node(int n, const DevNode<T>&, const Pose<T>&)
//   void sphere(int s, V3<T> c, T r)            s indexes the node-sorted table
//   void blocks(const DevNode<T>&, const Pose<T>&)  instead of sphere(), with kBlocks
//   void self(const DevSelfPair&, int i, int j, V3<T> ci, V3<T> cj, T rsum)
// `taylor` selects the collision check's sin/cos (a runtime, wave-uniform flag).
template <class T, class QS, class Vis>
__device__ __forceinline__ void run_fk(const KinArgs<T>& ka, const QS& qs, bool taylor, Vis& vis) {
  SlotFile<T> slots(ka);
  const Pose<T> base = base_pose(ka, qs);
  vis.begin(base, qs);
  Pose<T> prev = base;
  int n_walk = ka.n_nodes;
  if constexpr (is_ranged<Vis>::value) n_walk = vis.last_node() + 1 < n_walk ? vis.last_node() + 1 : n_walk;
  for (int n = 0; n < n_walk; ++n) {
    if (!wave_any(vis.alive())) return;
    const DevNode<T> nd = cload(ka.nodes, n);
    // parent_slot: -2 the previous node, -1 the base pose, >= 0 a kept pose
    const Pose<T> par = nd.parent_slot == -2 ? prev : (nd.parent_slot == -1 ? base : slots.get(nd.parent_slot));
    const V3<T> lt{nd.trans[0], nd.trans[1], nd.trans[2]};
    const Q4<T> oq{nd.quat[0], nd.quat[1], nd.quat[2], nd.quat[3]};
    Pose<T> P;
    switch (nd.kind) {
      case PMP_NODE_ROOT:
        P = base;
        break;
      case PMP_NODE_FIXED:
        P = compose(par, oq, lt, nd.identity != 0);
        break;
      case PMP_NODE_REVOLUTE: {
        T x = (T)qs.get(nd.q_col) * T(0.5);
        T s, c;
        if (taylor)
          taylor_sincos<T>(x, s, c);
        else
          exact_sincos(x, s, c);
        Q4<T> jq{s * nd.axis[0], s * nd.axis[1], s * nd.axis[2], c};
        Q4<T> lq = nd.identity ? jq : qmul(oq, jq);
        P = compose(par, lq, lt, false);
        break;
      }
      case PMP_NODE_PRISMATIC: {
        T v = (T)qs.get(nd.q_col);
        V3<T> t2{lt.x + nd.axis[0] * v, lt.y + nd.axis[1] * v, lt.z + nd.axis[2] * v};
        P = compose(par, oq, t2, true);
        break;
      }
      default: {  // PMP_NODE_TRANSLATION
        M3<T> R = qmat(par.q);
        P.q = par.q;
        P.t = add(par.t, mv(R, lt));
        break;
      }
    }
    if (nd.out_slot >= 0) slots.set(nd.out_slot, P);
    prev = P;
    vis.node(n, nd, P);
    if constexpr (uses_blocks<Vis>::value) {
      vis.blocks(nd, P);
    } else if (nd.sphere_begin < nd.sphere_end) {
      const M3<T> R = qmat(P.q);
      for (int s = nd.sphere_begin; s < nd.sphere_end; ++s) {
        const DevSphere<T> sp = cload(ka.spheres, s);
        V3<T> c = add(mv(R, V3<T>{sp.x, sp.y, sp.z}), P.t);
        vis.sphere(s, c, sp.r);
      }
    }
  }
  // Self-collision sphere pairs, group pair by group pair
  // (constraints/primitive_sphere_collision.cpp:222-254, 324-383).
  if constexpr (!Vis::kSelf) return;
  int k0 = 0, k1 = ka.n_pairs;
  if constexpr (is_ranged<Vis>::value) {
    k0 = vis.pair_lo();
    k1 = vis.pair_hi();
  }
  for (int k = k0; k < k1; ++k) {
    if (!wave_any(vis.alive())) return;
    const DevSelfPair pr = cload(ka.pairs, k);
    const Pose<T> P1 = slots.get(pr.slot1), P2 = slots.get(pr.slot2);
    const M3<T> R1 = qmat(P1.q), R2 = qmat(P2.q);
    if constexpr (is_pair_cull<Vis>::value) {
      // the groups' bounding spheres; |R v| <= grow |v| for a rotation
      // matrix of a quaternion with |q|^2 = n, grow = 1 + 2 |n - 1|
      const V3<T> g1 = add(mv(R1, V3<T>{pr.bound1[0], pr.bound1[1], pr.bound1[2]}), P1.t);
      const V3<T> g2 = add(mv(R2, V3<T>{pr.bound2[0], pr.bound2[1], pr.bound2[2]}), P2.t);
      const T n1 = P1.q.x * P1.q.x + P1.q.y * P1.q.y + P1.q.z * P1.q.z + P1.q.w * P1.q.w;
      const T n2 = P2.q.x * P2.q.x + P2.q.y * P2.q.y + P2.q.z * P2.q.z + P2.q.w * P2.q.w;
      const T dn = fabs(n1 - T(1)) > fabs(n2 - T(1)) ? fabs(n1 - T(1)) : fabs(n2 - T(1));
      const T reach = (T(pr.bound1[3]) + T(pr.bound2[3])) * (T(1) + T(2) * dn);
      if (!wave_any(vis.alive() && vis.pair_near(g1, g2, reach))) continue;
    }
    for (int i = pr.begin1; i < pr.end1; ++i) {
      const DevSphere<T> a = cload(ka.spheres, i);
      const V3<T> ci = add(mv(R1, V3<T>{a.x, a.y, a.z}), P1.t);
      for (int j = pr.begin2; j < pr.end2; ++j) {
        const DevSphere<T> bb = cload(ka.spheres, j);
        const V3<T> cj = add(mv(R2, V3<T>{bb.x, bb.y, bb.z}), P2.t);
        vis.self(pr, i, j, ci, cj, a.r + bb.r);
      }
    }
  }
}

}  // namespace pmp
