// Device-side table layout (internal to the library; the public boundary is
// include/plainmp_amd.h).  Tables are uploaded once per robot / scene and read
// by every lane with wave-uniform indices, so the compiler serves them from
// the scalar data cache into SGPRs; per-configuration state lives in VGPRs.
#pragma once

#include <stdint.h>

#include "../../include/plainmp_amd.h"

namespace pmp {

constexpr int kMaxSlots = 8;      // live ancestor poses kept in registers during FK
constexpr int kMaxCols = 32;      // q columns (joints + base)
constexpr int kMaxNodes = 256;
constexpr int kMaxSpheres = 4096;
constexpr int kMaxPrims = 512;
constexpr int kMaxSelfPairs = 1024;

template <class T>
struct alignas(16) DevNode {
  int32_t kind;
  int32_t parent_slot;   // -2: the previous node, -1: the base pose, >= 0: a kept pose
  int32_t out_slot;      // -1: pose not kept
  int32_t q_col;         // joint column, -1 otherwise
  int32_t sphere_begin;  // node-sorted sphere range
  int32_t sphere_end;
  int32_t identity;      // local rotation skipped on compose
  uint32_t col_mask;     // q columns that move this node (reference rptable_ row)
  T quat[4];
  T trans[3];
  T axis[3];
  T pad[2];
};

template <class T>
struct alignas(16) DevSphere {
  T x, y, z, r;
};

struct alignas(16) DevSphereMeta {
  int32_t out;    // index in SphereCollisionCst group order
  int32_t ext;    // 1 when checked against primitives (group not only-self)
  int32_t node;
  int32_t group;
};

struct alignas(16) DevSelfPair {
  int32_t slot1, begin1, end1, slot2;
  int32_t begin2, end2, node1, node2;
};

// Kernel arguments describing the robot (passed by value: kernarg -> SGPRs).
template <class T>
struct KinArgs {
  const DevNode<T>* nodes;
  const DevSphere<T>* spheres;
  const DevSphereMeta* meta;
  const DevSelfPair* pairs;
  int32_t n_nodes, n_spheres, n_pairs, n_cols;
  int32_t n_q_cols, base_type, n_ext_spheres;
  uint32_t rev_mask;  // q columns that rotate (revolute joints, base yaw/roll/pitch)
  T base_q[4];
  T base_t[3];
};

template <class T>
struct SceneArgs {
  const void* prims;  // DevPrim<T>*
  int32_t n_prims;
  int32_t n_cloud;    // number of cloud primitives (must be last in the list)
  const T* cloud_xyz; // cloud points [n][4] (x, y, z, |p|^2) for all clouds, concatenated
};

}  // namespace pmp
