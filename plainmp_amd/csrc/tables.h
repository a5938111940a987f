// Device-side table layout (internal to the library; the public boundary is
// include/plainmp_amd.h).  Tables are uploaded once per robot / scene and read
// by every lane with wave-uniform indices, so the compiler serves them from
// the scalar data cache into SGPRs; per-configuration state lives in VGPRs.
#pragma once

#include <stdint.h>

#include "../../include/plainmp_amd.h"

namespace pmp {

constexpr int kMaxSlots = 32;     // live poses kept during FK (LDS; the device's LDS may allow fewer)
constexpr int kMaxCols = 64;      // q columns (joints + base)
constexpr int kMaxNodes = 256;
constexpr int kMaxSpheres = 4096;
constexpr int kMaxPrims = 512;
constexpr int kMaxSelfPairs = 1024;
constexpr int kMaxClouds = 4;
constexpr int kCloudTile = 64;  // points per cloud tile
constexpr int kBlockSpheres = 8;  // spheres of one link tested together against a primitive
// XCDs of an MI355X; workgroups of a launch are handed to them in turn.  Only
// the locality of index mappings depends on it.
constexpr int kXcds = 8;

// Primitive kinds of the validity pass; the scene keeps a copy of its
// primitives sorted by kind so that each kind is a loop without branches.
enum PrimKind { kKindBoxAA = 0, kKindBoxRot, kKindCylZ, kKindCylRot, kKindSphere, kKindGround, kKindCloud, kNumKinds };

template <class T>
struct alignas(16) DevNode {
  int32_t kind;
  int32_t parent_slot;   // -2: the previous node, -1: the base pose, >= 0: a kept pose
  int32_t out_slot;      // -1: pose not kept
  int32_t q_col;         // joint column, -1 otherwise
  int32_t sphere_begin;  // node-sorted sphere range
  int32_t sphere_end;
  int32_t identity;      // local rotation skipped on compose
  int32_t block_begin;   // sphere blocks of this node (primitive-checked spheres only)
  int32_t block_end;
  int32_t ipad;
  uint64_t col_mask;     // q columns that move this node (reference rptable_ row)
  T quat[4];
  T trans[3];
  T axis[3];
  T n2_lo, n2_hi;  // bounds of |q|^2 of this node's world rotation assumed by the cull table
};

template <class T>
struct alignas(16) DevSphere {
  T x, y, z, r;
};

// Up to kBlockSpheres primitive-checked spheres of one link with a bounding
// sphere (g, G) in the link frame: G >= |s_k - g| + r_k for every member.
struct alignas(16) DevSphereBlock {
  int32_t n;
  int32_t ext[kBlockSpheres];  // index among the primitive-checked spheres
  int32_t sph[kBlockSpheres];  // index in the node-sorted sphere tables
  int32_t pad[3];
  float g[3], G;
  float s[kBlockSpheres][4];   // link-frame centre and radius
};

// A candidate primitive of a sphere block: the primitive's AABB grown by the
// block's reach (bounding radius times the rotation growth bound of its node,
// plus the band and a rounding slack), rounded outwards to FP32.
struct alignas(16) CullEntry {
  float lb[3];
  int32_t prim;  // index into the kind-sorted primitives
  float ub[3];
  int32_t orig;  // index in the scene's own primitive order
};
// Per block: its candidates of kind k are cull[at[k], at[k + 1]).
struct alignas(16) CullRange {
  int32_t at[8];
};
static_assert(kNumKinds + 1 == 8, "CullRange holds one offset per kind plus the end");

// A run of at most kItemSpheres consecutive node-sorted spheres of one node
// (every sphere, primitive-checked or not): the unit of work of the pair and
// adjoint kernels, which give each (configuration, item) its own lane.
constexpr int kItemSpheres = 8;
struct alignas(16) DevSphereItem {
  int32_t node, begin, end, pad;
};

// A unit of the evaluate kernels: the sphere blocks [b_lo, b_hi) against
// their candidate primitives and the self pairs [k_lo, k_hi), walking FK to
// node `last`.
struct alignas(16) DevEvalUnit {
  int32_t b_lo, b_hi, k_lo, k_hi;
  int32_t last, pad[3];
};

struct alignas(16) DevSphereMeta {
  int32_t out;    // index in SphereCollisionCst group order
  int32_t ext;    // index among the spheres checked against primitives, -1 for only-self groups
  int32_t node;
  int32_t group;
};

// Point-cloud tile: AABB and centre (FP32, for culling and the local frame of
// the matrix-core distances) and FP64 bounds for the exact re-decision.
struct alignas(16) CloudTile {
  float lb[3], ub[3], o[3];
  int32_t n;
};
struct alignas(16) CloudTile64 {
  double lb[3], ub[3];
};
// Record sizes are part of pmp_cloud_build_tiles' output contract.
static_assert(sizeof(CloudTile) == 48 && sizeof(CloudTile64) == 48, "cloud tile record layout");
// kCloudSuper consecutive (Morton-ordered) tiles under one FP32 AABB: a wave
// tests a super-tile once before its tiles.
constexpr int kCloudSuper = 16;
struct alignas(16) CloudSuper {
  float lb[3], ub[3];
  int32_t first, count;  // tile range [first, first + count)
};
static_assert(sizeof(CloudSuper) == 32, "cloud super-tile record layout");

// Where the FP32 validity pass leaves sphere-vs-cloud queries.
struct CloudWork {
  float* cen;                 // [n_ext][3][B] sphere centres
  float* cap2[kMaxClouds];    // [n_ext][B] squared reach (r + rc + band)^2, < 0 when not queried
  float* rr[kMaxClouds];      // [n_ext][B] r + rc
  int64_t B;
};

struct alignas(16) DevSelfPair {
  int32_t slot1, begin1, end1, slot2;
  int32_t begin2, end2, node1, node2;
  float bound1[4], bound2[4];  // each group's bounding sphere in its link frame: centre, radius (rounded up)
};

// Kernel arguments describing the robot (passed by value: kernarg -> SGPRs).
template <class T>
struct KinArgs {
  const DevNode<T>* nodes;
  const DevSphere<T>* spheres;
  const DevSphereMeta* meta;
  const DevSelfPair* pairs;
  const DevSphereBlock* blocks;
  const DevSphereItem* items;
  int32_t n_items;
  const DevEvalUnit* units;  // one unit per sphere block, then one per self pair
  int32_t n_units;
  const DevEvalUnit* unit_all;  // one unit: every sphere and self pair
  int32_t n_nodes, n_spheres, n_pairs, n_cols;
  int32_t n_q_cols, base_type, n_ext_spheres;
  uint64_t rev_mask;  // q columns that rotate (revolute joints, base yaw/roll/pitch)
  int32_t n_slots;    // kept-pose slots the FK walk uses
  int32_t slot_off;         // byte offset of the kept-pose slots in the launch's dynamic LDS
  int32_t slot_wave_bytes;  // kept-pose bytes per wave (n_slots * 7 * 64 * sizeof(T), or more)
  int32_t lds_limit;        // LDS bytes one workgroup may use on this device (launch sizing)
  T base_q[4];
  T base_t[3];
};

// Dynamic LDS a launch of `block` lanes reserves for the kept poses of FK in
// precision T (nothing for a robot that keeps none).
template <class T>
inline size_t slot_bytes(const KinArgs<T>& ka, int block) {
  return (size_t)ka.slot_wave_bytes * (size_t)(block / 64);
}
// The largest workgroup (a multiple of 64 lanes, at most `want`) whose
// kept-pose slots in precision T fit the LDS left beside `own` bytes; 64 at
// least (the robot upload checks that 64 FP64 lanes fit).
template <class T>
inline int slot_block(const KinArgs<T>& ka, int want, size_t own = 0) {
  int b = want;
  while (b > 64 && own + slot_bytes(ka, b) > (size_t)ka.lds_limit) b -= 64;
  return b;
}

template <class T>
struct SceneArgs {
  const void* prims;            // DevPrim<T>*; clouds carry their tile range in cloud_begin / cloud_count
  int32_t n_prims;
  int32_t n_cloud;              // number of cloud primitives
  const CloudTile* tiles;       // all clouds' tiles
  const CloudSuper* supers;     // all clouds' super-tiles
  int32_t super_at[kMaxClouds + 1];  // cloud i's super-tiles: supers[super_at[i], super_at[i + 1])
  const CloudTile64* tiles64;
  const float* tileB;           // [tile][4 MFMA][4 k][16 j] B operands in tile-local coordinates
  const double* pts64;          // [tile][64][4] FP64 world points (padding points far away)
  const void* vprims;           // DevPrim<T>* sorted by PrimKind (validity pass)
  int32_t vrange[kNumKinds + 1];  // kind k occupies vprims[vrange[k], vrange[k + 1])
  const CullEntry* cull;        // candidates of every sphere block (FP32 validity pass)
  const CullRange* cull_range;  // [n_blocks]
  const CullEntry* ecull;       // the same for evaluate: boxes grown by the 0.1 m cutoff
  const CullRange* ecull_range;
  const int32_t* vorig;         // kind-sorted index -> scene order
};

}  // namespace pmp
