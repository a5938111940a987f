// gfx950 kernels of the FK + sphere-SDF collision path.
//
// All batch kernels use one lane per configuration (or per edge waypoint):
// a wave carries 64 configurations through the same node / sphere /
// primitive loop, the tables are wave-uniform, and results are written in
// "batch index fastest" layout so that every store of a wave is one
// coalesced 256-byte line.
#include <float.h>
#include <limits.h>

#include "fk.hpp"
#include "launch.h"

namespace pmp {

constexpr float kCutoff = 0.1f;  // SphereCollisionCst::cutoff_dist_ (primitive_sphere_collision.hpp:115)
// Evaluate batches of at least this many 64-configuration chunks walk each
// configuration as one unit (the launch is large enough without splitting).
constexpr int64_t kWholeUnitChunks = 512;
// Edges are walked 64 states at a time; an edge needing more states than this
// (a step ratio below 2^-24) is reported invalid with n_checked = -1 instead.
constexpr double kMaxEdgeStates = 16777216.0;

template <class T>
__device__ __forceinline__ DevPrim<T> prim_at(const SceneArgs<T>& sc, int p) {
  return cload(reinterpret_cast<const DevPrim<T>*>(sc.prims), p);
}

__device__ __forceinline__ int64_t gid() { return (int64_t)blockIdx.x * blockDim.x + threadIdx.x; }

// Squared distance from c to the box [lb, ub] (0 inside).
__device__ __forceinline__ float box_dist2(const float (&lb)[3], const float (&ub)[3], V3<float> c) {
  const float dx = fmaxf(fmaxf(lb[0] - c.x, c.x - ub[0]), 0.f);
  const float dy = fmaxf(fmaxf(lb[1] - c.y, c.y - ub[1]), 0.f);
  const float dz = fmaxf(fmaxf(lb[2] - c.z, c.z - ub[2]), 0.f);
  return dx * dx + dy * dy + dz * dz;
}

// Squared distance between the boxes [lb, ub] and [l, u] (0 when they meet).
__device__ __forceinline__ float box_box_dist2(const float (&lb)[3], const float (&ub)[3], V3<float> l, V3<float> u) {
  const float dx = fmaxf(fmaxf(lb[0] - u.x, l.x - ub[0]), 0.f);
  const float dy = fmaxf(fmaxf(lb[1] - u.y, l.y - ub[1]), 0.f);
  const float dz = fmaxf(fmaxf(lb[2] - u.z, l.z - ub[2]), 0.f);
  return dx * dx + dy * dy + dz * dz;
}

// Minimum / maximum of v over the wave (every lane active; wave-uniform).
__device__ __forceinline__ float wave_minf(float v) {
#pragma unroll
  for (int s = 32; s > 0; s >>= 1) v = fminf(v, __shfl_xor(v, s, 64));
  return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(v)));
}
__device__ __forceinline__ float wave_maxf(float v) {
#pragma unroll
  for (int s = 32; s > 0; s >>= 1) v = fmaxf(v, __shfl_xor(v, s, 64));
  return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(v)));
}

// Minimum of v over the wave (every lane active; wave-uniform result).
__device__ __forceinline__ int wave_min(int v) {
#pragma unroll
  for (int s = 32; s > 0; s >>= 1) v = min(v, __shfl_xor(v, s, 64));
  return __builtin_amdgcn_readfirstlane(v);
}

// Maximum of v over the wave's active lanes (wave-uniform result).
__device__ __forceinline__ int wave_max_i(int v) {
  for (int m = 32; m > 0; m >>= 1) v = max(v, __shfl_xor(v, m, 64));
  return __builtin_amdgcn_readfirstlane(v);
}

// Four consecutive floats of a table read as one wave-uniform load.
struct alignas(16) Quad4 {
  float v[4];
};

// One kept tile t of cloud_block_min: the spheres' minima over its points.
template <bool kBounds>
__device__ __forceinline__ void cloud_tile_min(const SceneArgs<float>& sc, int t, const CloudTile& T,
                                               const V3<float> (&c)[kBlockSpheres], int n,
                                               float (&best)[kBlockSpheres], float (&lo)[kBlockSpheres]) {
  V3<float> cl[kBlockSpheres];
  float tb[kBlockSpheres];
#pragma unroll
  for (int k = 0; k < kBlockSpheres; ++k) {
    cl[k] = V3<float>{c[k].x - T.o[0], c[k].y - T.o[1], c[k].z - T.o[2]};
    tb[k] = FLT_MAX;
  }
  const Quad4* Bt = reinterpret_cast<const Quad4*>(sc.tileB + (int64_t)t * 4 * kCloudTile);
  for (int m = 0; m < 4; ++m) {
    for (int j4 = 0; j4 < 4; ++j4) {
      // points 16 m + 4 j4 .. + 3: components x, y, z, |p'|^2
      const Quad4 X = cload(Bt, m * 16 + 0 * 4 + j4), Y = cload(Bt, m * 16 + 1 * 4 + j4);
      const Quad4 Z = cload(Bt, m * 16 + 2 * 4 + j4), W = cload(Bt, m * 16 + 3 * 4 + j4);
#pragma unroll
      for (int jj = 0; jj < 4; ++jj) {
#pragma unroll
        for (int k = 0; k < kBlockSpheres; ++k) {
          if (k < n)
            tb[k] = fminf(tb[k], fmaf(X.v[jj], cl[k].x, fmaf(Y.v[jj], cl[k].y, fmaf(Z.v[jj], cl[k].z, W.v[jj]))));
        }
      }
    }
  }
  const float rho2 = kBounds ? tile_rho2(T) : 0.f;
#pragma unroll
  for (int k = 0; k < kBlockSpheres; ++k) {
    const float off = dot(cl[k], cl[k]);
    if constexpr (kBounds) {
      const float err = kCloudGemmErr * (off + rho2);
      best[k] = fminf(best[k], tb[k] + off + err);
      lo[k] = fminf(lo[k], tb[k] + off - err);
    } else {
      best[k] = fminf(best[k], tb[k] + off);
    }
  }
}

// Per sphere k < n of a block, min |c_k - p|^2 over the points of cloud oc
// that lie in tiles within sqrt(reach2) of the block's bounding centre g
// (FLT_MAX when none).  Super-tiles, then tiles, are skipped for the whole
// wave when no `live` lane's g reaches their box: 64 of them at a time are
// first tested against the box around the live lanes' g, a box a lane, and
// the survivors of that ballot are tested lane by lane.  A kept tile's points
// are compared in tile-local coordinates, |c - p|^2 = |c'|^2 + (|p'|^2 -
// 2 c'.p'), from the matrix-core pass's operands (-2 p', |p'|^2).  With
// kBounds the minimum is bracketed: best = min over tiles of (d2 + err), lo =
// min over tiles of (d2 - err), err = kCloudGemmErr (|c'|^2 + rho^2) of the
// tile.
template <bool kBounds>
__device__ __forceinline__ void cloud_block_min(const SceneArgs<float>& sc, int oc, const V3<float> (&c)[kBlockSpheres],
                                                int n, V3<float> g, float reach2, bool live,
                                                float (&best)[kBlockSpheres], float (&lo)[kBlockSpheres]) {
#pragma unroll
  for (int k = 0; k < kBlockSpheres; ++k) best[k] = lo[k] = FLT_MAX;
  if (!wave_any(live)) return;
  const V3<float> gl{wave_minf(live ? g.x : FLT_MAX), wave_minf(live ? g.y : FLT_MAX), wave_minf(live ? g.z : FLT_MAX)};
  const V3<float> gu{wave_maxf(live ? g.x : -FLT_MAX), wave_maxf(live ? g.y : -FLT_MAX),
                     wave_maxf(live ? g.z : -FLT_MAX)};
  const int lane = threadIdx.x & 63;
  const int s_end = sc.super_at[oc + 1];
  for (int s0 = sc.super_at[oc]; s0 < s_end; s0 += 64) {
    bool snear = false;
    if (s0 + lane < s_end) {
      const CloudSuper Sv = sc.supers[s0 + lane];  // a super-tile a lane: vector loads
      snear = box_box_dist2(Sv.lb, Sv.ub, gl, gu) <= reach2;
    }
    for (uint64_t smask = __ballot(snear); smask; smask &= smask - 1) {
      const CloudSuper S = cload(sc.supers, s0 + __builtin_ctzll(smask));
      if (!wave_any(live && box_dist2(S.lb, S.ub, g) <= reach2)) continue;
      const int t_end = S.first + S.count;
      for (int t0 = S.first; t0 < t_end; t0 += 64) {
        bool tnear = false;
        if (t0 + lane < t_end) {
          const CloudTile Tv = sc.tiles[t0 + lane];
          tnear = box_box_dist2(Tv.lb, Tv.ub, gl, gu) <= reach2;
        }
        for (uint64_t tmask = __ballot(tnear); tmask; tmask &= tmask - 1) {
          const int t = t0 + __builtin_ctzll(tmask);
          const CloudTile T = cload(sc.tiles, t);
          if (!wave_any(live && box_dist2(T.lb, T.ub, g) <= reach2)) continue;
          cloud_tile_min<kBounds>(sc, t, T, c, n, best, lo);
        }
      }
    }
  }
}

// The nearest point of cloud oc to c, per lane (an exhaustive search over
// the tiles, skipping those whose box is farther than the best so far), and
// its squared distance in best.
__device__ __forceinline__ V3<float> cloud_nearest(const SceneArgs<float>& sc, int oc, V3<float> c, float& best) {
  best = FLT_MAX;
  V3<float> bp = c;
  for (int si = sc.super_at[oc]; si < sc.super_at[oc + 1]; ++si) {
    const CloudSuper S = sc.supers[si];
    if (box_dist2(S.lb, S.ub, c) > best) continue;
    for (int t = S.first; t < S.first + S.count; ++t) {
      const CloudTile T = sc.tiles[t];
      if (box_dist2(T.lb, T.ub, c) > best) continue;
      const V3<float> cl{c.x - T.o[0], c.y - T.o[1], c.z - T.o[2]};
      const float off = dot(cl, cl);
      const float* Bt = sc.tileB + (int64_t)t * 4 * kCloudTile;
      for (int j = 0; j < kCloudTile; ++j) {
        const float* col = Bt + (j / 16) * 64 + (j % 16);
        const float bx = col[0], by = col[16], bz = col[32], bw = col[48];
        const float d2 = fmaf(bx, cl.x, fmaf(by, cl.y, fmaf(bz, cl.z, bw))) + off;
        if (d2 < best) {
          best = d2;
          bp = V3<float>{T.o[0] - 0.5f * bx, T.o[1] - 0.5f * by, T.o[2] - 0.5f * bz};
        }
      }
    }
  }
  return bp;
}

// ------------------------------------------------------------------ centres
template <class T>
struct CentersVis {
  static constexpr bool kSelf = false;
  const KinArgs<T>& ka;
  float* out;
  int64_t B, b;
  bool active;
  __device__ bool alive() const { return true; }
  template <class QS>
  __device__ void begin(const Pose<T>&, const QS&) {}
  __device__ void node(int, const DevNode<T>&, const Pose<T>&) {}
  __device__ void sphere(int s, V3<T> c, T) {
    if (!active) return;
    const int o = cload(ka.meta, s).out;
    out[(int64_t)(o * 3 + 0) * B + b] = (float)c.x;
    out[(int64_t)(o * 3 + 1) * B + b] = (float)c.y;
    out[(int64_t)(o * 3 + 2) * B + b] = (float)c.z;
  }
  __device__ void self(const DevSelfPair&, int, int, V3<T>, V3<T>, T) {}
};

template <class T>
__global__ void __launch_bounds__(256) centers_kernel(KinArgs<T> ka, const double* q, int64_t ldq, int64_t B,
                                                      int taylor, float* out) {
  const int64_t b = gid();
  const bool active = b < B;
  BatchQ qs{q, ldq, active ? b : B - 1};
  CentersVis<T> v{ka, out, B, b, active};
  run_fk(ka, qs, taylor != 0, v);
}

// ------------------------------------------------------------- link poses
struct PosesVis {
  static constexpr bool kSelf = false;
  double* out;
  int64_t B, b;
  bool active;
  __device__ bool alive() const { return true; }
  template <class QS>
  __device__ void begin(const Pose<double>&, const QS&) {}
  __device__ void node(int n, const DevNode<double>&, const Pose<double>& P) {
    if (!active) return;
    double v[7] = {P.t.x, P.t.y, P.t.z, P.q.x, P.q.y, P.q.z, P.q.w};
#pragma unroll
    for (int k = 0; k < 7; ++k) out[(int64_t)(n * 7 + k) * B + b] = v[k];
  }
  __device__ void sphere(int, V3<double>, double) {}
  __device__ void self(const DevSelfPair&, int, int, V3<double>, V3<double>, double) {}
};

__global__ void __launch_bounds__(256) poses_kernel(KinArgs<double> ka, const double* q, int64_t ldq, int64_t B,
                                                    int taylor, double* out) {
  const int64_t b = gid();
  const bool active = b < B;
  BatchQ qs{q, ldq, active ? b : B - 1};
  PosesVis v{out, B, b, active};
  run_fk(ka, qs, taylor != 0, v);
}

// -------------------------------------------------------------- clearance
struct ClearanceVis {
  static constexpr bool kSelf = false;
  const KinArgs<float>& ka;
  const SceneArgs<float>& sc;
  float* out;
  int64_t B, b;
  bool active;
  float* cen;  // [S][3][B] centres for the cloud pass, or null
  __device__ bool alive() const { return true; }
  template <class QS>
  __device__ void begin(const Pose<float>&, const QS&) {}
  __device__ void node(int, const DevNode<float>&, const Pose<float>&) {}
  __device__ void sphere(int s, V3<float> c, float r) {
    float best = INFINITY;
    for (int p = 0; p < sc.n_prims; ++p) {
      const DevPrim<float>& P = prim_at(sc, p);
      if (P.type == PMP_PRIM_CLOUD) continue;
      best = fminf(best, sdf_eval(P, c) - r);
    }
    if (!active) return;
    const int64_t o = cload(ka.meta, s).out;
    out[o * B + b] = best;
    if (cen) {
      cen[(o * 3 + 0) * B + b] = c.x;
      cen[(o * 3 + 1) * B + b] = c.y;
      cen[(o * 3 + 2) * B + b] = c.z;
    }
  }
  __device__ void self(const DevSelfPair&, int, int, V3<float>, V3<float>, float) {}
};

__global__ void __launch_bounds__(256) clearance_kernel(KinArgs<float> ka, SceneArgs<float> sc, const double* q,
                                                        int64_t ldq, int64_t B, int taylor, float* out, float* cen) {
  const int64_t b = gid();
  const bool active = b < B;
  BatchQ qs{q, ldq, active ? b : B - 1};
  ClearanceVis v{ka, sc, out, B, b, active, cen};
  run_fk(ka, qs, taylor != 0, v);
}

// --------------------------------------------------------------- validity
// FP32 pass: a collision decided outside the band ends the lane's work
// ("definite"); a decision inside the band marks the configuration for the
// FP64 re-decision ("uncertain").
// kInlineCloud: clouds are tested inside the walk (the edge pass, whose
// states have no batch-wide cloud pass of their own); otherwise the walk
// writes cloud queries for the matrix-core pass when use_cw is set.
// kFirstHit: the lanes are states of one edge and only the first certain
// collision in visiting order (`key`) matters, so a lane whose key is past
// the wave's first certain collision so far stops taking part.
template <bool kInlineCloud, bool kFirstHit = false>
struct ValidVis32T {
  static constexpr bool kSelf = true;
  static constexpr bool kBlocks = true;
  const KinArgs<float>& ka;
  const SceneArgs<float>& sc;
  float band;
  CloudWork cw{};        // cloud queries are written here when use_cw (the scene holds clouds)
  bool use_cw = false;
  int64_t b = 0;
  int key = 0;  // visiting index of the lane's state (kFirstHit)
  bool idle = false;
  bool definite = false;
  bool uncertain = false;
  // Candidates (block, primitive) found in the band; 2 also stands for an
  // uncertainty not tied to one candidate (a self pair, a rotation outside
  // the cull table's bounds).  With exactly one, the FP64 pass can re-decide
  // the state on that candidate alone: every other pair was decided outside
  // the band.
  int n_inb = 0;
  int inb_block = -1, inb_prim = -1;
  __device__ bool alive() const { return !definite && !idle; }
  template <class QS>
  __device__ void begin(const Pose<float>&, const QS&) {}
  __device__ void node(int, const DevNode<float>&, const Pose<float>&) {}
  __device__ void sphere(int, V3<float>, float) {}

  // One kind of primitive against the (up to eight) spheres of a block.  The
  // block's candidates carry the primitive's AABB grown by the block's reach
  // (bounding radius times the rotation growth bound, plus band and slack).
  // A candidate is skipped for the whole wave when no live lane's bounding
  // centre lies in that grown box: every member sphere is then outside the
  // AABB inflated by its radius plus the band, which is the reference's own
  // AABB pre-test failing with room to spare, so the skipped pairs can
  // neither collide nor fall in the band.  A candidate that survives is then
  // tested as a sphere of radius `reach` (the block's bounding radius times
  // the growth bound, plus band and a 1e-4 m slack) at the bounding centre,
  // against the primitive's own distance: the distance is 1-Lipschitz, so
  // a lane farther than that from the primitive has every member sphere
  // farther than its radius plus the band.  The ground's test is not a pure
  // distance (the reference's AABB slab is part of it) and skips this step.
  float reach = 0.0f;
  template <int K>
  __device__ __forceinline__ void kind_loop(const V3<float> (&c)[kBlockSpheres], const DevSphereBlock& bk, int n,
                                            V3<float> g, const CullRange& cr, int bi) {
    const DevPrim<float>* vp = reinterpret_cast<const DevPrim<float>*>(sc.vprims);
    for (int i = cr.at[K]; i < cr.at[K + 1]; ++i) {
      const CullEntry ce = cload(sc.cull, i);
      const bool out = g.x < ce.lb[0] || g.x > ce.ub[0] || g.y < ce.lb[1] || g.y > ce.ub[1] || g.z < ce.lb[2] ||
                       g.z > ce.ub[2];
      if (!wave_any(alive() && !out)) continue;
      const DevPrim<float> P = cload(vp, ce.prim);
      if constexpr (K != kKindGround) {
        bool nb = false;
        if (!wave_any(alive() && !out && collide_kind<K>(P, g, reach, 0.0f, nb))) continue;
      }
      bool ib = false;
#pragma unroll
      for (int k = 0; k < kBlockSpheres; ++k) {
        if (k < n) {
          bool inb = false;
          const bool col = collide_kind<K>(P, c[k], bk.s[k][3], band, inb);
          definite |= col && !inb;
          ib |= inb;
        }
      }
      if (ib) {
        uncertain = true;
        ++n_inb;
        inb_block = bi;
        inb_prim = ce.prim;
      }
    }
  }

  __device__ void blocks(const DevNode<float>& nd, const Pose<float>& P) {
    blocks_range(nd, P, nd.block_begin, nd.block_end);
  }
  // The node's sphere blocks [lo, hi).
  __device__ __forceinline__ void blocks_range(const DevNode<float>& nd, const Pose<float>& P, int lo, int hi) {
    if (lo >= hi) return;
    const M3<float> R = qmat(P.q);
    // The cull table assumes |q|^2 of this rotation within [n2_lo, n2_hi]
    // (so that |R v| <= grow |v|); a lane outside is left to the FP64 pass.
    const float n2 = P.q.x * P.q.x + P.q.y * P.q.y + P.q.z * P.q.z + P.q.w * P.q.w;
    if (!(n2 >= nd.n2_lo && n2 <= nd.n2_hi)) {
      uncertain = true;
      n_inb = 2;
    }
    const int n_solid = sc.vrange[kKindCloud];
    // |R v| <= growf |v| for a lane whose |q|^2 is in [n2_lo, n2_hi]
    const float growf = 1.0f + 2.0f * fmaxf(0.0f, fmaxf(nd.n2_hi - 1.0f, 1.0f - nd.n2_lo)) + 1e-6f;
    // A lane whose |q|^2 is in range has a bounded rotation: with a finite
    // translation none of its centres can be NaN.  The per-centre test runs
    // only when some live lane is not such a lane.
    const bool tfin = fabsf(P.t.x) <= FLT_MAX && fabsf(P.t.y) <= FLT_MAX && fabsf(P.t.z) <= FLT_MAX;
    const bool nan_test = n_solid > 0 && wave_any(alive() && !(n2 >= nd.n2_lo && n2 <= nd.n2_hi && tfin));
    for (int bi = lo; bi < hi; ++bi) {
      if constexpr (kFirstHit) {
        if (wave_any(definite && !idle)) {
          const int first = wave_min(definite && !idle ? key : INT_MAX);
          if (key > first) idle = true;
        }
      }
      if (!wave_any(alive())) return;
      const DevSphereBlock bk = cload(ka.blocks, bi);
      const int n = bk.n;
      V3<float> c[kBlockSpheres];
      bool bad = false;
#pragma unroll
      for (int k = 0; k < kBlockSpheres; ++k) {
        c[k] = P.t;
        if (k < n) c[k] = add(mv(R, V3<float>{bk.s[k][0], bk.s[k][1], bk.s[k][2]}), P.t);
      }
      if (nan_test) {
#pragma unroll
        for (int k = 0; k < kBlockSpheres; ++k)
          if (k < n) bad |= is_nan3(c[k]);
      }
      const V3<float> g = add(mv(R, V3<float>{bk.g[0], bk.g[1], bk.g[2]}), P.t);
      // NaN centres (NaN joint values: the reference's comparisons are all
      // false, so the state collides): left to the FP64 walk
      if (bad) {
        uncertain = true;
        n_inb = 2;
      }
      if (use_cw && !idle) {
        const int64_t B = cw.B;
#pragma unroll
        for (int k = 0; k < kBlockSpheres; ++k) {
          if (k < n) {
            const int64_t e = bk.ext[k];
            cw.cen[(e * 3 + 0) * B + b] = c[k].x;
            cw.cen[(e * 3 + 1) * B + b] = c[k].y;
            cw.cen[(e * 3 + 2) * B + b] = c[k].z;
          }
        }
        // clouds: queried by the matrix-core pass unless the AABB pre-test culls it
        const DevPrim<float>* vp = reinterpret_cast<const DevPrim<float>*>(sc.vprims);
        for (int p = sc.vrange[kKindCloud]; p < sc.vrange[kKindCloud + 1]; ++p) {
          const DevPrim<float> Pc = cload(vp, p);
          const int oc = (Pc.flags >> 4) & 15;
          // the cloud's output arrays, chosen with constant indices so that
          // the arguments stay in registers (an indexed kernel argument
          // would be copied to scratch)
          float* cap2 = cw.cap2[0];
          float* rro = cw.rr[0];
#pragma unroll
          for (int i = 1; i < kMaxClouds; ++i) {
            if (oc == i) {
              cap2 = cw.cap2[i];
              rro = cw.rr[i];
            }
          }
#pragma unroll
          for (int k = 0; k < kBlockSpheres; ++k) {
            if (k < n) {
              const float r = bk.s[k][3];
              const bool qd = alive() && !outside_aabb(Pc, c[k], r);
              const float Rr = r + Pc.p0;
              const int64_t at = (int64_t)bk.ext[k] * B + b;
              cap2[at] = qd ? (Rr + band) * (Rr + band) : -1.0f;
              rro[at] = qd ? Rr : -1.0f;
            }
          }
        }
      }
      if (!wave_any(alive())) continue;
      const CullRange cr = cload(sc.cull_range, bi);
      reach = bk.G * growf + band + 1e-4f;
      kind_loop<kKindBoxAA>(c, bk, n, g, cr, bi);
      kind_loop<kKindBoxRot>(c, bk, n, g, cr, bi);
      kind_loop<kKindCylZ>(c, bk, n, g, cr, bi);
      kind_loop<kKindCylRot>(c, bk, n, g, cr, bi);
      kind_loop<kKindSphere>(c, bk, n, g, cr, bi);
      kind_loop<kKindGround>(c, bk, n, g, cr, bi);
      if constexpr (kInlineCloud) cloud_inline(c, bk, n, g, nd.n2_hi, growf);
    }
  }

  // Every cloud against the spheres of a block, for the edge pass.  Super-
  // tiles, then tiles, are skipped when no live lane's block bounding sphere
  // (centre g, radius G, stretched by the rotation's norm bound) comes
  // within rc + band of their box; the points of a kept tile are compared
  // in tile-local coordinates, |c - p|^2 = |c'|^2 + (|p'|^2 - 2 c'.p'), from
  // the same operands as the matrix-core pass.  The decision is that of the
  // cloud pass (CloudSDF::is_outside, primitive_sdf.hpp:597-603, after the
  // AABB pre-test): collide iff min |c - p|^2 <= (r + rc)^2, with the band
  // around it, widened by the tile-local form's error bound, left to the
  // FP64 walk.
  __device__ __forceinline__ void cloud_inline(const V3<float> (&c)[kBlockSpheres], const DevSphereBlock& bk, int n,
                                               V3<float> g, float n2_hi, float growf) {
    const DevPrim<float>* vp = reinterpret_cast<const DevPrim<float>*>(sc.vprims);
    for (int p = sc.vrange[kKindCloud]; p < sc.vrange[kKindCloud + 1]; ++p) {
      if (!wave_any(alive())) return;
      const DevPrim<float> Pc = cload(vp, p);
      // every sphere of the block fails the reference's AABB pre-test when
      // the bounding centre is outside the AABB grown by the block's reach
      if (!wave_any(alive() && !outside_aabb(Pc, g, bk.G * growf + band + 1e-4f))) continue;
      const float rc = Pc.p0;
      const float reach = bk.G * n2_hi + rc + band + 1e-5f;
      float hi[kBlockSpheres], lo[kBlockSpheres];
      cloud_block_min<true>(sc, (Pc.flags >> 4) & 15, c, n, g, reach * reach, alive(), hi, lo);
#pragma unroll
      for (int k = 0; k < kBlockSpheres; ++k) {
        if (k < n) {
          const float r = bk.s[k][3];
          if (!alive() || outside_aabb(Pc, c[k], r)) continue;
          // the exact minimum lies in [lo, hi]: certain outside the band around R^2
          const float R = r + rc, R2 = R * R, w = band * (2.0f * R + band);
          if (hi[k] - R2 < -w) {
            definite = true;
          } else if (!(lo[k] - R2 > w)) {
            uncertain = true;
            n_inb = 2;
          }
        }
      }
    }
  }

  // A sphere pair can collide or fall in the band only within rs + band.
  static constexpr bool kPairCull = true;
  __device__ bool pair_near(V3<float> g1, V3<float> g2, float reach) const {
    const V3<float> d = sub(g1, g2);
    return !(sqrtf(dot(d, d)) - reach > band + 1e-4f);
  }
  __device__ void self(const DevSelfPair&, int, int, V3<float> ci, V3<float> cj, float rs) {
    if (!alive()) return;
    const V3<float> d = sub(ci, cj);
    const float m = dot(d, d) - rs * rs;
    if (fabsf(m) <= band * (2.0f * rs + band)) {
      uncertain = true;
      n_inb = 2;
    } else if (m < 0.0f) {
      definite = true;
    }
  }
};

using ValidVis32 = ValidVis32T<false>;

// The FP32 validity test of one evaluate unit (a sphere block, or a self
// link pair) of a configuration: FK stops at the unit's last node, blocks
// and self pairs outside the unit are skipped.  Clouds are not handled (the
// unit pass runs only for scenes without one).
struct ValidUnitVis : ValidVis32 {
  static constexpr bool kRanged = true;
  DevEvalUnit u{};
  __device__ int last_node() const { return u.last; }
  __device__ int pair_lo() const { return u.k_lo; }
  __device__ int pair_hi() const { return u.k_hi; }
  __device__ void blocks(const DevNode<float>& nd, const Pose<float>& P) {
    blocks_range(nd, P, max(nd.block_begin, u.b_lo), min(nd.block_end, u.b_hi));
  }
};

// FP64 re-decision with the reference's own tests: per sphere the AABB
// pre-test and is_outside (check_ext_collision), per self pair the strict
// squared-distance test (check_self_collision).  The group-sphere broad
// phase of the reference only culls, so it is not needed for the result.
// Spheres are visited by block with the FP32 pass's cull table: its grown
// boxes carry a 1e-4 m slack, far above the FP32/FP64 centre difference, so a
// candidate it drops is outside the reference's AABB pre-test in FP64 too.
// A wave may take a share of the work: candidates i with i % n_split ==
// split (self pairs with split 0).
struct ValidVis64 {
  static constexpr bool kSelf = true;
  static constexpr bool kBlocks = true;
  const KinArgs<double>& ka;
  const SceneArgs<double>& sc;
  int split = 0, n_split = 1;
  bool collided = false;
  __device__ bool alive() const { return !collided; }
  template <class QS>
  __device__ void begin(const Pose<double>&, const QS&) {}
  __device__ void node(int, const DevNode<double>&, const Pose<double>&) {}
  __device__ void sphere(int, V3<double>, double) {}

  template <int K>
  __device__ __forceinline__ void kind_loop(const V3<double> (&c)[kBlockSpheres], const double (&r)[kBlockSpheres],
                                            int n, V3<double> g, const CullRange& cr, bool nocull) {
    const DevPrim<double>* vp = reinterpret_cast<const DevPrim<double>*>(sc.vprims);
    const int lo = nocull ? sc.vrange[K] : cr.at[K], hi = nocull ? sc.vrange[K + 1] : cr.at[K + 1];
    for (int i = lo + split; i < hi; i += n_split) {
      int p = i;
      if (!nocull) {
        const CullEntry ce = cload(sc.cull, i);
        const bool out = g.x < ce.lb[0] || g.x > ce.ub[0] || g.y < ce.lb[1] || g.y > ce.ub[1] ||
                         g.z < ce.lb[2] || g.z > ce.ub[2];
        if (!wave_any(alive() && !out)) continue;
        p = ce.prim;
      }
      const DevPrim<double> P = cload(vp, p);
#pragma unroll
      for (int k = 0; k < kBlockSpheres; ++k)
        if (k < n && !outside_aabb(P, c[k], r[k]) && !is_outside_ref<double, K>(P, c[k], r[k])) collided = true;
    }
  }

  // CloudSDF::is_outside: sqrt(min |p - c|^2) - rc > r (primitive_sdf.hpp:597-603),
  // the minimum taken over the tiles that can reach within r + rc.
  __device__ void cloud(const DevPrim<double>& P, V3<double> c, double r) {
    if (outside_aabb(P, c, r)) return;
    const double R = r + P.p0, lim = R * R * (1.0 + 1e-12) + 1e-300;
    double best = DBL_MAX;
    for (int t = 0; t < P.cloud_count; ++t) {
      const CloudTile64 T = cload(sc.tiles64, P.cloud_begin + t);
      const double dx = fmax(fmax(T.lb[0] - c.x, c.x - T.ub[0]), 0.0);
      const double dy = fmax(fmax(T.lb[1] - c.y, c.y - T.ub[1]), 0.0);
      const double dz = fmax(fmax(T.lb[2] - c.z, c.z - T.ub[2]), 0.0);
      if (dx * dx + dy * dy + dz * dz > lim) continue;
      const double* pt = sc.pts64 + (int64_t)(P.cloud_begin + t) * kCloudTile * 4;
      for (int j = 0; j < kCloudTile; ++j) {
        const double ex = pt[4 * j] - c.x, ey = pt[4 * j + 1] - c.y, ez = pt[4 * j + 2] - c.z;
        const double d2 = ex * ex + ey * ey + ez * ez;
        if (d2 < best) best = d2;
      }
    }
    if (!(sqrt(best) - P.p0 > r)) collided = true;
  }

  __device__ void blocks(const DevNode<double>& nd, const Pose<double>& P) {
    if (nd.block_begin >= nd.block_end) return;
    const M3<double> R = qmat(P.q);
    // outside the norm bounds the cull table assumed: test every primitive
    const double n2 = P.q.x * P.q.x + P.q.y * P.q.y + P.q.z * P.q.z + P.q.w * P.q.w;
    const bool nocull = wave_any(alive() && !(n2 >= nd.n2_lo && n2 <= nd.n2_hi));
    const int n_solid = sc.vrange[kKindCloud];
    for (int bi = nd.block_begin; bi < nd.block_end; ++bi) {
      if (!wave_any(alive())) return;
      const DevSphereBlock bk = cload(ka.blocks, bi);
      const int n = bk.n;
      V3<double> c[kBlockSpheres];
      double r[kBlockSpheres];
      bool bad = false;
#pragma unroll
      for (int k = 0; k < kBlockSpheres; ++k) {
        c[k] = P.t;
        r[k] = 0.0;
        if (k < n) {
          const DevSphere<double> sp = cload(ka.spheres, bk.sph[k]);
          c[k] = add(mv(R, V3<double>{sp.x, sp.y, sp.z}), P.t);
          r[k] = sp.r;
          bad |= is_nan3(c[k]);
        }
      }
      if (bad && n_solid > 0) collided = true;  // NaN centres collide with every primitive
      const V3<double> g = add(mv(R, V3<double>{bk.g[0], bk.g[1], bk.g[2]}), P.t);
      const CullRange cr = cload(sc.cull_range, bi);
      kind_loop<kKindBoxAA>(c, r, n, g, cr, nocull);
      kind_loop<kKindBoxRot>(c, r, n, g, cr, nocull);
      kind_loop<kKindCylZ>(c, r, n, g, cr, nocull);
      kind_loop<kKindCylRot>(c, r, n, g, cr, nocull);
      kind_loop<kKindSphere>(c, r, n, g, cr, nocull);
      kind_loop<kKindGround>(c, r, n, g, cr, nocull);
      const DevPrim<double>* vp = reinterpret_cast<const DevPrim<double>*>(sc.vprims);
      for (int p = sc.vrange[kKindCloud] + split; p < sc.vrange[kKindCloud + 1]; p += n_split) {
        const DevPrim<double> Pc = cload(vp, p);
#pragma unroll
        for (int k = 0; k < kBlockSpheres; ++k)
          if (k < n && !collided) cloud(Pc, c[k], r[k]);
      }
    }
  }
  __device__ void self(const DevSelfPair&, int, int, V3<double> ci, V3<double> cj, double rs) {
    if (collided || split != 0) return;
    const V3<double> d = sub(ci, cj);
    if (dot(d, d) < rs * rs) collided = true;
  }
};

// FP64 re-decision of a state on the one candidate the FP32 pass found in
// the band: the spheres of sphere block `block` against kind-sorted
// primitive `prim`, both per lane, tested when the walk reaches the node
// that carries the block.  Every other pair of the state was decided by the
// FP32 pass outside the band, so this pair decides the state.
struct PairVis64 {
  static constexpr bool kSelf = false;
  static constexpr bool kBlocks = true;
  const KinArgs<double>& ka;
  const SceneArgs<double>& sc;
  int block = 0, prim = 0;
  bool idle = false;
  bool done = false;
  bool collided = false;
  __device__ bool alive() const { return !idle && !done; }
  template <class QS>
  __device__ void begin(const Pose<double>&, const QS&) {}
  __device__ void node(int, const DevNode<double>&, const Pose<double>&) {}
  __device__ void sphere(int, V3<double>, double) {}
  __device__ void blocks(const DevNode<double>& nd, const Pose<double>& P) {
    if (idle || done || block < nd.block_begin || block >= nd.block_end) return;
    done = true;
    const M3<double> R = qmat(P.q);
    const DevSphereBlock& bk = ka.blocks[block];  // per-lane index: vector loads
    const DevPrim<double>& pr = reinterpret_cast<const DevPrim<double>*>(sc.vprims)[prim];
    const int n = bk.n;
    for (int k = 0; k < n; ++k) {
      const DevSphere<double> sp = ka.spheres[bk.sph[k]];
      const V3<double> c = add(mv(R, V3<double>{sp.x, sp.y, sp.z}), P.t);
      if (!outside_aabb(pr, c, sp.r) && !is_outside_ref<double>(pr, c, sp.r)) collided = true;
    }
  }
  __device__ void self(const DevSelfPair&, int, int, V3<double>, V3<double>, double) {}
};

// The FP32 walk's band covers its rounding for joint values within these
// ranges (an angle's FP32 rounding is |q| 2^-24, so within 4 pi it stays
// below 1e-6 rad; a translation within 16 m, below 1e-6 m).  A state with
// a value outside them is decided by the FP64 walk.
constexpr double kFp32AngleRange = 4.0 * M_PI;
constexpr double kFp32LinearRange = 16.0;
__device__ __forceinline__ bool fp32_range_col(uint64_t rev_mask, int col, double v) {
  return fabs(v) <= (((rev_mask >> col) & 1ull) ? kFp32AngleRange : kFp32LinearRange);
}

// valid[b] of a configuration with a joint value outside the FP32 walk's
// ranges: the FP64 walk decides it, whatever the FP32 tests said.
constexpr uint8_t kValidFp64 = 3;
template <class QS>
__device__ __forceinline__ bool fp32_range(const KinArgs<float>& ka, const QS& qs) {
  bool in = true;
  for (int j = 0; j < ka.n_cols; ++j) in &= fp32_range_col(ka.rev_mask, j, qs.get(j));
  return in;
}

// valid[b]: 0 collision, 1 free, 2 pending FP64 re-decision, kValidFp64
// decided by the FP64 walk only.
__global__ void __launch_bounds__(256) valid_kernel(KinArgs<float> k32, SceneArgs<float> s32, const double* q,
                                                    int64_t ldq, int64_t B, float band, uint8_t* valid, CloudWork cw) {
  const int64_t b = gid();
  const bool active = b < B;
  BatchQ qs{q, ldq, active ? b : B - 1};
  const bool taylor = k32.base_type == PMP_BASE_FIXED;  // constraints/primitive.hpp:59-61
  ValidVis32 v{k32, s32, band};
  v.cw = cw;
  v.use_cw = s32.n_cloud > 0 && active;
  v.b = b;
  v.idle = !active;
  run_fk(k32, qs, taylor, v);
  if (active) valid[b] = !fp32_range(k32, qs) ? kValidFp64 : (v.definite ? 0 : (v.uncertain ? 2 : 1));
}

// Batches of up to kValidUnitsMaxBatch configurations: a lane per
// (configuration, unit), the unit being grid row y, so that a batch of a few
// thousand configurations still spreads over the GPU.  Each lane ORs its
// outcome into flags[b] (1: a certain collision, 2: a test in the band);
// valid_units_finish_kernel turns the flags into the answer valid_kernel
// gives (a certain collision anywhere decides, else any band test leaves
// the configuration to the FP64 pass) and clears them for the next launch.
// With clouds in the scene each lane writes the cloud queries of its unit's
// sphere blocks (the units' blocks partition the spheres).
__global__ void __launch_bounds__(64) valid_units_kernel(KinArgs<float> k32, SceneArgs<float> s32, const double* q,
                                                         int64_t ldq, int64_t B, float band, int32_t* flags,
                                                         CloudWork cw) {
  const int64_t b = gid();
  const bool active = b < B;
  BatchQ qs{q, ldq, active ? b : B - 1};
  ValidUnitVis v{{k32, s32, band}};
  v.u = cload(k32.units, (int)blockIdx.y);
  v.cw = cw;
  v.use_cw = s32.n_cloud > 0 && active;
  v.b = b;
  v.idle = !active;
  run_fk(k32, qs, k32.base_type == PMP_BASE_FIXED, v);
  // unit 0 also reports a joint value outside the FP32 walk's ranges (4)
  const int f = (v.definite ? 1 : 0) | (v.uncertain ? 2 : 0) | (blockIdx.y == 0 && !fp32_range(k32, qs) ? 4 : 0);
  if (active && f != 0) atomicOr(flags + b, f);
}

__global__ void __launch_bounds__(256) valid_units_finish_kernel(int32_t* flags, int64_t B, uint8_t* valid) {
  const int64_t b = gid();
  if (b >= B) return;
  const int f = flags[b];
  valid[b] = (f & 4) ? kValidFp64 : ((f & 1) ? 0 : ((f & 2) ? 2 : 1));
  flags[b] = 0;
}

// FP64 re-decision of the configurations left pending (2) by the FP32 pass
// and of those it could not decide (kValidFp64).
// With unit_flags (a scene without clouds, after the unit pass) it first
// turns the flags into valid[b] as valid_units_finish_kernel does, and
// clears them.
__global__ void __launch_bounds__(64) valid_recheck_kernel(KinArgs<double> k64, SceneArgs<double> s64,
                                                           const double* q, int64_t ldq, int64_t B, uint8_t* valid,
                                                           int32_t* n_recheck, int32_t* unit_flags) {
  const int64_t b = gid();
  if (b >= B) return;
  if (unit_flags) {
    const int f = unit_flags[b];
    unit_flags[b] = 0;
    if (!(f & 4) && (f & 3) != 2) {
      valid[b] = (f & 1) ? 0 : 1;
      return;
    }
  } else if (valid[b] < 2) {
    return;
  }
  BatchQ qs{q, ldq, b};
  ValidVis64 v{k64, s64};
  run_fk(k64, qs, k64.base_type == PMP_BASE_FIXED, v);
  valid[b] = v.collided ? 0 : 1;
  if (n_recheck) atomicAdd(n_recheck, 1);
}

// ------------------------------------------------------------------ edges
// Persistent kernel: each wave takes the next edge from a global counter,
// walks its waypoints 64 at a time in the validator's visiting order and
// stops at the first chunk holding an invalid waypoint.  The edge's answer
// does not depend on the visiting order (every waypoint must be valid); the
// farthest-first order only makes the first chunk the most informative.
__device__ __forceinline__ double edge_step_ratio(const EdgeArgs& ea, int D, int64_t e) {
  // Step ratio (ompl/motion_validator.cpp:89-126), FP64.
  if (ea.vtype == PMP_VALIDATOR_BOX) {
    double max_diff = -INFINITY, diff_longest = 0.0;
    int longest = 0;
    for (int i = 0; i < D; ++i) {
      const double diff = cload(ea.s2, (int64_t)i * ea.lds + e) - cload(ea.s1, (int64_t)i * ea.lds + e);
      const double asd = fabs(diff) * cload(ea.inv_width, i);
      if (asd > max_diff) {
        max_diff = asd;
        longest = i;
        diff_longest = diff;
      }
    }
    return fabs(diff_longest) < 1e-6 ? INFINITY : cload(ea.width, longest) / fabs(diff_longest);
  }
  double d2 = 0.0;
  for (int i = 0; i < D; ++i) {
    const double diff = cload(ea.s1, (int64_t)i * ea.lds + e) - cload(ea.s2, (int64_t)i * ea.lds + e);
    d2 += diff * diff;
  }
  return ea.resolution / sqrt(d2);
}

// Waypoint k (0-based) of an edge with n_test states: k = 0 is s2 itself,
// k >= 1 is interior position k + 1 of the farthest-first sequence when
// n_test <= 128 and position k of the linear sweep otherwise
// (ompl/motion_validator.cpp:44-78).
__device__ __forceinline__ EdgeQ edge_waypoint(const EdgeArgs& ea, int64_t e, int64_t n_test, double ratio,
                                               int64_t k) {
  double t = 0.0;
  if (k > 0) t = n_test <= 128 ? (double)ea.seq[(n_test - 1) * 128 + (k + 1)] * ratio : (double)k * ratio;
  return EdgeQ{ea.s1, ea.s2, ea.lds, e, t, k == 0};
}

// An edge's endpoints held across its wave: lane i < D holds column i of s1
// (a) and of s2 (b).  They are fetched once with vector loads, and each
// per-joint read is a lane read into a scalar register, which keeps the
// streamed edge data out of the scalar cache the cull and primitive tables
// live in.
struct EdgeEnds {
  double a, b;
};
__device__ __forceinline__ EdgeEnds load_edge_ends(const EdgeArgs& ea, int D, int64_t e, int lane) {
  EdgeEnds r{0.0, 0.0};
  if (lane < D) {
    r.a = ea.s1[(int64_t)lane * ea.lds + e];
    r.b = ea.s2[(int64_t)lane * ea.lds + e];
  }
  return r;
}
__device__ __forceinline__ double lane_f64(double v, int i) {
  const long long u = __builtin_bit_cast(long long, v);
  const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)u, i);
  const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(u >> 32), i);
  return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
}
struct EdgeQv {
  EdgeEnds ends;
  double t;
  bool at_end;
  __device__ __forceinline__ double get(int col) const {
    const double a = lane_f64(ends.a, col), b = lane_f64(ends.b, col);
    return at_end ? b : a + (b - a) * t;
  }
};
// edge_step_ratio and edge_waypoint over the wave-held endpoints: the same
// arithmetic on the same values.
__device__ __forceinline__ double edge_step_ratio(const EdgeArgs& ea, int D, const EdgeEnds& en) {
  if (ea.vtype == PMP_VALIDATOR_BOX) {
    double max_diff = -INFINITY, diff_longest = 0.0;
    int longest = 0;
    for (int i = 0; i < D; ++i) {
      const double diff = lane_f64(en.b, i) - lane_f64(en.a, i);
      const double asd = fabs(diff) * cload(ea.inv_width, i);
      if (asd > max_diff) {
        max_diff = asd;
        longest = i;
        diff_longest = diff;
      }
    }
    return fabs(diff_longest) < 1e-6 ? INFINITY : cload(ea.width, longest) / fabs(diff_longest);
  }
  double d2 = 0.0;
  for (int i = 0; i < D; ++i) {
    const double diff = lane_f64(en.a, i) - lane_f64(en.b, i);
    d2 += diff * diff;
  }
  return ea.resolution / sqrt(d2);
}
__device__ __forceinline__ EdgeQv edge_waypoint(const EdgeArgs& ea, const EdgeEnds& en, int64_t n_test, double ratio,
                                                int64_t k) {
  double t = 0.0;
  if (k > 0) t = n_test <= 128 ? (double)ea.seq[(n_test - 1) * 128 + (k + 1)] * ratio : (double)k * ratio;
  return EdgeQv{en, t, k == 0};
}

// valid[e]: 0 invalid, 1 valid, 2 pending FP64 re-decision.  n_checked[e]:
// the states the reference's checkMotion evaluates, i.e. the visiting-order
// index of the first invalid state plus one, or every state of a valid edge
// (0 for an edge whose step ratio is infinite, -1 for an edge that would need
// more than kMaxEdgeStates states).  A pending edge queues its in-band states
// that can matter (all of them, or those before the first certain collision)
// and that collision's index; the FP64 pass decides from them.
// Register budget: four waves per SIMD without clouds (the by-index grid
// of the autotuned schedule holds four per SIMD), three with them.
template <bool kCloud>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(kCloud ? 3 : 4))) edges_kernel(KinArgs<float> k32, SceneArgs<float> s32, KinArgs<double> k64,
                                                    SceneArgs<double> s64, EdgeArgs ea, float band) {
  const int lane = threadIdx.x & 63;
  const bool taylor = k32.base_type == PMP_BASE_FIXED;
  // A grid with a wave for every edge gives each wave its edge by index:
  // workgroups go to the XCDs in turn, so XCD x takes a contiguous range of
  // edges (neighbouring edges share their endpoints' cache lines in that
  // XCD's L2) and no wave waits on the work counter.  A smaller grid pulls
  // edges from the counter.  Every lane takes part in the atomic (lane 0
  // adds 1, the others 0) and the index is read from the first lane into a
  // scalar register, so it and every branch on it are wave-uniform.
  const int W = blockDim.x >> 6;
  const bool by_index = (int64_t)gridDim.x * W >= ea.E;
  int e;
  if (by_index) {
    const int g = blockIdx.x, x = g % kXcds;
    int base = 0;
    for (int xx = 0; xx < x; ++xx) base += ((int)gridDim.x - xx + kXcds - 1) / kXcds;
    e = (base + g / kXcds) * W + __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  } else {
    e = __builtin_amdgcn_readfirstlane(atomicAdd(ea.counter, lane == 0 ? 1 : 0));
  }
  for (; e < ea.E; e = by_index ? ea.E : __builtin_amdgcn_readfirstlane(atomicAdd(ea.counter, lane == 0 ? 1 : 0))) {
    const EdgeEnds ends = load_edge_ends(ea, k32.n_cols, e, lane);
    const double ratio = edge_step_ratio(ea, k32.n_cols, ends);
    // endpoints outside the FP32 walk's ranges (so are its waypoints: they
    // lie between them): the FP64 pass walks every state
    bool far = false;
    for (int i = 0; i < k32.n_cols; ++i)
      far |= !(fp32_range_col(k32.rev_mask, i, lane_f64(ends.a, i)) &&
               fp32_range_col(k32.rev_mask, i, lane_f64(ends.b, i)));
    int state = 1;
    int64_t checked = 0;
    int n_pend = 0;        // queued in-band states, -1 once more than kPendStates
    int kdef = INT_MAX;    // first certain collision
    if (ratio != INFINITY) {
      if (!(ratio > 0.0) || 1.0 / ratio > kMaxEdgeStates) {  // NaN states, or more states than an edge may hold
        state = 0;
        checked = -1;
      } else if (far) {
        state = 2;
        n_pend = -1;
      } else {
        const int64_t n_test = (int64_t)floor(1.0 / ratio) + 2;
        const int64_t n_wp = n_test - 1;  // s2 plus the interior waypoints
        bool unc = false;                 // an in-band state that matters seen so far
        int qk = 0, qb = -1, qp = 0;      // queue entry `lane` (lanes < kPendStates) in registers
        for (int64_t base = 0; base < n_wp; base += 64) {
          const int64_t k = base + lane;
          const bool has = k < n_wp;
          ValidVis32T<kCloud, true> v{k32, s32, band};
          v.idle = !has;
          v.key = (int)k;
          run_fk(k32, edge_waypoint(ea, ends, n_test, ratio, has ? k : 0), taylor, v);
          const bool def = has && v.definite;
          if (__any(def)) kdef = wave_min(def ? (int)k : INT_MAX);
          const bool in_band = has && v.uncertain && !v.definite && (int)k < kdef;
          const uint64_t m = __ballot(in_band);
          if (m) {
            unc = true;
            const int cnt = __popcll(m);
            if (n_pend >= 0 && n_pend + cnt <= kPendStates) {
              const int pb = v.n_inb == 1 ? v.inb_block : -1;
              if (in_band) {
                const int64_t slot = (int64_t)e * kPendStates + n_pend + __popcll(m & ((1ull << lane) - 1));
                ea.pend[slot] = (int)k;
                ea.pend_pair[2 * slot] = pb;
                ea.pend_pair[2 * slot + 1] = v.inb_prim;
              }
              // lane n_pend + r takes the entry of the in-band lane of rank r
              const int r = lane - n_pend;
              int src = 0;
              if (r >= 0 && r < cnt) {
                uint64_t mm = m;
                for (int i = 0; i < r; ++i) mm &= mm - 1;
                src = __builtin_ctzll(mm);
              }
              const int sk = __shfl((int)k, src, 64), sb = __shfl(pb, src, 64), sp = __shfl(v.inb_prim, src, 64);
              if (r >= 0 && r < cnt) {
                qk = sk;
                qb = sb;
                qp = sp;
              }
              n_pend += cnt;
            } else {
              n_pend = -1;
            }
          }
          if (kdef != INT_MAX) {
            state = unc ? 2 : 0;
            checked = kdef + 1;
            break;
          }
          checked = base + 64 < n_wp ? base + 64 : n_wp;
        }
        if (state == 1 && unc) state = 2;
        // A pending edge whose queued states each have one in-band candidate
        // is re-decided here, in FP64 on those candidates as the FP64 pass's
        // one-candidate path does, and never reaches that pass.
        if (state == 2 && n_pend > 0) {
          const bool has = lane < n_pend;
          const int k = has ? qk : 0;
          const int pb = has ? qb : 0, pp = has ? qp : 0;
          if (!__any(has && pb < 0)) {
            PairVis64 v{k64, s64, pb, pp};
            v.idle = !has;
            run_fk(k64, edge_waypoint(ea, ends, n_test, ratio, k), k64.base_type == PMP_BASE_FIXED, v);
            const bool hit = has && v.collided;
            const int first = min(kdef, __any(hit) ? wave_min(hit ? k : INT_MAX) : INT_MAX);
            state = first == INT_MAX ? 1 : 0;
            checked = first == INT_MAX ? n_wp : (int64_t)first + 1;
            if (lane == 0) atomicAdd(ea.counter + kCntInline, 1);
          }
        }
      }
    }
    if (lane == 0) {
      ea.valid[e] = (uint8_t)state;
      if (ea.n_checked) ea.n_checked[e] = (int32_t)checked;
      if (state == 2) {
        ea.pend_info[2 * (int64_t)e] = n_pend;
        ea.pend_info[2 * (int64_t)e + 1] = kdef;
        ea.plist[atomicAdd(ea.counter + kCntPending, 1)] = e;
      }
    }
  }
}

// FP64 decision of the edges left pending by edges_kernel, read from its
// pending list, one workgroup of kRecheckWaves waves per edge and the list
// shared out over the grid.  With its in-band states queued, the first wave
// re-decides just those (a lane each); the first colliding state is then the
// smaller of the FP64 hits and the FP32 pass's first certain collision.  An edge that queued too many
// walks its states again: a round evaluates up to kRecheckWaves chunks of 64
// states at once, each chunk's candidates shared among kRecheckWaves / chunks
// waves, and a round that finds a collision ends the walk.  Every branch
// around a barrier depends only on the list entry, so it is uniform over the
// workgroup.
constexpr int kRecheckWaves = 4;
constexpr int kRecheckGrid = 32;  // workgroups; few, so the idle launch and its end-of-launch atomic stay cheap
__global__ void __launch_bounds__(64 * kRecheckWaves) edges_recheck_kernel(KinArgs<double> k64,
                                                                           SceneArgs<double> s64, EdgeArgs ea) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int n_list = ea.counter[kCntPending];
  __shared__ int first;
  for (int i = blockIdx.x; i < n_list; i += gridDim.x) {
    const int64_t e = ea.plist[i];
    const double ratio = edge_step_ratio(ea, k64.n_cols, e);
    const int64_t n_test = (int64_t)floor(1.0 / ratio) + 2;
    const int64_t n_wp = n_test - 1;
    const int n_pend = ea.pend_info[2 * e];
    if (threadIdx.x == 0) first = n_pend >= 0 ? ea.pend_info[2 * e + 1] : INT_MAX;
    __syncthreads();
    if (n_pend >= 0) {
      // One wave: sharing out so few states' candidates among the waves
      // measured slower than one wave walking them all.
      if (w == 0) {
        const bool has = lane < n_pend;
        const int64_t slot = e * kPendStates + lane;
        const int64_t k = has ? ea.pend[slot] : 0;
        const int pb = has ? ea.pend_pair[2 * slot] : 0, pp = has ? ea.pend_pair[2 * slot + 1] : 0;
        bool hit;
        if (__any(pb < 0)) {  // a state whose uncertainty is not one candidate: full FP64 walk
          ValidVis64 v{k64, s64};
          v.collided = !has;
          run_fk(k64, edge_waypoint(ea, e, n_test, ratio, k), k64.base_type == PMP_BASE_FIXED, v);
          hit = has && v.collided;
        } else {
          PairVis64 v{k64, s64, pb, pp};
          v.idle = !has;
          run_fk(k64, edge_waypoint(ea, e, n_test, ratio, k), k64.base_type == PMP_BASE_FIXED, v);
          hit = has && v.collided;
        }
        if (__any(hit)) {
          const int kmin = wave_min(hit ? (int)k : INT_MAX);
          if (lane == 0) atomicMin(&first, kmin);
        }
      }
      __syncthreads();
    } else {
      const int64_t n_chunks = (n_wp + 63) / 64;
      const int nw = (int)(blockDim.x >> 6);  // kRecheckWaves, fewer when the kept poses need the LDS
      const int per_round = n_chunks < nw ? (int)n_chunks : nw;
      const int n_split = nw / per_round;
      for (int64_t c0 = 0; c0 < n_chunks; c0 += per_round) {
        const int64_t chunk = c0 + w / n_split;
        if (w / n_split < per_round && chunk < n_chunks) {
          const int64_t k = chunk * 64 + lane;
          const bool has = k < n_wp;
          ValidVis64 v{k64, s64};
          v.split = w % n_split;
          v.n_split = n_split;
          v.collided = !has;
          run_fk(k64, edge_waypoint(ea, e, n_test, ratio, has ? k : 0), k64.base_type == PMP_BASE_FIXED, v);
          const bool hit = has && v.collided;
          if (__any(hit)) {
            const int kmin = wave_min(hit ? (int)k : INT_MAX);
            if (lane == 0) atomicMin(&first, kmin);
          }
        }
        __syncthreads();
        const int f = first;
        __syncthreads();
        if (f != INT_MAX) break;  // uniform: every wave read the same value between the barriers
      }
    }
    if (threadIdx.x == 0) {
      ea.valid[e] = first == INT_MAX ? 1 : 0;
      if (ea.n_checked) ea.n_checked[e] = (int32_t)(first == INT_MAX ? n_wp : (int64_t)first + 1);
    }
    __syncthreads();  // thread 0 has read `first` before the next entry resets it
  }
  // The last workgroup to finish resets the counters for the next launch.
  __syncthreads();
  if (threadIdx.x == 0) {
    __threadfence();
    if (atomicAdd(ea.counter + kCntDone, 1) == (int)gridDim.x - 1) {
      ea.counter[kCntRecheck] = n_list + ea.counter[kCntInline];
      ea.counter[kCntInline] = 0;
      ea.counter[kCntPending] = 0;
      ea.counter[kCntWork] = 0;
      ea.counter[kCntDone] = 0;
    }
  }
}

// ------------------------------------------------- Jacobian-carrying visitors
// Joint frames (world axis, origin) per q column, filled while walking the
// chain; base columns are set from the base pose.  Column j moves a point c
// on a node whose col_mask has bit j by
//   revolute: a_j x (c - o_j)      prismatic: a_j
// (KinematicModel::get_attached_point_jacobian, kinematics/algorithm.cpp:346-433;
// the base columns are the analytic form of its finite differences).
// The base's own q columns (a planar or floating base), as (column, axis,
// origin) frames: translations along x, y, z, then the roll / pitch / yaw
// axes of the base's Rz(yaw) Ry(pitch) Rx(roll) rotation.
template <class QS, class F>
__device__ __forceinline__ void base_columns(const KinArgs<float>& ka, const Pose<float>& base, const QS& qs, F&& f) {
  const int n = ka.n_q_cols;
  if (ka.base_type == PMP_BASE_PLANAR) {
    f(n, V3<float>{1.f, 0.f, 0.f}, base.t);
    f(n + 1, V3<float>{0.f, 1.f, 0.f}, base.t);
    f(n + 2, V3<float>{0.f, 0.f, 1.f}, base.t);
  } else if (ka.base_type == PMP_BASE_FLOATING) {
    const float pitch = (float)qs.get(n + 4), yaw = (float)qs.get(n + 5);
    float sp, cp, sy, cy;
    sincosf(pitch, &sp, &cp);
    sincosf(yaw, &sy, &cy);
    f(n, V3<float>{1.f, 0.f, 0.f}, base.t);
    f(n + 1, V3<float>{0.f, 1.f, 0.f}, base.t);
    f(n + 2, V3<float>{0.f, 0.f, 1.f}, base.t);
    f(n + 3, V3<float>{cy * cp, sy * cp, -sp}, base.t);  // Rz(yaw) Ry(pitch) x
    f(n + 4, V3<float>{-sy, cy, 0.f}, base.t);           // Rz(yaw) y
    f(n + 5, V3<float>{0.f, 0.f, 1.f}, base.t);
  }
}

template <int MAXC>
struct Frames {
  float ax[MAXC], ay[MAXC], az[MAXC], ox[MAXC], oy[MAXC], oz[MAXC];
  __device__ __forceinline__ void set(int col, V3<float> a, V3<float> o) {
#pragma unroll
    for (int j = 0; j < MAXC; ++j)
      if (j == col) {
        ax[j] = a.x;
        ay[j] = a.y;
        az[j] = a.z;
        ox[j] = o.x;
        oy[j] = o.y;
        oz[j] = o.z;
      }
  }
  template <class QS>
  __device__ __forceinline__ void set_base(const KinArgs<float>& ka, const Pose<float>& base, const QS& qs) {
    base_columns(ka, base, qs, [&](int col, V3<float> a, V3<float> o) { set(col, a, o); });
  }
  __device__ __forceinline__ V3<float> col(int j, uint64_t mask, uint64_t rev, V3<float> c) const {
    V3<float> r{0.f, 0.f, 0.f};
#pragma unroll
    for (int k = 0; k < MAXC; ++k)
      if (k == j && ((mask >> k) & 1ull)) {
        const V3<float> a{ax[k], ay[k], az[k]};
        r = ((rev >> k) & 1ull) ? cross(a, V3<float>{c.x - ox[k], c.y - oy[k], c.z - oz[k]}) : a;
      }
    return r;
  }
};

// ------------------------------------------------------------ evaluate rows
// evaluate_ext_collision / evaluate_self_collision
// (constraints/primitive_sphere_collision.cpp:256-383) split into units: a
// unit is a sphere item (the external minimum over its spheres x every
// primitive, visited in the reference's (group, primitive, sphere) key order)
// or a self group pair (minimum over its sphere pairs in visiting order).  A
// wave evaluates one unit for 64 consecutive configurations and leaves
// (value, winner) per lane; the last wave to finish a chunk (device-scope
// counter after a release fence) reduces the units in visiting order and
// recomputes the winning Jacobian rows with one more FK walk.
template <int MAXC>
struct EvalUnitVis {
  static constexpr bool kSelf = true;
  static constexpr bool kRanged = true;
  static constexpr bool kBlocks = true;
  const KinArgs<float>& ka;
  const SceneArgs<float>& sc;
  DevEvalUnit u{};
  float best = kCutoff, sbest = kCutoff;
  int bkey = 0x7fffffff, bidx = -1, sidx = -1;
  V3<float> bc{0.f, 0.f, 0.f}, sc1{0.f, 0.f, 0.f}, sc2{0.f, 0.f, 0.f};  // the winners' world centres
  float ereach = 0.0f;  // the current block's bounding radius, grown, plus a slack
  __device__ bool alive() const { return true; }
  __device__ int last_node() const { return u.last; }
  __device__ int pair_lo() const { return u.k_lo; }
  __device__ int pair_hi() const { return u.k_hi; }
  template <class QS>
  __device__ void begin(const Pose<float>&, const QS&) {}
  __device__ void node(int, const DevNode<float>&, const Pose<float>&) {}
  __device__ void sphere(int, V3<float>, float) {}

  // Candidates of one kind (all primitives of that kind when `all`): the
  // reference skips a pair whose sphere, grown by the cutoff, is outside the
  // primitive's AABB; the evaluate cull table grows the boxes by the block's
  // radius on top, so a candidate the wave skips is one every lane skips.
  template <int K>
  __device__ __forceinline__ void kind_loop(const V3<float> (&c)[kBlockSpheres], const DevSphereBlock& bk, int n,
                                            const DevSphereMeta (&m)[kBlockSpheres], V3<float> g,
                                            const CullRange& cr, bool all) {
    const DevPrim<float>* vp = reinterpret_cast<const DevPrim<float>*>(sc.vprims);
    const int lo = all ? sc.vrange[K] : cr.at[K], hi = all ? sc.vrange[K + 1] : cr.at[K + 1];
    for (int i = lo; i < hi; ++i) {
      int p, orig;
      if (all) {
        p = i;
        orig = cload(sc.vorig, i);
      } else {
        const CullEntry ce = cload(sc.ecull, i);
        const bool out = g.x < ce.lb[0] || g.x > ce.ub[0] || g.y < ce.lb[1] || g.y > ce.ub[1] ||
                         g.z < ce.lb[2] || g.z > ce.ub[2];
        if (!wave_any(!out)) continue;
        p = ce.prim;
        orig = ce.orig;
      }
      const DevPrim<float> P = cload(vp, p);
      // every member sphere's value is at least the bounding centre's
      // distance minus the block's reach (the SDFs are 1-Lipschitz): a wave
      // whose lanes all have their best below that bound skips the candidate
      if (!all && !wave_any(!(sdf_eval_kind<K>(P, g) - ereach > best))) continue;
#pragma unroll
      for (int k = 0; k < kBlockSpheres; ++k) {
        if (k < n) {
          const float r = bk.s[k][3];
          if (outside_aabb(P, c[k], r + kCutoff)) continue;
          const float v = sdf_eval_kind<K>(P, c[k]) - r;
          if (v <= best) {  // a new best, or a tie that the visiting order settles
            const int key = (m[k].group * sc.n_prims + orig) * ka.n_spheres + m[k].out;
            if (v < best || (bidx >= 0 && key < bkey)) {
              best = v;
              bkey = key;
              bidx = bk.sph[k] * kMaxPrims + orig;
              bc = c[k];
            }
          }
        }
      }
    }
  }

  __device__ void blocks(const DevNode<float>& nd, const Pose<float>& P) {
    const int lo = max(nd.block_begin, u.b_lo), hi = min(nd.block_end, u.b_hi);
    if (lo >= hi || sc.n_prims == 0) return;
    const M3<float> R = qmat(P.q);
    // outside the norm bounds the cull table assumed: every primitive is a candidate
    const float n2 = P.q.x * P.q.x + P.q.y * P.q.y + P.q.z * P.q.z + P.q.w * P.q.w;
    const bool all = wave_any(!(n2 >= nd.n2_lo && n2 <= nd.n2_hi));
    const float growf = 1.0f + 2.0f * fmaxf(0.0f, fmaxf(nd.n2_hi - 1.0f, 1.0f - nd.n2_lo)) + 1e-6f;
    for (int bi = lo; bi < hi; ++bi) {
      const DevSphereBlock bk = cload(ka.blocks, bi);
      ereach = bk.G * growf + 1e-4f;
      const int n = bk.n;
      V3<float> c[kBlockSpheres];
      DevSphereMeta m[kBlockSpheres];
#pragma unroll
      for (int k = 0; k < kBlockSpheres; ++k) {
        c[k] = P.t;
        m[k] = DevSphereMeta{0, 0, 0, 0};
        if (k < n) {
          c[k] = add(mv(R, V3<float>{bk.s[k][0], bk.s[k][1], bk.s[k][2]}), P.t);
          m[k] = cload(ka.meta, bk.sph[k]);
        }
      }
      const V3<float> g = add(mv(R, V3<float>{bk.g[0], bk.g[1], bk.g[2]}), P.t);
      const CullRange cr = cload(sc.ecull_range, bi);
      kind_loop<kKindBoxAA>(c, bk, n, m, g, cr, all);
      kind_loop<kKindBoxRot>(c, bk, n, m, g, cr, all);
      kind_loop<kKindCylZ>(c, bk, n, m, g, cr, all);
      kind_loop<kKindCylRot>(c, bk, n, m, g, cr, all);
      kind_loop<kKindSphere>(c, bk, n, m, g, cr, all);
      kind_loop<kKindGround>(c, bk, n, m, g, cr, all);
      if (sc.n_cloud > 0) cloud_loop(c, bk, n, m, g, nd.n2_hi, all);
    }
  }

  // Clouds: sqrt(min |c - p|^2) - rc - r (CloudSDF::evaluate,
  // primitive_sdf.hpp:585-590, minus the sphere radius) after the reference's
  // AABB prefilter grown by the cutoff.  Only points within r + rc + cutoff of
  // a sphere can give a value below the cutoff, so the minimum is taken over
  // the tiles within that reach of the block (all tiles when the rotation is
  // outside the norm bounds).
  __device__ __forceinline__ void cloud_loop(const V3<float> (&c)[kBlockSpheres], const DevSphereBlock& bk, int n,
                                             const DevSphereMeta (&m)[kBlockSpheres], V3<float> g, float n2_hi,
                                             bool all) {
    const DevPrim<float>* vp = reinterpret_cast<const DevPrim<float>*>(sc.vprims);
    for (int p = sc.vrange[kKindCloud]; p < sc.vrange[kKindCloud + 1]; ++p) {
      const DevPrim<float> Pc = cload(vp, p);
      const int orig = cload(sc.vorig, p);
      const float rc = Pc.p0;
      const float reach = bk.G * n2_hi + rc + kCutoff + 1e-5f;
      float d2[kBlockSpheres], unused[kBlockSpheres];
      cloud_block_min<false>(sc, (Pc.flags >> 4) & 15, c, n, g, all ? FLT_MAX : reach * reach, true, d2, unused);
#pragma unroll
      for (int k = 0; k < kBlockSpheres; ++k) {
        if (k < n) {
          const float r = bk.s[k][3];
          if (d2[k] == FLT_MAX || outside_aabb(Pc, c[k], r + kCutoff)) continue;
          const float v = sqrtf(d2[k]) - rc - r;
          const int key = (m[k].group * sc.n_prims + orig) * ka.n_spheres + m[k].out;
          if (v < best || (bidx >= 0 && v == best && key < bkey)) {
            best = v;
            bkey = key;
            bidx = bk.sph[k] * kMaxPrims + orig;
            bc = c[k];
          }
        }
      }
    }
  }

  // A sphere pair's value is at least the groups' distance minus their
  // reach; only a value below the lane's best self value changes it.
  static constexpr bool kPairCull = true;
  __device__ bool pair_near(V3<float> g1, V3<float> g2, float reach) const {
    const V3<float> d = sub(g1, g2);
    return !(sqrtf(dot(d, d)) - reach > sbest + 1e-5f);
  }
  __device__ void self(const DevSelfPair&, int i, int j, V3<float> ci, V3<float> cj, float rs) {
    const V3<float> d = sub(ci, cj);
    const float dist = sqrtf(dot(d, d)) - rs;
    if (dist < sbest) {
      sbest = dist;
      sidx = i * kMaxSpheres + j;
      sc1 = ci;
      sc2 = cj;
    }
  }
};

// FK walk for the Jacobian rows of the winners: it keeps the joint frames
// and computes the world centres of just the (up to three) winning spheres,
// and stops at the last node any lane of the wave needs.
template <int MAXC>
struct EvalRowVis {
  static constexpr bool kSelf = false;
  static constexpr bool kRanged = true;
  static constexpr bool kBlocks = true;
  const KinArgs<float>& ka;
  Frames<MAXC> fr;
  int s0 = -1, s1 = -1, s2 = -1;
  int last = 0;  // wave-uniform
  V3<float> c0{0.f, 0.f, 0.f}, c1{0.f, 0.f, 0.f}, c2{0.f, 0.f, 0.f};
  __device__ bool alive() const { return true; }
  __device__ int last_node() const { return last; }
  __device__ int pair_lo() const { return 0; }
  __device__ int pair_hi() const { return 0; }
  template <class QS>
  __device__ void begin(const Pose<float>& base, const QS& qs) { fr.set_base(ka, base, qs); }
  __device__ void node(int, const DevNode<float>& nd, const Pose<float>& P) {
    if (nd.q_col < 0) return;
    fr.set(nd.q_col, qrot(P.q, V3<float>{nd.axis[0], nd.axis[1], nd.axis[2]}), P.t);
  }
  __device__ void sphere(int, V3<float>, float) {}
  __device__ void blocks(const DevNode<float>& nd, const Pose<float>& P) {
    const bool h0 = s0 >= nd.sphere_begin && s0 < nd.sphere_end;
    const bool h1 = s1 >= nd.sphere_begin && s1 < nd.sphere_end;
    const bool h2 = s2 >= nd.sphere_begin && s2 < nd.sphere_end;
    if (!wave_any(h0 || h1 || h2)) return;
    const M3<float> R = qmat(P.q);
    auto centre = [&](int s) {
      const DevSphere<float> sp = ka.spheres[s];  // per-lane index: a vector load
      return add(mv(R, V3<float>{sp.x, sp.y, sp.z}), P.t);
    };
    if (h0) c0 = centre(s0);
    if (h1) c1 = centre(s1);
    if (h2) c2 = centre(s2);
  }
  __device__ void self(const DevSelfPair&, int, int, V3<float>, V3<float>, float) {}
};

// FK walk that builds the Jacobian rows of the winners column by column as
// the joint frames appear, from the winners' centres and directions known
// before the walk (ext: the primitive's gradient g0 at c0; self: the unit
// normal n from c2 to c1): no frame is kept, column j of a row is
// d . (a_j x (c - o_j)) for a rotating column and d . a_j otherwise, when j
// moves the winner's sphere (its node's column mask).
template <int MAXC>
struct EvalRowAccVis {
  static constexpr bool kSelf = false;
  static constexpr bool kRanged = true;
  static constexpr bool kBlocks = true;
  const KinArgs<float>& ka;
  int last = 0;  // wave-uniform
  V3<float> c0{0.f, 0.f, 0.f}, c1{0.f, 0.f, 0.f}, c2{0.f, 0.f, 0.f}, g0{0.f, 0.f, 0.f}, n{0.f, 0.f, 0.f};
  uint64_t m0 = 0, m1 = 0, m2 = 0;  // column masks of the winners' nodes (0: no winner)
  float re[MAXC], rs[MAXC];
  __device__ bool alive() const { return true; }
  __device__ int last_node() const { return last; }
  __device__ int pair_lo() const { return 0; }
  __device__ int pair_hi() const { return 0; }
  __device__ __forceinline__ void column(int col, V3<float> a, V3<float> o) {
    const bool rev = (ka.rev_mask >> col) & 1ull;
    auto cv = [&](V3<float> c) { return rev ? cross(a, sub(c, o)) : a; };
    const float ve = ((m0 >> col) & 1ull) ? dot(g0, cv(c0)) : 0.f;
    const float v1 = ((m1 >> col) & 1ull) ? dot(n, cv(c1)) : 0.f;
    const float v2 = ((m2 >> col) & 1ull) ? dot(n, cv(c2)) : 0.f;
#pragma unroll
    for (int j = 0; j < MAXC; ++j)
      if (j == col) {
        re[j] = ve;
        rs[j] = v1 - v2;
      }
  }
  template <class QS>
  __device__ void begin(const Pose<float>& base, const QS& qs) {
#pragma unroll
    for (int j = 0; j < MAXC; ++j) re[j] = rs[j] = 0.f;
    base_columns(ka, base, qs, [&](int col, V3<float> a, V3<float> o) { column(col, a, o); });
  }
  __device__ void node(int, const DevNode<float>& nd, const Pose<float>& P) {
    if (nd.q_col < 0) return;
    column(nd.q_col, qrot(P.q, V3<float>{nd.axis[0], nd.axis[1], nd.axis[2]}), P.t);
  }
  __device__ void sphere(int, V3<float>, float) {}
  __device__ void blocks(const DevNode<float>&, const Pose<float>&) {}
  __device__ void self(const DevSelfPair&, int, int, V3<float>, V3<float>, float) {}
};

struct EvalOut {
  float* vals;  // [cst_dim][ld]
  float* jac;   // [cst_dim][D][ld]
  int64_t ld;
  const DevEvalUnit* units;
  int n_units;
  // trajectories (offsets == nullptr: plain evaluate)
  const int64_t* offsets;
  int n_traj;
  const double* w2;
  float* cost;
  float* grad;
};

// Smoothness gradient of waypoint n and, from the trajectory's first lane,
// its cost (smoothcost_fullmat, nlp_solver/sqp_based_solver.py:74-96):
//   cost_i = sum_{interior t} sum_j w_j^2 a_{t,j}^2,  a_t = x_{t-1} - 2 x_t + x_{t+1},
//   d cost / d x_t = 2 w^2 (a_{t-1} - 2 a_t + a_{t+1})   (terms of interior indices only).
// Second differences are formed in FP64 (they cancel), terms accumulate in
// FP32 in a fixed order.
template <int MAXC>
__device__ void smooth_terms(const double* X, int64_t ldx, int64_t n, int D, const EvalOut& o) {
  int lo = 0, hi = o.n_traj;  // offsets[lo] <= n < offsets[hi]
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (o.offsets[mid] <= n) lo = mid;
    else hi = mid;
  }
  const int64_t b0 = o.offsets[lo], T = o.offsets[lo + 1] - b0, t = n - b0;
  auto x = [&](int64_t s, int j) { return X[(int64_t)j * ldx + b0 + s]; };
  auto acc = [&](int64_t s, int j) { return x(s - 1, j) - 2.0 * x(s, j) + x(s + 1, j); };
#pragma unroll
  for (int j = 0; j < MAXC; ++j) {
    if (j < D) {
      double g = 0.0;
      if (t - 1 >= 1 && t - 1 <= T - 2) g += acc(t - 1, j);
      if (t >= 1 && t <= T - 2) g -= 2.0 * acc(t, j);
      if (t + 1 >= 1 && t + 1 <= T - 2) g += acc(t + 1, j);
      o.grad[(int64_t)j * o.ld + n] = (float)(2.0 * o.w2[j] * g);
    }
  }
  if (t == 0) {
    float c = 0.f;
    for (int64_t s = 1; s + 1 < T; ++s)
      for (int j = 0; j < D; ++j) {
        const double a = acc(s, j);
        c += (float)(o.w2[j] * a * a);
      }
    o.cost[lo] = c;
  }
}

// Reduction of a configuration's unit results in the reference's order
// (external minimum by (group, primitive, sphere) key, self minimum by
// visiting order; units hold increasing self-pair ranges, so unit order is
// visiting order), then the Jacobian rows of the winners from one more FK
// walk, and for trajectories the smoothness terms.  Unit k's results are at
// xv[k * stride], xi, sv, si (any address space); (best, bidx, sbest, sidx)
// start the reduction.
template <int MAXC, class QS>
__device__ __forceinline__ void eval_finish(const KinArgs<float>& ka, const SceneArgs<float>& sc, const QS& qs, const double* q,
                            int64_t ldq, int64_t b, const EvalOut& o, float best, int bidx, float sbest, int sidx,
                            const float* xv, const int* xi, const float* sv, const int* si, int64_t stride,
                            int n_red) {
  const int D = ka.n_cols, NP = sc.n_prims;
  int bkey = 0x7fffffff;
  for (int k = 0; k < n_red; ++k) {
    const float val = xv[k * stride];
    const int idx = xi[k * stride];
    if (idx >= 0) {
      const DevSphereMeta m = cload(ka.meta, idx / kMaxPrims);
      const int key = (m.group * NP + idx % kMaxPrims) * ka.n_spheres + m.out;
      if (val < best || (bidx >= 0 && val == best && key < bkey)) {
        best = val;
        bkey = key;
        bidx = idx;
      }
    }
    const float sval = sv[k * stride];
    const int sid = si[k * stride];
    if (sid >= 0 && sval < sbest) {
      sbest = sval;
      sidx = sid;
    }
  }
  EvalRowVis<MAXC> rv{ka};
  int need = -1;  // the last node this lane's winners need
  if (bidx >= 0) {
    rv.s0 = bidx / kMaxPrims;
    need = cload(ka.meta, rv.s0).node;
  }
  if (sidx >= 0) {
    rv.s1 = sidx / kMaxSpheres;
    rv.s2 = sidx % kMaxSpheres;
    need = max(need, max(cload(ka.meta, rv.s1).node, cload(ka.meta, rv.s2).node));
  }
  // lanes of the wave that reach this point walk together (inactive lanes
  // have returned); no winner in the wave: the rows are zero without FK
  const int last = wave_max_i(need);
  if (last >= 0) {
    rv.last = last;
    run_fk(ka, qs, false, rv);
  }
  int row = 0;
  if (NP > 0) {
    o.vals[b] = best;
    V3<float> g{0.f, 0.f, 0.f};
    uint64_t mask = 0;
    if (bidx >= 0) {
      const DevPrim<float> P = prim_at(sc, bidx % kMaxPrims);
      if (P.type == PMP_PRIM_CLOUD) {
        // the gradient of sqrt(min |c - p|^2) at the nearest point, which the
        // reference's forward difference (eps 1e-6) approximates
        float d2;
        const V3<float> d = sub(rv.c0, cloud_nearest(sc, (P.flags >> 4) & 15, rv.c0, d2));
        g = scale(1.0f / sqrtf(dot(d, d)), d);
      } else {
        g = sdf_grad(P, rv.c0);
      }
      mask = cload(ka.nodes, cload(ka.meta, rv.s0).node).col_mask;
    }
#pragma unroll
    for (int j = 0; j < MAXC; ++j)
      if (j < D) o.jac[(int64_t)j * o.ld + b] = dot(g, rv.fr.col(j, mask, ka.rev_mask, rv.c0));
    row = 1;
  }
  if (ka.n_pairs > 0) {
    o.vals[(int64_t)row * o.ld + b] = sbest;
    V3<float> n{0.f, 0.f, 0.f};
    uint64_t m1 = 0, m2 = 0;
    if (sidx >= 0) {
      const V3<float> d = sub(rv.c1, rv.c2);
      n = scale(1.0f / sqrtf(dot(d, d)), d);
      m1 = cload(ka.nodes, cload(ka.meta, rv.s1).node).col_mask;
      m2 = cload(ka.nodes, cload(ka.meta, rv.s2).node).col_mask;
    }
#pragma unroll
    for (int j = 0; j < MAXC; ++j)
      if (j < D)
        o.jac[((int64_t)row * D + j) * o.ld + b] =
            dot(n, sub(rv.fr.col(j, m1, ka.rev_mask, rv.c1), rv.fr.col(j, m2, ka.rev_mask, rv.c2)));
  }
  if (o.offsets) smooth_terms<MAXC>(q, ldq, b, D, o);
}

// The rows of one walk's winners (whole mode, where the unit walk kept the
// winners' centres): the directions are formed first, then a walk that keeps
// no frames builds the rows column by column (EvalRowAccVis).
template <int MAXC, class QS>
__device__ __forceinline__ void eval_finish_whole(const KinArgs<float>& ka, const SceneArgs<float>& sc, const QS& qs,
                                                  const double* q, int64_t ldq, int64_t b, const EvalOut& o,
                                                  const EvalUnitVis<MAXC>& v) {
  const int D = ka.n_cols, NP = sc.n_prims;
  EvalRowAccVis<MAXC> rv{ka};
#pragma unroll
  for (int j = 0; j < MAXC; ++j) rv.re[j] = rv.rs[j] = 0.f;
  int need = -1;  // the last node this lane's winners need
  if (v.bidx >= 0) {
    const int node0 = cload(ka.meta, v.bidx / kMaxPrims).node;
    need = node0;
    rv.c0 = v.bc;
    rv.m0 = cload(ka.nodes, node0).col_mask;
    const DevPrim<float> P = prim_at(sc, v.bidx % kMaxPrims);
    if (P.type == PMP_PRIM_CLOUD) {
      // the gradient of sqrt(min |c - p|^2) at the nearest point, which the
      // reference's forward difference (eps 1e-6) approximates
      float d2;
      const V3<float> d = sub(v.bc, cloud_nearest(sc, (P.flags >> 4) & 15, v.bc, d2));
      rv.g0 = scale(1.0f / sqrtf(dot(d, d)), d);
    } else {
      rv.g0 = sdf_grad(P, v.bc);
    }
  }
  if (v.sidx >= 0) {
    const int n1 = cload(ka.meta, v.sidx / kMaxSpheres).node, n2 = cload(ka.meta, v.sidx % kMaxSpheres).node;
    need = max(need, max(n1, n2));
    rv.c1 = v.sc1;
    rv.c2 = v.sc2;
    const V3<float> d = sub(v.sc1, v.sc2);
    rv.n = scale(1.0f / sqrtf(dot(d, d)), d);
    rv.m1 = cload(ka.nodes, n1).col_mask;
    rv.m2 = cload(ka.nodes, n2).col_mask;
  }
  // lanes of the wave that reach this point walk together (inactive lanes
  // have returned); no winner in the wave: the rows are zero without FK
  const int last = wave_max_i(need);
  if (last >= 0) {
    rv.last = last;
    run_fk(ka, qs, false, rv);
  }
  int row = 0;
  if (NP > 0) {
    o.vals[b] = v.best;
#pragma unroll
    for (int j = 0; j < MAXC; ++j)
      if (j < D) o.jac[(int64_t)j * o.ld + b] = rv.re[j];
    row = 1;
  }
  if (ka.n_pairs > 0) {
    o.vals[(int64_t)row * o.ld + b] = v.sbest;
#pragma unroll
    for (int j = 0; j < MAXC; ++j)
      if (j < D) o.jac[((int64_t)row * D + j) * o.ld + b] = rv.rs[j];
  }
  if (o.offsets) smooth_terms<MAXC>(q, ldq, b, D, o);
}

// Whole mode: one wave per chunk of 64 configurations walks every sphere
// block and self pair of them and reduces in place.
template <int MAXC>
__global__ void __launch_bounds__(64) eval_units_kernel(KinArgs<float> ka, SceneArgs<float> sc, const double* q,
                                                        int64_t ldq, int64_t B, EvalOut o) {
  const int lane = threadIdx.x;
  const int64_t b = (int64_t)blockIdx.x * 64 + lane;
  const bool active = b < B;
  BatchQ qs{q, ldq, active ? b : B - 1};
  EvalUnitVis<MAXC> v{ka, sc, cload(o.units, 0)};
  run_fk(ka, qs, false, v);
  if (!active) return;
  eval_finish_whole<MAXC>(ka, sc, qs, q, ldq, b, o, v);
}

// Split mode: one workgroup per chunk of 64 configurations; its waves take
// the units in turn, leave their results in LDS, and after a barrier the
// first wave reduces them (no device-scope fence or atomic: a version with a
// wave per (chunk, unit) that met through a device-scope counter took
// longer).  LDS holds
// 16 B x 64 x n_units.
constexpr int kEvalGroupWaves = 8;
constexpr int kEvalGroupMaxUnits = 64;  // 64 KiB of LDS
template <int MAXC>
__global__ void __launch_bounds__(64 * kEvalGroupWaves) eval_group_kernel(KinArgs<float> ka, SceneArgs<float> sc,
                                                                          const double* q, int64_t ldq, int64_t B,
                                                                          EvalOut o) {
  HIP_DYNAMIC_SHARED(int32_t, eval_lds)
  const int U = o.n_units;
  float* xv = reinterpret_cast<float*>(eval_lds);
  int* xi = eval_lds + U * 64;
  float* sv = reinterpret_cast<float*>(eval_lds + 2 * U * 64);
  int* si = eval_lds + 3 * U * 64;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, W = blockDim.x >> 6;
  const int64_t b = (int64_t)blockIdx.x * 64 + lane;
  const bool active = b < B;
  BatchQ qs{q, ldq, active ? b : B - 1};
  for (int u = w; u < U; u += W) {
    EvalUnitVis<MAXC> v{ka, sc, cload(o.units, u)};
    run_fk(ka, qs, false, v);
    xv[u * 64 + lane] = v.best;
    xi[u * 64 + lane] = v.bidx;
    sv[u * 64 + lane] = v.sbest;
    si[u * 64 + lane] = v.sidx;
  }
  __syncthreads();
  if (w != 0 || !active) return;
  eval_finish<MAXC>(ka, sc, qs, q, ldq, b, o, kCutoff, -1, kCutoff, -1, xv + lane, xi + lane, sv + lane, si + lane, 64,
                    U);
}

// The pair and adjoint kernels give each (configuration, sphere item) its
// own lane: the wave's item is uniform (grid y), FK runs only down to the
// item's node, and the item's spheres meet every primitive.  Lanes b of a
// wave are consecutive, so every store and gradient load is coalesced.
// kCloud: the scene holds a point cloud (a separate kernel, so that the
// cloud search does not weigh on the register allocation of the others).
template <int MAXC, bool kCloud>
struct PairVis {
  static constexpr bool kSelf = false;
  const KinArgs<float>& ka;
  const SceneArgs<float>& sc;
  Frames<MAXC> fr;
  float* dist;
  float* grad;
  int64_t B, b;
  bool active;
  DevSphereItem it;
  int p_lo, p_hi;  // primitives of this lane's share
  bool done = false;
  __device__ bool alive() const { return !done; }
  template <class QS>
  __device__ void begin(const Pose<float>& base, const QS& qs) { fr.set_base(ka, base, qs); }
  __device__ void node(int n, const DevNode<float>& nd, const Pose<float>& P) {
    if (n == it.node) done = true;  // its spheres follow in this same step
    if (nd.q_col < 0) return;
    fr.set(nd.q_col, qrot(P.q, V3<float>{nd.axis[0], nd.axis[1], nd.axis[2]}), P.t);
  }
  __device__ void sphere(int s, V3<float> c, float r) {
    if (s < it.begin || s >= it.end) return;
    const DevSphereMeta m = cload(ka.meta, s);
    const uint64_t mask = cload(ka.nodes, m.node).col_mask;
    const int D = ka.n_cols, NP = sc.n_prims;
    V3<float> J[MAXC];
#pragma unroll
    for (int j = 0; j < MAXC; ++j) J[j] = fr.col(j, mask, ka.rev_mask, c);
    for (int p = p_lo; p < p_hi; ++p) {
      const DevPrim<float> P = prim_at(sc, p);
      float d;
      V3<float> g{0.f, 0.f, 0.f};
      if (kCloud && P.type == PMP_PRIM_CLOUD) {
        // CloudSDF::evaluate (primitive_sdf.hpp:585-590): sqrt of the nearest
        // squared distance minus r_cloud; its gradient at the nearest point
        float d2;
        const V3<float> dv = sub(c, cloud_nearest(sc, (P.flags >> 4) & 15, c, d2));
        d = sqrtf(d2) - P.p0 - r;
        g = scale(1.0f / sqrtf(dot(dv, dv)), dv);
      } else {
        d = sdf_eval(P, c) - r;
        if (grad) g = sdf_grad(P, c);
      }
      // the outputs are written once and not read back here: streaming
      // (non-temporal) stores keep them from displacing the tables in L2
      const int64_t row = (int64_t)m.out * NP + p;
      if (active) __builtin_nontemporal_store(d, dist + row * B + b);
      if (grad) {
#pragma unroll
        for (int j = 0; j < MAXC; ++j)
          if (j < D && active) __builtin_nontemporal_store(dot(g, J[j]), grad + (row * D + j) * B + b);
      }
    }
  }
  __device__ void self(const DevSelfPair&, int, int, V3<float>, V3<float>, float) {}
};

template <int MAXC, bool kCloud>
__global__ void __launch_bounds__(256) pair_kernel(KinArgs<float> ka, SceneArgs<float> sc, const double* q,
                                                   int64_t ldq, int64_t B, float* dist, float* grad) {
  const int64_t b = gid();
  const bool active = b < B;
  BatchQ qs{q, ldq, active ? b : B - 1};
  // grid z splits the primitives (p_per each), so that a batch of a few
  // thousand configurations still puts several waves on every SIMD: the
  // launch is bound by its stores, and more waves keep more of them in flight
  const int p_per = (sc.n_prims + (int)gridDim.z - 1) / (int)gridDim.z;
  const int p_lo = (int)blockIdx.z * p_per, p_hi = min(sc.n_prims, p_lo + p_per);
  PairVis<MAXC, kCloud> v{ka, sc, {}, dist, grad, B, b, active, cload(ka.items, blockIdx.y), p_lo, p_hi};
  run_fk(ka, qs, false, v);
}

// Jacobian of a point fixed to one link (rotation ignored): the joint frames
// of every node up to the link, then a_j x (p - o_j) (or a_j) for the columns
// that move it (KinematicModel::get_jacobian / get_attached_point_jacobian,
// kinematics/algorithm.cpp:212-433).
template <int MAXC>
struct LinkJacVis {
  static constexpr bool kSelf = false;
  const KinArgs<float>& ka;
  Frames<MAXC> fr;
  int target;
  V3<float> off;
  V3<float> p{0.f, 0.f, 0.f};
  Pose<float> lp{};  // the link's pose
  uint64_t mask = 0;
  bool done = false;
  __device__ bool alive() const { return !done; }
  template <class QS>
  __device__ void begin(const Pose<float>& base, const QS& qs) { fr.set_base(ka, base, qs); }
  __device__ void node(int n, const DevNode<float>& nd, const Pose<float>& P) {
    if (nd.q_col >= 0) fr.set(nd.q_col, qrot(P.q, V3<float>{nd.axis[0], nd.axis[1], nd.axis[2]}), P.t);
    if (n == target) {
      p = add(qrot(P.q, off), P.t);
      lp = P;
      mask = nd.col_mask;
      done = true;
    }
  }
  __device__ void sphere(int, V3<float>, float) {}
  __device__ void self(const DevSelfPair&, int, int, V3<float>, V3<float>, float) {}
};

template <int MAXC>
__global__ void __launch_bounds__(256) link_jac_kernel(KinArgs<float> ka, const double* q, int64_t ldq, int64_t B,
                                                       int target, float ox, float oy, float oz, float* jac,
                                                       float* pose) {
  const int64_t b = gid();
  const bool active = b < B;
  BatchQ qs{q, ldq, active ? b : B - 1};
  LinkJacVis<MAXC> v{ka, {}, target, V3<float>{ox, oy, oz}};
  run_fk(ka, qs, false, v);
  if (!active) return;
  const int D = ka.n_cols;
  if (pose) {
    const float pv[7] = {v.lp.t.x, v.lp.t.y, v.lp.t.z, v.lp.q.x, v.lp.q.y, v.lp.q.z, v.lp.q.w};
#pragma unroll
    for (int k = 0; k < 7; ++k) pose[(int64_t)k * B + b] = pv[k];
  }
#pragma unroll
  for (int j = 0; j < MAXC; ++j) {
    if (j < D) {
      const V3<float> c = v.fr.col(j, v.mask, ka.rev_mask, v.p);
      jac[((int64_t)0 * D + j) * B + b] = c.x;
      jac[((int64_t)1 * D + j) * B + b] = c.y;
      jac[((int64_t)2 * D + j) * B + b] = c.z;
    }
  }
}

// Reverse mode: per sphere the upstream weights of its primitive distances
// are folded into one world-space force g = sum_p w_p grad sdf_p(c); each
// q column then receives g . (a_j x (c - o_j)) (or g . a_j).  Each
// (configuration, item) lane writes its partial column sums; a second pass
// adds the items in a fixed order.
template <int MAXC>
struct VjpVis {
  static constexpr bool kSelf = false;
  const KinArgs<float>& ka;
  const SceneArgs<float>& sc;
  Frames<MAXC> fr;
  const float* gd;
  int64_t B, b;
  DevSphereItem it;
  int p_lo, p_hi;  // primitives of this lane's share
  float acc[MAXC];
  bool done = false;
  __device__ bool alive() const { return !done; }
  template <class QS>
  __device__ void begin(const Pose<float>& base, const QS& qs) { fr.set_base(ka, base, qs); }
  __device__ void node(int n, const DevNode<float>& nd, const Pose<float>& P) {
    if (n == it.node) done = true;
    if (nd.q_col < 0) return;
    fr.set(nd.q_col, qrot(P.q, V3<float>{nd.axis[0], nd.axis[1], nd.axis[2]}), P.t);
  }
  __device__ void sphere(int s, V3<float> c, float) {
    if (s < it.begin || s >= it.end) return;
    const DevSphereMeta m = cload(ka.meta, s);
    const uint64_t mask = cload(ka.nodes, m.node).col_mask;
    const int NP = sc.n_prims;
    V3<float> g{0.f, 0.f, 0.f};
    for (int p = p_lo; p < p_hi; ++p) {
      const DevPrim<float> P = prim_at(sc, p);
      const float w = gd[((int64_t)m.out * NP + p) * B + b];
      V3<float> gp;
      if (P.type == PMP_PRIM_CLOUD) {
        float d2;
        const V3<float> dv = sub(c, cloud_nearest(sc, (P.flags >> 4) & 15, c, d2));
        gp = scale(1.0f / sqrtf(dot(dv, dv)), dv);
      } else {
        gp = sdf_grad(P, c);
      }
      g = add(g, scale(w, gp));
    }
#pragma unroll
    for (int j = 0; j < MAXC; ++j) acc[j] += dot(g, fr.col(j, mask, ka.rev_mask, c));
  }
  __device__ void self(const DevSelfPair&, int, int, V3<float>, V3<float>, float) {}
};

template <int MAXC>
__global__ void __launch_bounds__(256) vjp_kernel(KinArgs<float> ka, SceneArgs<float> sc, const double* q,
                                                  int64_t ldq, int64_t B, const float* gd, float* part) {
  const int64_t b = gid();
  const bool active = b < B;
  const int64_t bb = active ? b : B - 1;
  BatchQ qs{q, ldq, bb};
  // grid z splits the primitives into shares, as in pair_kernel; each
  // (item, share) writes its own partial column sums
  const int p_per = (sc.n_prims + (int)gridDim.z - 1) / (int)gridDim.z;
  const int p_lo = (int)blockIdx.z * p_per, p_hi = min(sc.n_prims, p_lo + p_per);
  VjpVis<MAXC> v{ka, sc, {}, gd, B, bb, cload(ka.items, blockIdx.y), p_lo, p_hi};
#pragma unroll
  for (int j = 0; j < MAXC; ++j) v.acc[j] = 0.f;
  run_fk(ka, qs, false, v);
  if (!active) return;
  const int D = ka.n_cols;
  const int64_t part_row = (int64_t)blockIdx.y * gridDim.z + blockIdx.z;
#pragma unroll
  for (int j = 0; j < MAXC; ++j)
    if (j < D) part[(part_row * D + j) * B + b] = v.acc[j];
}

// Sum of the n_parts partial column sums (items x shares, in that order:
// deterministic).
__global__ void __launch_bounds__(256) vjp_reduce_kernel(const float* part, int n_parts, int D, int64_t B,
                                                         float* grad_q) {
  const int64_t i = gid();  // j * B + b
  if (i >= (int64_t)D * B) return;
  float s = 0.f;
  for (int it = 0; it < n_parts; ++it) s += part[(int64_t)it * D * B + i];
  grad_q[i] = s;
}

// ---------------------------------------------------- point queries (FP64)
// SDF of the union of `n` primitives at points (UnionSDF::evaluate: min;
// UnionSDF::is_outside: all outside), collision/primitive_sdf.hpp:155-170.
__global__ void __launch_bounds__(256) prims_eval_kernel(const DevPrim<double>* prims, int n,
                                                         const double* cloud_xyz, const double* pts, int64_t N,
                                                         int mode, double radius, double* out) {
  const int64_t i = gid();
  if (i >= N) return;
  const V3<double> c{pts[i], pts[N + i], pts[2 * N + i]};
  double val = DBL_MAX;
  bool outside = true;
  for (int k = 0; k < n; ++k) {
    const DevPrim<double>& P = prims[k];
    double d;
    if (P.type == PMP_PRIM_CLOUD) {
      double best = DBL_MAX;
      for (int j = 0; j < P.cloud_count; ++j) {
        const double* y = cloud_xyz + 4 * (int64_t)(P.cloud_begin + j);
        const double dx = y[0] - c.x, dy = y[1] - c.y, dz = y[2] - c.z;
        const double s = dx * dx + dy * dy + dz * dz;
        if (s < best) best = s;
      }
      d = sqrt(best) - P.p0;
      if (!(d > radius)) outside = false;
    } else {
      d = sdf_eval(P, c);
      if (!is_outside_ref(P, c, radius)) outside = false;
    }
    val = d < val ? d : val;
  }
  out[i] = mode == 0 ? val : (outside ? 1.0 : 0.0);
}

// Exact nearest neighbour, brute force with LDS tiles (KDTree::query / sqdist).
__global__ void __launch_bounds__(256) nearest_kernel(const double* pts, int64_t N, const double* qry, int64_t M,
                                                      double* sqd, int32_t* idx) {
  __shared__ double tile[3][256];
  const int64_t i = gid();
  const bool active = i < M;
  V3<double> c{0, 0, 0};
  if (active) c = {qry[i], qry[M + i], qry[2 * M + i]};
  double best = DBL_MAX;
  int32_t bi = -1;
  for (int64_t base = 0; base < N; base += 256) {
    const int64_t j = base + threadIdx.x;
    __syncthreads();
    if (j < N) {
      tile[0][threadIdx.x] = pts[j];
      tile[1][threadIdx.x] = pts[N + j];
      tile[2][threadIdx.x] = pts[2 * N + j];
    }
    __syncthreads();
    const int cnt = (int)(N - base < 256 ? N - base : 256);
    for (int k = 0; k < cnt; ++k) {
      const double dx = tile[0][k] - c.x, dy = tile[1][k] - c.y, dz = tile[2][k] - c.z;
      const double s = dx * dx + dy * dy + dz * dz;
      if (s < best) {
        best = s;
        bi = (int32_t)(base + k);
      }
    }
  }
  if (active) {
    sqd[i] = best;
    idx[i] = bi;
  }
}

// ------------------------------------------------------------ launchers
static inline unsigned blocks_for(int64_t n, int t) { return (unsigned)((n + t - 1) / t); }

hipError_t launch_centers(const KinArgs<float>& ka, const double* q, int64_t ldq, int64_t B, int taylor, float* out,
                          hipStream_t st) {
  const int bk = slot_block(ka, 256);
  hipLaunchKernelGGL(centers_kernel<float>, dim3(blocks_for(B, bk)), dim3(bk), slot_bytes(ka, bk), st, ka, q, ldq, B,
                     taylor, out);
  return hipGetLastError();
}
hipError_t launch_poses(const KinArgs<double>& ka, const double* q, int64_t ldq, int64_t B, int taylor, double* out,
                        hipStream_t st) {
  const int bk = slot_block(ka, 256);
  hipLaunchKernelGGL(poses_kernel, dim3(blocks_for(B, bk)), dim3(bk), slot_bytes(ka, bk), st, ka, q, ldq, B, taylor,
                     out);
  return hipGetLastError();
}
hipError_t launch_clearance(const KinArgs<float>& ka, const SceneArgs<float>& sc, const double* q, int64_t ldq,
                            int64_t B, int taylor, float* out, float* cen, hipStream_t st) {
  const int bk = slot_block(ka, 256);
  hipLaunchKernelGGL(clearance_kernel, dim3(blocks_for(B, bk)), dim3(bk), slot_bytes(ka, bk), st, ka, sc, q, ldq, B,
                     taylor, out, cen);
  return hipGetLastError();
}
hipError_t launch_valid(const KinArgs<float>& k32, const KinArgs<double>& k64, const SceneArgs<float>& s32,
                        const SceneArgs<double>& s64, const double* q, int64_t ldq, int64_t B, float band,
                        uint8_t* valid, int32_t* n_recheck, int block, const CloudWork& cw,
                        const CloudRange* clouds, int n_clouds, int32_t* flags, int32_t* unit_flags,
                        hipStream_t st) {
  if (block <= 0) block = 256;
  block = slot_block(k32, block);
  // the unit flags become answers in the FP64 pass (no clouds) or in the
  // first cloud's merge
  const int64_t Q = (int64_t)k32.n_ext_spheres * B;
  const bool units = unit_flags && k32.n_units > 0;
  const bool fused = units && n_clouds == 0;
  const bool merged = units && n_clouds > 0 && Q > 0;
  if (units) {
    hipLaunchKernelGGL(valid_units_kernel, dim3(blocks_for(B, 64), (unsigned)k32.n_units), dim3(64),
                       slot_bytes(k32, 64), st, k32, s32, q, ldq, B, band, unit_flags, cw);
    if (!fused && !merged)
      hipLaunchKernelGGL(valid_units_finish_kernel, dim3(blocks_for(B, 256)), dim3(256), 0, st, unit_flags, B, valid);
  } else {
    hipLaunchKernelGGL(valid_kernel, dim3(blocks_for(B, block)), dim3(block), slot_bytes(k32, block), st, k32, s32, q,
                       ldq, B, band, valid, cw);
  }
  // sphere-vs-cloud queries on the matrix cores, one pass per cloud
  for (int oc = 0; oc < n_clouds && Q > 0; ++oc) {
    hipError_t e = launch_cloud_valid(s32.tiles, s32.tileB, s32.supers, clouds[oc], cw.cen, cw.cap2[oc], cw.rr[oc],
                                      B, Q, band, valid, flags, merged && oc == 0 ? unit_flags : nullptr, st);
    if (e != hipSuccess) return e;
  }
  hipLaunchKernelGGL(valid_recheck_kernel, dim3(blocks_for(B, 64)), dim3(64), slot_bytes(k64, 64), st, k64, s64, q,
                     ldq, B, valid, n_recheck, fused ? unit_flags : nullptr);
  return hipGetLastError();
}
hipError_t launch_edges(const KinArgs<float>& k32, const KinArgs<double>& k64, const SceneArgs<float>& s32,
                        const SceneArgs<double>& s64, const EdgeArgs& ea, float band, int grid, int block,
                        hipStream_t st, hipEvent_t ev_before, hipEvent_t ev_after) {
  if (ev_before) {
    hipError_t e = hipEventRecord(ev_before, st);
    if (e != hipSuccess) return e;
  }
  // scenes with a point cloud take the variant that tests clouds in the walk;
  // a wave's FP32 walk and its FP64 re-decisions after it share the wave's
  // kept-pose chunk, sized for FP64
  KinArgs<float> kw = k32;
  kw.slot_wave_bytes = k64.slot_wave_bytes;
  const size_t lds = slot_bytes(k64, block);
  if (s32.n_cloud > 0)
    hipLaunchKernelGGL(edges_kernel<true>, dim3(grid), dim3(block), lds, st, kw, s32, k64, s64, ea, band);
  else
    hipLaunchKernelGGL(edges_kernel<false>, dim3(grid), dim3(block), lds, st, kw, s32, k64, s64, ea, band);
  if (ev_after) {
    hipError_t e = hipEventRecord(ev_after, st);
    if (e != hipSuccess) return e;
  }
  const unsigned rgrid = (unsigned)(ea.E < kRecheckGrid ? ea.E : kRecheckGrid);
  const int rblock = slot_block(k64, 64 * kRecheckWaves, sizeof(int));
  hipLaunchKernelGGL(edges_recheck_kernel, dim3(rgrid), dim3(rblock), slot_bytes(k64, rblock), st, k64, s64, ea);
  return hipGetLastError();
}

#define PMP_DISPATCH_MAXC(ncols, KERNEL, ...)                                                 \
  do {                                                                                        \
    if ((ncols) <= 8)                                                                         \
      hipLaunchKernelGGL(KERNEL<8>, __VA_ARGS__);                                             \
    else if ((ncols) <= 16)                                                                   \
      hipLaunchKernelGGL(KERNEL<16>, __VA_ARGS__);                                            \
    else if ((ncols) <= 32)                                                                   \
      hipLaunchKernelGGL(KERNEL<32>, __VA_ARGS__);                                            \
    else                                                                                      \
      hipLaunchKernelGGL(KERNEL<64>, __VA_ARGS__);                                            \
  } while (0)

#define PMP_DISPATCH_MAXC_CLOUD(ncols, KERNEL, CLOUD, ...)                                   \
  do {                                                                                        \
    if ((ncols) <= 8)                                                                         \
      hipLaunchKernelGGL((KERNEL<8, CLOUD>), __VA_ARGS__);                                    \
    else if ((ncols) <= 16)                                                                   \
      hipLaunchKernelGGL((KERNEL<16, CLOUD>), __VA_ARGS__);                                   \
    else if ((ncols) <= 32)                                                                   \
      hipLaunchKernelGGL((KERNEL<32, CLOUD>), __VA_ARGS__);                                   \
    else                                                                                      \
      hipLaunchKernelGGL((KERNEL<64, CLOUD>), __VA_ARGS__);                                   \
  } while (0)

__global__ void fill_f32_kernel(float* p, float v, int64_t n) {
  const int64_t i = gid();
  if (i < n) p[i] = v;
}
__global__ void fill_counts_kernel(int32_t* p, int32_t v, int64_t n) {
  const int64_t i = gid();
  if (i < n) p[i] = v;
}
// SequentialCst::evaluate (constraints/sequential_constraint.cpp:63-127) of
// the SQP inequality sequence (nlp_solver/sqp_based_solver.py:48-71), one
// thread per column (waypoint n, joint d) of its trajectory's matrix.  Rows:
// the collision constraint's C rows of every waypoint (t-major), then the
// motion-step box pairs q_t - q_{t+1} + w, q_{t+1} - q_t + w of every segment.
// Column (t, d) holds, in row order, the C collision entries J + 1e-12, the
// (-1, +1) of segment t - 1 and the (+1, -1) of segment t, so its position in
// the compressed arrays follows from (t, d) alone.  Consecutive columns own
// consecutive ranges of the value and row arrays (trajectories follow each
// other there too), so a workgroup stages its columns' entries in LDS and
// writes its whole range with consecutive lanes on consecutive entries.
// Threads past the columns close the column pointers of empty trajectories.
constexpr int kSqpBlock = 256;
__global__ void __launch_bounds__(kSqpBlock) sqp_ineq_kernel(SqpArgs a) {
  HIP_DYNAMIC_SHARED(double, sqp_lds)
  __shared__ int64_t range[2];  // the workgroup's first entry and the end of its last
  const int64_t i = gid();
  const int D = a.D, C = a.C;
  double* sv = sqp_lds;
  int32_t* sr = reinterpret_cast<int32_t*>(sqp_lds + (int64_t)kSqpBlock * (C + 4));
  const int64_t i_first = (int64_t)blockIdx.x * kSqpBlock, n_cols = a.N * D;
  const int64_t i_last = (i_first + kSqpBlock < n_cols ? i_first + kSqpBlock : n_cols) - 1;
  const bool col = i < n_cols;
  // the column: trajectory lo, waypoint t of T, joint d; its entries start at g
  int lo = 0;
  int64_t n = 0, b0 = 0, T = 0, t = 0, before = 0, g = 0;
  int d = 0, et = 0;
  const bool box = a.msbox != nullptr;
  if (col) {
    n = i / D;
    d = (int)(i % D);
    int hi = a.n_traj;  // offsets[lo] <= n < offsets[hi]
    while (hi - lo > 1) {
      const int mid = (lo + hi) >> 1;
      if (a.offsets[mid] <= n) lo = mid;
      else hi = mid;
    }
    b0 = a.offsets[lo];
    T = a.offsets[lo + 1] - b0;
    t = n - b0;
    const int64_t j = t * D + d;
    // box entries of the columns of the waypoints before t, and of each column of t
    const int64_t e_prev = box ? 2 * (t > 0 ? t - 1 : 0) + 2 * (t < T - 1 ? t : T - 1) : 0;
    et = box ? 2 * (t > 0) + 2 * (t < T - 1) : 0;
    before = j * C + (int64_t)D * e_prev + (int64_t)d * et;
    int32_t* cp = a.colptr + D * b0 + lo;
    cp[j] = (int32_t)before;
    if (j == D * T - 1) cp[D * T] = (int32_t)(before + C + et);
    g = a.nnz_off[lo] + before;
    if (i == i_first) range[0] = g;
    if (i == i_last) range[1] = g + C + et;
  } else {
    const int64_t tr = i - n_cols;
    if (tr < a.n_traj && a.offsets[tr + 1] == a.offsets[tr]) a.colptr[D * a.offsets[tr] + tr] = 0;
  }
  __syncthreads();
  if (col) {
    // the column's entries in LDS at their offset from the workgroup's first
    const int64_t o = g - range[0];
    for (int c = 0; c < C; ++c) {
      sr[o + c] = (int32_t)(t * C + c);
      sv[o + c] = (double)a.jac[((int64_t)c * D + d) * a.N + n] + 1e-12;
    }
    double* cv = a.c + a.row_off[lo];
    if (d == 0)
      for (int c = 0; c < C; ++c) cv[t * C + c] = (double)a.vals[(int64_t)c * a.N + n];
    if (box) {
      const int64_t m0 = C * T;
      int64_t k = o + C;
      if (t > 0) {
        const int64_t r = m0 + (t - 1) * 2 * D;
        sr[k] = (int32_t)(r + d);
        sv[k] = -1.0;
        sr[k + 1] = (int32_t)(r + D + d);
        sv[k + 1] = 1.0;
        k += 2;
      }
      if (t < T - 1) {
        const int64_t r = m0 + t * 2 * D;
        sr[k] = (int32_t)(r + d);
        sv[k] = 1.0;
        sr[k + 1] = (int32_t)(r + D + d);
        sv[k + 1] = -1.0;
        const double q1 = a.X[(int64_t)d * a.ldx + n], q2 = a.X[(int64_t)d * a.ldx + n + 1], w = a.msbox[d];
        cv[r + d] = q1 - q2 + w;
        cv[r + D + d] = q2 - q1 + w;
      }
    }
  }
  __syncthreads();
  if (i_first >= n_cols) return;  // a workgroup of threads past the columns (uniform)
  const int64_t base = range[0], cnt = range[1] - range[0];
  for (int64_t e = threadIdx.x; e < cnt; e += kSqpBlock) {
    a.values[base + e] = sv[e];
    a.rowind[base + e] = sr[e];
  }
}

hipError_t launch_link_jacobian(const KinArgs<float>& ka, const double* q, int64_t ldq, int64_t B, int link,
                                const float (&off)[3], float* jac, float* pose, hipStream_t st) {
  if (B > 0) {
    const int bk = slot_block(ka, 256);
    PMP_DISPATCH_MAXC(ka.n_cols, link_jac_kernel, dim3(blocks_for(B, bk)), dim3(bk), slot_bytes(ka, bk), st, ka, q,
                      ldq, B, link, off[0], off[1], off[2], jac, pose);
  }
  return hipGetLastError();
}

int edges_resident_waves_per_cu(bool cloud, int block, size_t lds) {
  int n = 0;
  const hipError_t err =
      cloud ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, edges_kernel<true>, block, lds)
            : hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, edges_kernel<false>, block, lds);
  return err == hipSuccess ? n * (block / 64) : 0;
}

hipError_t launch_sqp_ineq(const SqpArgs& a, hipStream_t st) {
  const int64_t n = a.N * a.D + a.n_traj;
  if (a.C < 0 || a.C > 2) return hipErrorInvalidValue;  // a column holds at most C + 4 <= 6 entries
  const size_t lds = (size_t)kSqpBlock * (a.C + 4) * (sizeof(double) + sizeof(int32_t));
  if (n > 0) hipLaunchKernelGGL(sqp_ineq_kernel, dim3(blocks_for(n, kSqpBlock)), dim3(kSqpBlock), lds, st, a);
  return hipGetLastError();
}

hipError_t launch_fill_i32(int32_t* p, int32_t v, int64_t n, hipStream_t st) {
  if (n > 0) hipLaunchKernelGGL(fill_counts_kernel, dim3(blocks_for(n, 256)), dim3(256), 0, st, p, v, n);
  return hipGetLastError();
}
hipError_t launch_eval_units(const KinArgs<float>& ka, const SceneArgs<float>& sc, const double* q, int64_t ldq,
                             int64_t B, const EvalArgs& e, hipStream_t st) {
  const int64_t chunks = (B + 63) / 64;
  EvalOut o{};
  o.vals = e.vals;
  o.jac = e.jac;
  o.ld = B;
  // Small batches give each configuration many units (one per sphere block
  // and per self pair, the units of a chunk in one workgroup) so that the
  // launch fills the GPU; from kWholeUnitChunks chunks on (or beyond
  // kEvalGroupMaxUnits units), one unit walks the whole configuration.
  const int64_t whole = e.whole_chunks > 0 ? e.whole_chunks : kWholeUnitChunks;
  const int waves = ka.n_units < kEvalGroupWaves ? ka.n_units : kEvalGroupWaves;
  const size_t own = sizeof(int32_t) * 4 * 64 * (size_t)ka.n_units;  // (value, winner) pairs per unit and lane
  const bool split = chunks < whole && ka.n_units > 1 && ka.n_units <= kEvalGroupMaxUnits &&
                     own + slot_bytes(ka, 64 * waves) <= (size_t)ka.lds_limit;
  o.units = split ? ka.units : ka.unit_all;
  o.n_units = split ? ka.n_units : 1;
  o.offsets = e.offsets;
  o.n_traj = e.n_traj;
  o.w2 = e.w2;
  o.cost = e.cost;
  o.grad = e.grad;
  if (e.offsets)  // empty trajectories have no lane of their own
    hipLaunchKernelGGL(fill_f32_kernel, dim3(blocks_for(e.n_traj, 256)), dim3(256), 0, st, e.cost, 0.0f,
                       (int64_t)e.n_traj);
  if (B == 0) return hipGetLastError();
  if (split) {
    KinArgs<float> kg = ka;  // slots after the unit partials
    kg.slot_off = (int32_t)own;
    PMP_DISPATCH_MAXC(ka.n_cols, eval_group_kernel, dim3((unsigned)chunks), dim3(64 * waves),
                      own + slot_bytes(ka, 64 * waves), st, kg, sc, q, ldq, B, o);
  } else {
    PMP_DISPATCH_MAXC(ka.n_cols, eval_units_kernel, dim3((unsigned)chunks), dim3(64), slot_bytes(ka, 64), st, ka, sc,
                      q, ldq, B, o);
  }
  return hipGetLastError();
}
// Waves the pair launch aims for: four per SIMD of a 256-CU MI355X.
constexpr int64_t kPairWaves = 4096;
hipError_t launch_pairs(const KinArgs<float>& ka, const SceneArgs<float>& sc, const double* q, int64_t ldq,
                        int64_t B, float* dist, float* grad, int shares, hipStream_t st) {
  if (ka.n_items == 0 || sc.n_prims == 0) return hipSuccess;
  const int bk = slot_block(ka, 256);
  const size_t lds = slot_bytes(ka, bk);
  // primitive shares: the schedule's, or enough for about kPairWaves waves;
  // at least two primitives a share
  const int64_t waves = (B + 63) / 64 * ka.n_items;
  const int64_t want = shares > 0 ? shares : (kPairWaves + waves - 1) / waves, most = (sc.n_prims + 1) / 2;
  const int nz = (int)(want < most ? want : most) < 1 ? 1 : (int)(want < most ? want : most);
  const dim3 grid(blocks_for(B, bk), (unsigned)ka.n_items, (unsigned)nz);
  if (sc.n_cloud > 0)
    PMP_DISPATCH_MAXC_CLOUD(ka.n_cols, pair_kernel, true, grid, dim3(bk), lds, st, ka, sc, q, ldq, B, dist, grad);
  else
    PMP_DISPATCH_MAXC_CLOUD(ka.n_cols, pair_kernel, false, grid, dim3(bk), lds, st, ka, sc, q, ldq, B, dist, grad);
  return hipGetLastError();
}
int vjp_shares(const KinArgs<float>& ka, int n_prims, int64_t B, int shares) {
  // the schedule's shares, or about three waves per SIMD (one round at the
  // adjoint's register use); at least two primitives a share
  const int64_t waves = (B + 63) / 64 * ka.n_items;
  const int64_t want = shares > 0 ? shares : (waves > 0 ? 3072 / waves : 1), most = (n_prims + 1) / 2;
  const int64_t z = want < most ? want : most;
  return z < 1 ? 1 : (int)z;
}
hipError_t launch_vjp(const KinArgs<float>& ka, const SceneArgs<float>& sc, const double* q, int64_t ldq, int64_t B,
                      const float* gd, float* part, float* grad_q, int shares, hipStream_t st) {
  const int nz = vjp_shares(ka, sc.n_prims, B, shares);
  if (ka.n_items > 0) {
    const int bk = slot_block(ka, 256);
    PMP_DISPATCH_MAXC(ka.n_cols, vjp_kernel, dim3(blocks_for(B, bk), (unsigned)ka.n_items, (unsigned)nz), dim3(bk),
                      slot_bytes(ka, bk), st, ka, sc, q, ldq, B, gd, part);
  }
  hipLaunchKernelGGL(vjp_reduce_kernel, dim3(blocks_for((int64_t)ka.n_cols * B, 256)), dim3(256), 0, st, part,
                     ka.n_items * nz, ka.n_cols, B, grad_q);
  return hipGetLastError();
}
hipError_t launch_prims_eval(const DevPrim<double>* prims, int n, const double* cloud_xyz, const double* pts,
                             int64_t N, int mode, double radius, double* out, hipStream_t st) {
  hipLaunchKernelGGL(prims_eval_kernel, dim3(blocks_for(N, 256)), dim3(256), 0, st, prims, n, cloud_xyz, pts, N,
                     mode, radius, out);
  return hipGetLastError();
}
hipError_t launch_nearest(const double* pts, int64_t N, const double* qry, int64_t M, double* sqd, int32_t* idx,
                          hipStream_t st) {
  hipLaunchKernelGGL(nearest_kernel, dim3(blocks_for(M, 256)), dim3(256), 0, st, pts, N, qry, M, sqd, idx);
  return hipGetLastError();
}

}  // namespace pmp
