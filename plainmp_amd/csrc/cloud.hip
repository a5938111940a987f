// Sphere-vs-point-cloud distances on the matrix cores.
//
// A cloud is stored as tiles of 64 points (Morton order, so each tile is
// spatially compact) with the tile's AABB and centre o.  For a sphere centre
// c the squared distance to point p is
//     |c - p|^2 = |c - o|^2 + ( |p - o|^2 - 2 (c - o) . (p - o) )
// and the bracket over 16 centres x 16 points is one K = 4 MFMA:
//     A (16 x 4) = [ c - o , 1 ],   B (4 x 16) = [ -2 (p - o) ; |p - o|^2 ].
// Working in tile-local coordinates keeps every term near the tile size; the
// FP32 cancellation error is bounded per tile by kCloudGemmErr (|c'|^2 +
// rho^2) (fk.hpp), which validity decisions add to their band.  A wave
// handles 16 queries; a tile is skipped when no query's cap reaches the
// tile's AABB.
#include <float.h>

#include "fk.hpp"
#include "launch.h"

namespace pmp {

typedef float floatx4 __attribute__((ext_vector_type(4)));

// Squared distance between two boxes (0 when they overlap).
__device__ __forceinline__ float box_box_d2(float alx, float aly, float alz, float aux, float auy, float auz,
                                            float blx, float bly, float blz, float bux, float buy, float buz) {
  const float dx = fmaxf(fmaxf(alx - bux, blx - aux), 0.f);
  const float dy = fmaxf(fmaxf(aly - buy, bly - auy), 0.f);
  const float dz = fmaxf(fmaxf(alz - buz, blz - auz), 0.f);
  return dx * dx + dy * dy + dz * dz;
}

__device__ __forceinline__ float wave_min16(float v) {
  for (int s = 1; s < 16; s <<= 1) v = fminf(v, __shfl_xor(v, s, 64));
  return v;
}
__device__ __forceinline__ float wave_max16(float v) {
  for (int s = 1; s < 16; s <<= 1) v = fmaxf(v, __shfl_xor(v, s, 64));
  return v;
}

// Decision arguments of the bounds instantiation: rr[q] = r + rc of a live
// query (< 0 when not queried), the band, and the per-configuration flag
// word the decisions are ORed into.
struct CloudDecide {
  const float* rr;
  float band;
  int32_t* flags;
  const int32_t* unit_flags;  // the unit pass's words instead of `skip`, when not yet turned into valid[]
};

// min |c - p|^2 over the cloud's tiles for the queries q = e * B + b
// (e < Q / B spheres, b < B configurations) whose cap2[q] >= 0; centres in
// cen[(e * 3 + k) * B + b].  Writes out_d2[q] = min squared distance
// (FLT_MAX when no point lies within the cap).  With kBounds (validity) the
// minimum is bracketed instead, between min over tiles of (d2 - err) and of
// (d2 + err), err = kCloudGemmErr (|c'|^2 + rho^2) per tile, and each live
// query is decided from its bracket in place, ORing into dec.flags[b]; a wave
// whose 16 queries are all unqueried or skipped does nothing.  `skip`
// (optional, per configuration b): configurations already decided (0) are
// not evaluated.
//
// A wave takes 16 consecutive spheres of one configuration (spheres are
// ordered by link, so they are close together; consecutive waves take the
// next configurations) and culls by the box around
// its queries grown by their largest reach: 64 super-tiles at a time are
// tested against that box, a lane each, and a ballot lists the survivors;
// then a survivor's tiles, a lane each; a surviving tile is then tested per
// query and, if any query reaches it, compared on the matrix cores.
template <bool kBounds>
__global__ void __launch_bounds__(256) cloud_min_kernel(const CloudTile* tiles, const float* tileB,
                                                        const CloudSuper* supers, int super_begin, int n_super,
                                                        const float* cen, const float* cap2, int64_t B, int64_t Q,
                                                        const uint8_t* skip, float* out_d2, CloudDecide dec) {
  const int lane = threadIdx.x & 63;
  const int64_t S = Q / B;  // spheres per configuration
  const int64_t groups = (S + 15) / 16;
  // Waves in configuration-fastest order, workgroups handed to the XCDs in
  // turn and each XCD given a contiguous range: neighbouring configurations
  // read the same lines of the [sphere][B] query arrays, in one L2.
  const int g = blockIdx.x, xcd = g % kXcds;
  int64_t first = 0;
  for (int x = 0; x < xcd; ++x) first += ((int64_t)gridDim.x - x + kXcds - 1) / kXcds;
  const int64_t w = (first + g / kXcds) * (blockDim.x >> 6) + (threadIdx.x >> 6);
  if (w >= groups * B) return;  // wave-uniform
  const int64_t b = w % B, e0 = (w / B) * 16;
  const int i = lane & 15, k = lane >> 4;
  // skipped: already decided by a certain collision (unit words: bit 0
  // without bit 2)
  const bool skipped = kBounds && dec.unit_flags ? (dec.unit_flags[b] & 5) == 1 : (skip && skip[b] == 0);
  // the A-row query of this lane; a wave none of whose queries is live ends
  // here without writing: the decisions read a query's bracket only when the
  // query is live (rr >= 0, written with cap2), and a skipped configuration
  // is already decided
  float capi = -1.f;
  if (e0 + i < S && !skipped) capi = cap2 ? cap2[(e0 + i) * B + b] : FLT_MAX;
  if (!__any(capi >= 0.f)) return;
  float cx = 0.f, cy = 0.f, cz = 0.f;
  if (e0 + i < S) {
    const int64_t e = e0 + i;
    cx = cen[(e * 3 + 0) * B + b];
    cy = cen[(e * 3 + 1) * B + b];
    cz = cen[(e * 3 + 2) * B + b];
  }
  // centres of the four output rows 4k + r held by this lane
  float rx[4], ry[4], rz[4];
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int64_t e = e0 + 4 * k + r;
    rx[r] = ry[r] = rz[r] = 0.f;
    if (e < S) {
      rx[r] = cen[(e * 3 + 0) * B + b];
      ry[r] = cen[(e * 3 + 1) * B + b];
      rz[r] = cen[(e * 3 + 2) * B + b];
    }
  }
  float best[4] = {FLT_MAX, FLT_MAX, FLT_MAX, FLT_MAX};
  float lo[4] = {FLT_MAX, FLT_MAX, FLT_MAX, FLT_MAX};
  const floatx4 zero = {0.f, 0.f, 0.f, 0.f};
  {
    // the box around the live queries and their largest cap
    const bool live = capi >= 0.f;
    const float qlx = wave_min16(live ? cx : FLT_MAX), qly = wave_min16(live ? cy : FLT_MAX),
                qlz = wave_min16(live ? cz : FLT_MAX);
    const float qux = wave_max16(live ? cx : -FLT_MAX), quy = wave_max16(live ? cy : -FLT_MAX),
                quz = wave_max16(live ? cz : -FLT_MAX);
    const float cap = wave_max16(capi);
    for (int s0 = 0; s0 < n_super; s0 += 64) {
      bool near = false;
      if (s0 + lane < n_super) {
        const CloudSuper Sv = supers[super_begin + s0 + lane];
        near = box_box_d2(Sv.lb[0], Sv.lb[1], Sv.lb[2], Sv.ub[0], Sv.ub[1], Sv.ub[2], qlx, qly, qlz, qux, quy, quz) <=
               cap;
      }
      uint64_t smask = __ballot(near);
      while (smask) {
        const int sj = __builtin_ctzll(smask);
        smask &= smask - 1;
        const CloudSuper Su = cload(supers, super_begin + s0 + sj);
        // the super-tile's tiles, a lane each
        bool tnear = false;
        if (lane < Su.count) {
          const CloudTile Tv = tiles[Su.first + lane];
          tnear = box_box_d2(Tv.lb[0], Tv.lb[1], Tv.lb[2], Tv.ub[0], Tv.ub[1], Tv.ub[2], qlx, qly, qlz, qux, quy,
                             quz) <= cap;
        }
        uint64_t tmask = __ballot(tnear);
        while (tmask) {
          const int tj = __builtin_ctzll(tmask);
          tmask &= tmask - 1;
          const int t = Su.first + tj;
          const CloudTile T = cload(tiles, t);
          const float dx = fmaxf(fmaxf(T.lb[0] - cx, cx - T.ub[0]), 0.f);
          const float dy = fmaxf(fmaxf(T.lb[1] - cy, cy - T.ub[1]), 0.f);
          const float dz = fmaxf(fmaxf(T.lb[2] - cz, cz - T.ub[2]), 0.f);
          if (!__any(dx * dx + dy * dy + dz * dz <= capi)) continue;
          const float a = k == 0 ? cx - T.o[0] : (k == 1 ? cy - T.o[1] : (k == 2 ? cz - T.o[2] : 1.f));
          const float* Bt = tileB + (int64_t)t * 256;
          floatx4 d0 = __builtin_amdgcn_mfma_f32_16x16x4f32(a, Bt[lane], zero, 0, 0, 0);
          floatx4 d1 = __builtin_amdgcn_mfma_f32_16x16x4f32(a, Bt[64 + lane], zero, 0, 0, 0);
          floatx4 d2 = __builtin_amdgcn_mfma_f32_16x16x4f32(a, Bt[128 + lane], zero, 0, 0, 0);
          floatx4 d3 = __builtin_amdgcn_mfma_f32_16x16x4f32(a, Bt[192 + lane], zero, 0, 0, 0);
          const float rho2 = kBounds ? tile_rho2(T) : 0.f;
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const float ox = rx[r] - T.o[0], oy = ry[r] - T.o[1], oz = rz[r] - T.o[2];
            const float off = ox * ox + oy * oy + oz * oz;
            const float m = fminf(fminf(d0[r], d1[r]), fminf(d2[r], d3[r])) + off;
            if constexpr (kBounds) {
              const float err = kCloudGemmErr * (off + rho2);
              best[r] = fminf(best[r], m + err);
              lo[r] = fminf(lo[r], m - err);
            } else {
              best[r] = fminf(best[r], m);
            }
          }
        }
      }
    }
  }
  // minimum over the 16 lanes that share the rows (columns j = lane & 15)
#pragma unroll
  for (int r = 0; r < 4; ++r)
    for (int s = 1; s < 16; s <<= 1) {
      best[r] = fminf(best[r], __shfl_xor(best[r], s, 64));
      if constexpr (kBounds) lo[r] = fminf(lo[r], __shfl_xor(lo[r], s, 64));
    }
  if constexpr (kBounds) {
    // Collision decision per live query: collide iff sqrt(d2) - rc <= r,
    // i.e. d2 <= (r + rc)^2 (CloudSDF::is_outside, primitive_sdf.hpp:597-603),
    // where the exact minimum lies in [lo, best]: certain outside the band
    // around R^2, bit 0 (a collision) or nothing; bit 1 (in the band)
    // otherwise.  The wave's 16 queries are of one configuration: one OR.
    int bits = 0;
    if (i == 0) {
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int64_t e = e0 + 4 * k + r;
        if (e < S) {
          const float R = dec.rr[e * B + b];
          if (R >= 0.f) {
            const float R2 = R * R, w = dec.band * (2.f * R + dec.band);
            if (best[r] - R2 < -w)
              bits |= 1;
            else if (!(lo[r] - R2 > w))
              bits |= 2;
          }
        }
      }
    }
    const int f = (__ballot(bits & 1) ? 1 : 0) | (__ballot(bits & 2) ? 2 : 0);
    if (lane == 0 && f) atomicOr(dec.flags + b, f);
  } else if (i == 0) {
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int64_t e = e0 + 4 * k + r;
      if (e < S) out_d2[e * B + b] = best[r];
    }
  }
}

// valid[b] in {0, 1, 2} merged with the cloud flags; 3 (a configuration
// the FP64 walk decides whatever the FP32 tests say) is kept.  With
// unit_flags, valid[b] is first formed from the unit pass's word as
// valid_units_finish_kernel does (and the word cleared).  The flags are
// cleared for the next cloud or launch (they are zero on entry).
__global__ void __launch_bounds__(256) cloud_merge_kernel(uint8_t* valid, int32_t* flags, int32_t* unit_flags,
                                                          int64_t B) {
  const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= B) return;
  const int f = flags[b];
  flags[b] = 0;
  uint8_t v;
  if (unit_flags) {
    const int u = unit_flags[b];
    unit_flags[b] = 0;
    v = (u & 4) ? 3 : ((u & 1) ? 0 : ((u & 2) ? 2 : 1));
    valid[b] = v;
  } else {
    v = valid[b];
  }
  if (v == 0 || v == 3) return;
  if (f & 1)
    valid[b] = 0;
  else if ((f & 2) || v == 2)
    valid[b] = 2;
}

__global__ void __launch_bounds__(256) fill_i32_kernel(int32_t* p, int32_t v, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// clearance[s * B + b] = min(clearance, sqrt(d2) - rc - r_s)
__global__ void __launch_bounds__(256) cloud_clearance_kernel(float* clearance, const float* d2,
                                                              const float* radius_out, float rc, int64_t B,
                                                              int64_t S) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= S * B) return;
  const float v = sqrtf(d2[q]) - rc - radius_out[q / B];
  clearance[q] = fminf(clearance[q], v);
}

hipError_t launch_cloud_clearance(float* clearance, const float* d2, const float* radius_out, float rc, int64_t B,
                                  int64_t S, hipStream_t st) {
  hipLaunchKernelGGL(cloud_clearance_kernel, dim3((unsigned)((S * B + 255) / 256)), dim3(256), 0, st, clearance, d2,
                     radius_out, rc, B, S);
  return hipGetLastError();
}

// The FP32 validity pass has filled cen / cap2 / rr; the bracket of the
// squared distance is written after rr (the workspace holds 2 Q floats
// there: upper ends, then lower ends).
hipError_t launch_cloud_valid(const CloudTile* tiles, const float* tileB, const CloudSuper* supers,
                              const CloudRange& cr, const float* cen, const float* cap2, const float* rr, int64_t B,
                              int64_t Q, float band, uint8_t* valid, int32_t* flags, int32_t* unit_flags,
                              hipStream_t st) {
  auto blocks = [](int64_t n, int t) { return (unsigned)((n + t - 1) / t); };
  const int64_t waves = (Q / B + 15) / 16 * B;  // 16 spheres of one configuration per wave
  hipLaunchKernelGGL(cloud_min_kernel<true>, dim3(blocks(waves * 64, 256)), dim3(256), 0, st, tiles, tileB, supers,
                     cr.super_begin, cr.n_super, cen, cap2, B, Q, valid, (float*)nullptr,
                     CloudDecide{rr, band, flags, unit_flags});
  hipLaunchKernelGGL(cloud_merge_kernel, dim3(blocks(B, 256)), dim3(256), 0, st, valid, flags, unit_flags, B);
  return hipGetLastError();
}

hipError_t launch_cloud_min(const CloudTile* tiles, const float* tileB, const CloudSuper* supers,
                            const CloudRange& cr, const float* cen, const float* cap2, int64_t B, int64_t Q,
                            float* out_d2, hipStream_t st) {
  const int64_t waves = (Q / B + 15) / 16 * B;  // 16 spheres of one configuration per wave
  hipLaunchKernelGGL(cloud_min_kernel<false>, dim3((unsigned)((waves * 64 + 255) / 256)), dim3(256), 0, st, tiles,
                     tileB, supers, cr.super_begin, cr.n_super, cen, cap2, B, Q, (const uint8_t*)nullptr, out_d2,
                     CloudDecide{nullptr, 0.f, nullptr, nullptr});
  return hipGetLastError();
}

}  // namespace pmp
