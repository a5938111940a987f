// Device build of a point cloud's tile tables, the structure the cloud
// distance kernels walk in place of the reference's KD-tree
// (collision/kdtree.cpp:40-65, built by CloudSDF's constructor,
// collision/primitive_sdf.hpp:568-583).
//
// Steps, all on the stream given:
//   1. bounds of the points (two-level min / max, exact in FP64);
//   2. a 30-bit Morton code per point from its position in the bounds
//      (10 bits per axis), packed with the point index as a 64-bit key
//      (code << 30 | index), so that sorting the keys orders the points by
//      code and, within a code, by index;
//   3. a bitonic sort of the keys, padded to a power of two (at least
//      kSortChunk) with all-ones keys: stages whose partner distance is below
//      kSortChunk run in LDS, the others over global memory, two stages a
//      pass where both partner distances are at least kSortChunk;
//   4. one wave per tile of kCloudTile consecutive sorted points: FP64 bounds,
//      the FP32 centre and outward-rounded FP32 bounds, the matrix-core B
//      operands (-2 p', |p'|^2) of the points in tile-local coordinates
//      p' = p - o, and the FP64 world points (padding points far away);
//   5. super-tiles of kCloudSuper consecutive tiles under one FP32 AABB.
// Every value is a min / max, a rounding of one FP64 operation or a fixed
// sequence of them (no contraction into FMA), so the tables do not depend
// on the launch shape.
#include <float.h>
#include <stdint.h>

#include "launch.h"

#pragma clang fp contract(off)

namespace pmp {

namespace {

constexpr int kBoundsBlock = 256;
constexpr int kBoundsGrid = 256;  // partial bounds: one record of 6 per workgroup
constexpr int64_t kIndexBits = 30;

// Interleave the low 10 bits of v into every third bit.
__device__ inline uint32_t spread10(uint32_t v) {
  v &= 1023u;
  v = (v | (v << 16)) & 0x030000FFu;
  v = (v | (v << 8)) & 0x0300F00Fu;
  v = (v | (v << 4)) & 0x030C30C3u;
  v = (v | (v << 2)) & 0x09249249u;
  return v;
}

// The neighbouring FP32 value towards -inf / +inf of a finite x.
__device__ inline float next_down(float x) {
  if (x == 0.f) return -FLT_TRUE_MIN;
  const int32_t b = __float_as_int(x);
  return __int_as_float(x > 0.f ? b - 1 : b + 1);
}
__device__ inline float next_up(float x) {
  if (x == 0.f) return FLT_TRUE_MIN;
  const int32_t b = __float_as_int(x);
  return __int_as_float(x > 0.f ? b + 1 : b - 1);
}

__global__ void __launch_bounds__(kBoundsBlock) bounds_partial_kernel(const double* pts, int64_t n, double* part) {
  __shared__ double red[6][kBoundsBlock];
  double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    for (int k = 0; k < 3; ++k) {
      const double v = pts[3 * i + k];
      lo[k] = fmin(lo[k], v);
      hi[k] = fmax(hi[k], v);
    }
  for (int k = 0; k < 3; ++k) {
    red[k][threadIdx.x] = lo[k];
    red[3 + k][threadIdx.x] = hi[k];
  }
  __syncthreads();
  for (int s = kBoundsBlock / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s)
      for (int k = 0; k < 3; ++k) {
        red[k][threadIdx.x] = fmin(red[k][threadIdx.x], red[k][threadIdx.x + s]);
        red[3 + k][threadIdx.x] = fmax(red[3 + k][threadIdx.x], red[3 + k][threadIdx.x + s]);
      }
    __syncthreads();
  }
  if (threadIdx.x < 6) part[6 * blockIdx.x + threadIdx.x] = red[threadIdx.x][0];
}

// One workgroup: the bounds from the partial records, then the keys.
__global__ void __launch_bounds__(kBoundsBlock) bounds_final_kernel(const double* part, int n_part, double* lohi) {
  __shared__ double red[6][kBoundsBlock];
  double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
  for (int b = threadIdx.x; b < n_part; b += blockDim.x)
    for (int k = 0; k < 3; ++k) {
      lo[k] = fmin(lo[k], part[6 * b + k]);
      hi[k] = fmax(hi[k], part[6 * b + 3 + k]);
    }
  for (int k = 0; k < 3; ++k) {
    red[k][threadIdx.x] = lo[k];
    red[3 + k][threadIdx.x] = hi[k];
  }
  __syncthreads();
  for (int s = kBoundsBlock / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s)
      for (int k = 0; k < 3; ++k) {
        red[k][threadIdx.x] = fmin(red[k][threadIdx.x], red[k][threadIdx.x + s]);
        red[3 + k][threadIdx.x] = fmax(red[3 + k][threadIdx.x], red[3 + k][threadIdx.x + s]);
      }
    __syncthreads();
  }
  if (threadIdx.x < 6) lohi[threadIdx.x] = red[threadIdx.x][0];
}

__global__ void __launch_bounds__(256) keys_kernel(const double* pts, int64_t n, int64_t npad, const double* lohi,
                                                   uint64_t* keys) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= npad) return;
  if (i >= n) {
    keys[i] = ~0ull;
    return;
  }
  uint32_t c[3];
  for (int k = 0; k < 3; ++k) {
    const double lo = lohi[k], span = lohi[3 + k] - lo;
    const double u = span > 0 ? (pts[3 * i + k] - lo) / span : 0.0;
    c[k] = (uint32_t)fmin(1023.0, fmax(0.0, u * 1023.0));
  }
  const uint64_t code = spread10(c[0]) | (spread10(c[1]) << 1) | (spread10(c[2]) << 2);
  keys[i] = (code << kIndexBits) | (uint64_t)i;
}

// Compare-exchange of keys i < l in the bitonic network of stage k: the
// block of size k that holds i is sorted ascending when i & k == 0.
__device__ inline void cmpx(uint64_t& a, uint64_t& b, bool ascending) {
  if ((a > b) == ascending) {
    const uint64_t t = a;
    a = b;
    b = t;
  }
}

// Stages k = 2 .. kSortChunk of the network, inside each chunk of
// kSortChunk keys (the whole sort of a chunk).
__global__ void __launch_bounds__(kSortChunk / 2) sort_chunk_kernel(uint64_t* keys) {
  __shared__ uint64_t s[kSortChunk];
  const int64_t base = (int64_t)blockIdx.x * kSortChunk;
  const int t = threadIdx.x;
  s[t] = keys[base + t];
  s[t + kSortChunk / 2] = keys[base + t + kSortChunk / 2];
  __syncthreads();
  for (int k = 2; k <= kSortChunk; k <<= 1)
    for (int j = k >> 1; j > 0; j >>= 1) {
      const int i = 2 * j * (t / j) + (t % j);
      cmpx(s[i], s[i + j], ((base + i) & k) == 0);
      __syncthreads();
    }
  keys[base + t] = s[t];
  keys[base + t + kSortChunk / 2] = s[t + kSortChunk / 2];
}

// Stage k >= 2 kSortChunk, partner distances j = kSortChunk / 2 .. 1, in LDS.
__global__ void __launch_bounds__(kSortChunk / 2) merge_chunk_kernel(uint64_t* keys, int64_t k) {
  __shared__ uint64_t s[kSortChunk];
  const int64_t base = (int64_t)blockIdx.x * kSortChunk;
  const int t = threadIdx.x;
  s[t] = keys[base + t];
  s[t + kSortChunk / 2] = keys[base + t + kSortChunk / 2];
  __syncthreads();
  const bool ascending = (base & k) == 0;  // k > kSortChunk: one direction per chunk
  for (int j = kSortChunk / 2; j > 0; j >>= 1) {
    const int i = 2 * j * (t / j) + (t % j);
    cmpx(s[i], s[i + j], ascending);
    __syncthreads();
  }
  keys[base + t] = s[t];
  keys[base + t + kSortChunk / 2] = s[t + kSortChunk / 2];
}

// Stage k, partner distance j >= kSortChunk: one compare-exchange per thread.
__global__ void __launch_bounds__(256) merge_global_kernel(uint64_t* keys, int64_t npad, int64_t k, int64_t j) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= npad / 2) return;
  const int64_t i = 2 * j * (t / j) + (t % j);
  uint64_t a = keys[i], b = keys[i + j];
  const bool ascending = (i & k) == 0;
  if ((a > b) == ascending) {
    keys[i] = b;
    keys[i + j] = a;
  }
}

// Stages j and j / 2 of stage k (j / 2 >= kSortChunk) in one pass: a thread
// holds the four keys i0, i0 + j / 2, i0 + j, i0 + 3 j / 2 (i0 with both bits
// clear), compares across j, then across j / 2.  All four lie in one block
// of size k, so they share a direction.
__global__ void __launch_bounds__(256) merge_global2_kernel(uint64_t* keys, int64_t npad, int64_t k, int64_t j) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= npad / 4) return;
  const int64_t h = j >> 1;
  const int64_t i0 = 2 * j * (t / h) + (t % h);
  const bool ascending = (i0 & k) == 0;
  uint64_t a = keys[i0], b = keys[i0 + h], c = keys[i0 + j], d = keys[i0 + j + h];
  cmpx(a, c, ascending);
  cmpx(b, d, ascending);
  cmpx(a, b, ascending);
  cmpx(c, d, ascending);
  keys[i0] = a;
  keys[i0 + h] = b;
  keys[i0 + j] = c;
  keys[i0 + j + h] = d;
}

__device__ inline double wave_min(double v) {
  for (int m = 32; m > 0; m >>= 1) v = fmin(v, __shfl_xor(v, m));
  return v;
}
__device__ inline double wave_max(double v) {
  for (int m = 32; m > 0; m >>= 1) v = fmax(v, __shfl_xor(v, m));
  return v;
}

// One wave per tile; lane jj holds point jj of the tile.
__global__ void __launch_bounds__(256) tiles_kernel(const double* pts, int64_t n, const uint64_t* keys, int64_t nt,
                                                    CloudTile* t32, CloudTile64* t64, float* tb, double* p64) {
  const int64_t t = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  if (t >= nt) return;  // wave-uniform
  const int jj = threadIdx.x & 63;
  const int64_t b = t * kCloudTile;
  const bool live = b + jj < n;
  double pw[3] = {0.0, 0.0, 0.0};
  if (live) {
    const int64_t idx = (int64_t)(keys[b + jj] & (((uint64_t)1 << kIndexBits) - 1));
    for (int k = 0; k < 3; ++k) pw[k] = pts[3 * idx + k];
  }
  double lb[3], ub[3], o[3];
  float of[3];
  for (int k = 0; k < 3; ++k) {
    lb[k] = wave_min(live ? pw[k] : 1e300);
    ub[k] = wave_max(live ? pw[k] : -1e300);
    o[k] = 0.5 * (lb[k] + ub[k]);
    of[k] = (float)o[k];
  }
  if (jj == 0) {
    CloudTile T;
    CloudTile64 T64;
    for (int k = 0; k < 3; ++k) {
      T64.lb[k] = lb[k];
      T64.ub[k] = ub[k];
      T.o[k] = of[k];
      T.lb[k] = next_down((float)lb[k]);
      T.ub[k] = next_up((float)ub[k]);
    }
    T.n = (int32_t)((n - b) < kCloudTile ? (n - b) : kCloudTile);
    t32[t] = T;
    t64[t] = T64;
  }
  // B operand [m][k][j] of point 16 m + j, and the FP64 point
  float* col = tb + t * 4 * kCloudTile + (jj / 16) * 64 + (jj % 16);
  double* pp = p64 + (t * kCloudTile + jj) * 4;
  if (live) {
    double pl[3];
    for (int k = 0; k < 3; ++k) pl[k] = (double)(float)(pw[k] - (double)of[k]);
    double n2 = 0.0;
    for (int k = 0; k < 3; ++k) n2 = n2 + pl[k] * pl[k];
    for (int k = 0; k < 3; ++k) col[k * 16] = (float)(-2.0 * pl[k]);
    col[3 * 16] = (float)n2;
    for (int k = 0; k < 3; ++k) pp[k] = pw[k];
  } else {  // padding: never the nearest
    for (int k = 0; k < 3; ++k) col[k * 16] = 0.f;
    col[3 * 16] = 1e30f;
    for (int k = 0; k < 3; ++k) pp[k] = 1e30;
  }
  pp[3] = 0.0;
}

// One thread per super-tile; tiles and supers are indexed within the cloud,
// `first` is written as tile_base + the cloud's tile index.
__global__ void __launch_bounds__(256) supers_kernel(const CloudTile* t32, int64_t nt, int32_t tile_base,
                                                     CloudSuper* sup) {
  const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t a = s * kCloudSuper;
  if (a >= nt) return;
  CloudSuper S;
  for (int k = 0; k < 3; ++k) {
    S.lb[k] = INFINITY;
    S.ub[k] = -INFINITY;
  }
  S.first = tile_base + (int32_t)a;
  S.count = (int32_t)((nt - a) < kCloudSuper ? (nt - a) : kCloudSuper);
  for (int64_t i = a; i < a + S.count; ++i)
    for (int k = 0; k < 3; ++k) {
      S.lb[k] = fminf(S.lb[k], t32[i].lb[k]);
      S.ub[k] = fmaxf(S.ub[k], t32[i].ub[k]);
    }
  sup[s] = S;
}

}  // namespace

CloudBuildSizes cloud_build_sizes(int64_t n) {
  CloudBuildSizes z;
  z.n_tiles = (n + kCloudTile - 1) / kCloudTile;
  z.n_supers = (z.n_tiles + kCloudSuper - 1) / kCloudSuper;
  int64_t npad = kSortChunk;
  while (npad < n) npad <<= 1;
  z.n_keys = npad;
  // keys, then the partial bounds and the bounds
  z.scratch_bytes = npad * (int64_t)sizeof(uint64_t) + (6 * kBoundsGrid + 6) * (int64_t)sizeof(double);
  return z;
}

hipError_t launch_cloud_build(const double* pts, int64_t n, int32_t tile_base, CloudTile* t32, CloudTile64* t64,
                              float* tb, double* p64, CloudSuper* sup, void* scratch, hipStream_t st) {
  const CloudBuildSizes z = cloud_build_sizes(n);
  uint64_t* keys = (uint64_t*)scratch;
  double* part = (double*)(keys + z.n_keys);
  double* lohi = part + 6 * kBoundsGrid;
  int64_t nb = (n + kBoundsBlock - 1) / kBoundsBlock;
  const int grid = (int)(nb < kBoundsGrid ? nb : kBoundsGrid);
  hipLaunchKernelGGL(bounds_partial_kernel, dim3(grid), dim3(kBoundsBlock), 0, st, pts, n, part);
  hipLaunchKernelGGL(bounds_final_kernel, dim3(1), dim3(kBoundsBlock), 0, st, (const double*)part, grid, lohi);
  hipLaunchKernelGGL(keys_kernel, dim3((unsigned)((z.n_keys + 255) / 256)), dim3(256), 0, st, pts, n, z.n_keys,
                     (const double*)lohi, keys);
  const unsigned chunks = (unsigned)(z.n_keys / kSortChunk);
  hipLaunchKernelGGL(sort_chunk_kernel, dim3(chunks), dim3(kSortChunk / 2), 0, st, keys);
  for (int64_t k = 2 * kSortChunk; k <= z.n_keys; k <<= 1) {
    int64_t j = k >> 1;
    for (; j >= 2 * kSortChunk; j >>= 2)
      hipLaunchKernelGGL(merge_global2_kernel, dim3((unsigned)((z.n_keys / 4 + 255) / 256)), dim3(256), 0, st, keys,
                         z.n_keys, k, j);
    if (j >= kSortChunk)
      hipLaunchKernelGGL(merge_global_kernel, dim3((unsigned)((z.n_keys / 2 + 255) / 256)), dim3(256), 0, st, keys,
                         z.n_keys, k, j);
    hipLaunchKernelGGL(merge_chunk_kernel, dim3(chunks), dim3(kSortChunk / 2), 0, st, keys, k);
  }
  hipLaunchKernelGGL(tiles_kernel, dim3((unsigned)((z.n_tiles * 64 + 255) / 256)), dim3(256), 0, st, pts, n,
                     (const uint64_t*)keys, z.n_tiles, t32, t64, tb, p64);
  hipLaunchKernelGGL(supers_kernel, dim3((unsigned)((z.n_supers + 255) / 256)), dim3(256), 0, st,
                     (const CloudTile*)t32, z.n_tiles, tile_base, sup);
  return hipGetLastError();
}

}  // namespace pmp
