// Host-memory stand-ins for the HIP runtime calls and kernel launchers that
// the library's host code (C ABI table builders, host mirror) makes, so that
// this code can run on a CPU under AddressSanitizer / UndefinedBehaviorSanitizer
// (tests/test_host_sanitize.py).  TEST INFRASTRUCTURE ONLY: device memory is
// zeroed host memory, copies are memcpy, kernels do nothing (their outputs
// keep whatever the buffers held), streams and events are tokens.
#include <hip/hip_runtime_api.h>

#include <cstdlib>
#include <cstring>

#include "../../plainmp_amd/csrc/ik.h"
#include "../../plainmp_amd/csrc/launch.h"

extern "C" {
hipError_t hipGetDeviceCount(int* n) {
  *n = 1;
  return hipSuccess;
}
hipError_t hipSetDevice(int) { return hipSuccess; }
hipError_t hipGetDevice(int* d) {
  *d = 0;
  return hipSuccess;
}
hipError_t hipMemcpy(void* dst, const void* src, size_t n, hipMemcpyKind) {
  if (n) std::memmove(dst, src, n);
  return hipSuccess;
}
hipError_t hipGetDevicePropertiesR0600(hipDeviceProp_tR0600* p, int) {
  std::memset(p, 0, sizeof(*p));
  p->multiProcessorCount = 256;
  p->sharedMemPerBlock = 65536;
  std::strcpy(p->name, "host stub");
  return hipSuccess;
}
hipError_t hipDeviceSynchronize(void) { return hipSuccess; }
hipError_t hipMalloc(void** p, size_t n) {
  *p = std::calloc(n ? n : 1, 1);
  return *p ? hipSuccess : hipErrorOutOfMemory;
}
hipError_t hipFree(void* p) {
  std::free(p);
  return hipSuccess;
}
hipError_t hipMemcpyAsync(void* dst, const void* src, size_t n, hipMemcpyKind, hipStream_t) {
  if (n) std::memmove(dst, src, n);
  return hipSuccess;
}
hipError_t hipStreamSynchronize(hipStream_t) { return hipSuccess; }
hipError_t hipStreamCreate(hipStream_t* s) {
  *s = reinterpret_cast<hipStream_t>(std::malloc(16));
  return hipSuccess;
}
hipError_t hipStreamCreateWithFlags(hipStream_t* s, unsigned int) {
  *s = reinterpret_cast<hipStream_t>(std::malloc(16));
  return hipSuccess;
}
hipError_t hipStreamDestroy(hipStream_t s) {
  std::free(s);
  return hipSuccess;
}
hipError_t hipStreamIsCapturing(hipStream_t, hipStreamCaptureStatus* st) {
  *st = hipStreamCaptureStatusNone;
  return hipSuccess;
}
hipError_t hipStreamWaitEvent(hipStream_t, hipEvent_t, unsigned int) { return hipSuccess; }
hipError_t hipStreamBeginCapture(hipStream_t, hipStreamCaptureMode) { return hipSuccess; }
hipError_t hipStreamEndCapture(hipStream_t, hipGraph_t* g) {
  *g = reinterpret_cast<hipGraph_t>(std::malloc(16));
  return hipSuccess;
}
hipError_t hipGraphInstantiate(hipGraphExec_t* e, hipGraph_t, hipGraphNode_t*, char*, size_t) {
  *e = reinterpret_cast<hipGraphExec_t>(std::malloc(16));
  return hipSuccess;
}
hipError_t hipGraphLaunch(hipGraphExec_t, hipStream_t) { return hipSuccess; }
hipError_t hipGraphDestroy(hipGraph_t g) {
  std::free(g);
  return hipSuccess;
}
hipError_t hipGraphExecDestroy(hipGraphExec_t e) {
  std::free(e);
  return hipSuccess;
}
hipError_t hipEventCreate(hipEvent_t* e) {
  *e = reinterpret_cast<hipEvent_t>(std::malloc(16));
  return hipSuccess;
}
hipError_t hipEventCreateWithFlags(hipEvent_t* e, unsigned) { return hipEventCreate(e); }
hipError_t hipEventDestroy(hipEvent_t e) {
  std::free(e);
  return hipSuccess;
}
hipError_t hipEventRecord(hipEvent_t, hipStream_t) { return hipSuccess; }
hipError_t hipEventSynchronize(hipEvent_t) { return hipSuccess; }
hipError_t hipEventElapsedTime(float* ms, hipEvent_t, hipEvent_t) {
  *ms = 0.f;
  return hipSuccess;
}
const char* hipGetErrorString(hipError_t) { return "host stub error"; }
}

namespace pmp {
static void* g_stub_scratch = nullptr;
hipError_t noalloc_scratch_init(int) {
  if (!g_stub_scratch) g_stub_scratch = std::calloc(kNoAllocScratchBytes, 1);
  return hipSuccess;
}
void* noalloc_scratch(int) { return g_stub_scratch; }
CloudBuildSizes cloud_build_sizes(int64_t n) {
  CloudBuildSizes z;
  z.n_tiles = (n + kCloudTile - 1) / kCloudTile;
  z.n_supers = (z.n_tiles + 15) / 16;
  int64_t npad = 512;
  while (npad < n) npad <<= 1;
  z.n_keys = npad;
  z.scratch_bytes = npad * 8 + 4096 * 8;
  return z;
}
int edges_resident_waves_per_cu(bool, int block, size_t) { return 16 * 64 / block * (block / 64); }
hipError_t launch_fill_i32(int32_t* p, int32_t v, int64_t n, hipStream_t) {
  for (int64_t i = 0; i < n; ++i) p[i] = v;
  return hipSuccess;
}
hipError_t launch_centers(const KinArgs<float>&, const double*, int64_t, int64_t, int, float*, hipStream_t) {
  return hipSuccess;
}
hipError_t launch_poses(const KinArgs<double>&, const double*, int64_t, int64_t, int, double*, hipStream_t) {
  return hipSuccess;
}
hipError_t launch_clearance(const KinArgs<float>&, const SceneArgs<float>&, const double*, int64_t, int64_t, int,
                            float*, float*, hipStream_t) {
  return hipSuccess;
}
hipError_t launch_valid(const KinArgs<float>&, const KinArgs<double>&, const SceneArgs<float>&,
                        const SceneArgs<double>&, const double*, int64_t, int64_t B, float, uint8_t* valid, int32_t*,
                        int, const CloudWork&, const CloudRange*, int, int32_t*, int32_t*, hipStream_t) {
  for (int64_t b = 0; b < B; ++b) valid[b] = 1;
  return hipSuccess;
}
hipError_t launch_cloud_valid(const CloudTile*, const float*, const CloudSuper*, const CloudRange&, const float*,
                              const float*, const float*, int64_t, int64_t, float, uint8_t*, int32_t*, int32_t*,
                              hipStream_t) {
  return hipSuccess;
}
hipError_t launch_cloud_min(const CloudTile*, const float*, const CloudSuper*, const CloudRange&, const float*,
                            const float*, int64_t, int64_t, float*, hipStream_t) {
  return hipSuccess;
}
hipError_t launch_cloud_clearance(float*, const float*, const float*, float, int64_t, int64_t, hipStream_t) {
  return hipSuccess;
}
hipError_t launch_edges(const KinArgs<float>&, const KinArgs<double>&, const SceneArgs<float>&,
                        const SceneArgs<double>&, const EdgeArgs& ea, float, int, int, hipStream_t, hipEvent_t,
                        hipEvent_t) {
  for (int64_t e = 0; e < ea.E; ++e) {
    ea.valid[e] = 1;
    if (ea.n_checked) ea.n_checked[e] = 0;
  }
  return hipSuccess;
}
hipError_t launch_eval_units(const KinArgs<float>&, const SceneArgs<float>&, const double*, int64_t, int64_t,
                             const EvalArgs&, hipStream_t) {
  return hipSuccess;
}
hipError_t launch_link_jacobian(const KinArgs<float>&, const double*, int64_t, int64_t, int, const float (&)[3],
                                float*, float*, hipStream_t) {
  return hipSuccess;
}
hipError_t launch_nearest(const double*, int64_t, const double*, int64_t, double*, int32_t*, hipStream_t) {
  return hipSuccess;
}
hipError_t launch_pairs(const KinArgs<float>&, const SceneArgs<float>&, const double*, int64_t, int64_t, float*,
                        float*, int, hipStream_t) {
  return hipSuccess;
}
hipError_t launch_prims_eval(const DevPrim<double>*, int, const double*, const double*, int64_t, int, double,
                             double*, hipStream_t) {
  return hipSuccess;
}
hipError_t launch_sqp_ineq(const SqpArgs&, hipStream_t) { return hipSuccess; }
int vjp_shares(const KinArgs<float>&, int, int64_t, int) { return 1; }
hipError_t launch_vjp(const KinArgs<float>&, const SceneArgs<float>&, const double*, int64_t, int64_t, const float*,
                      float*, float*, int, hipStream_t) {
  return hipSuccess;
}
hipError_t launch_cloud_build(const double*, int64_t, int32_t, CloudTile*, CloudTile64*, float*, double*, CloudSuper*,
                              void*, hipStream_t) {
  return hipSuccess;
}
hipError_t launch_ik_step(const IkArgs&, hipStream_t) { return hipSuccess; }
}  // namespace pmp
