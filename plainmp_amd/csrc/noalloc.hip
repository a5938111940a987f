// Scratch memory of the entry point that allocates nothing
// (pmp_is_valid_batch_noalloc): one statically sized device array, of which
// every GPU the library runs on holds its own copy from module load until the
// process ends.  Host code never uses the array's name as an address: the
// device address of each GPU's copy is looked up once per device, with that
// device current, when a context is created on it (noalloc_scratch_init) and
// kept here.
#include <mutex>

#include "launch.h"

namespace pmp {

static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {
constexpr int kMaxDevices = 64;
void* g_scratch[kMaxDevices] = {};
std::mutex g_scratch_mu;
}  // namespace

hipError_t noalloc_scratch_init(int device) {
  if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch[device]) return hipSuccess;
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess) return e;
  g_scratch[device] = p;
  return hipSuccess;
}

void* noalloc_scratch(int device) {
  if (device < 0 || device >= kMaxDevices) return nullptr;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  return g_scratch[device];
}

}  // namespace pmp
