// Scratch of the allocation-free trace (gpugi_trace_noalloc): one file-scope
// device array per GPU, looked up once per device and then handed to the wide
// trace kernel as its batch-queue counters and spill stacks.
//
// Layout: [0, 1 KiB) the per-XCD batch counters (kWideQueueWords), then the
// spill stacks, [depth][thread] over the grid.  The array lives as long as the
// process; every call on a device shares it, so calls that overlap on one GPU
// race (the caller orders them, e.g. on one stream).
#include <hip/hip_runtime.h>

#include <mutex>

#include "launch.h"

namespace gpugi {

static_assert(kNoAllocQueueBytes >= kWideQueueWords * sizeof(uint32_t), "queue counters fit the header");

static __device__ __align__(256) unsigned char scratch[GPUGI_NOALLOC_SCRATCH_BYTES];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratchMutex;
void* g_scratchPtr[kMaxDevices];  // device address of `scratch` on each device (null: not looked up)
}  // namespace

hipError_t NoAllocScratchAddress(int device, void** out) {
  if (device < 0 || device >= kMaxDevices || !out) return hipErrorInvalidValue;
  std::lock_guard<std::mutex> lock(g_scratchMutex);
  if (!g_scratchPtr[device]) {
    int current = 0;
    hipError_t e = hipGetDevice(&current);
    if (e != hipSuccess) return e;
    if (current != device && (e = hipSetDevice(device)) != hipSuccess) return e;
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (current != device) (void)hipSetDevice(current);
    if (e != hipSuccess) return e;
    if (!p) return hipErrorInvalidSymbol;
    g_scratchPtr[device] = p;
  }
  *out = g_scratchPtr[device];
  return hipSuccess;
}

}  // namespace gpugi
