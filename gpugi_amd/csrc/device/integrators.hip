// Path tracer, light tracer and bidirectional path tracer kernels.
//
// Compiled with -ffp-contract=off.  Each kernel restates the GLSL main() it
// replaces with the same random-number draw order, so that with the same
// FrameSeed and light samples a render matches a CPU restatement pixel for
// pixel when the reference traversal is selected.
#include <hip/hip_runtime.h>

#include "integrator_common.h"

namespace gpugi {

// pathtracer.comp:16-110 with path regeneration (persistent_items over the
// pixels of all 8x8 tiles): a wave no longer runs as long as its longest
// path.  Every pixel's path is the same computation as in a
// one-pixel-per-thread launch, and each pixel is written by one lane.
__global__ void k_queue_reset(uint32_t* q) {
  if (threadIdx.x == 0) q[0] = 0u;
}

// Five waves per SIMD (at most 96 VGPRs, a 16-byte spill in the wide variant)
// instead of the four that 115 VGPRs allow: clutter room -4.5 % per iteration.
template <bool WIDE>
__global__ __launch_bounds__(kTileThreads) __attribute__((amdgpu_waves_per_eu(5))) void k_pathtracer(
    DrawParams p, uint32_t* __restrict__ queue) {
  __shared__ uint2 stk[WIDE ? kWideLdsStack * kTileThreads : 1];
  const int tid = threadIdx.x;
  const uint32_t linear = blockIdx.x * kTileThreads + tid;
  TraceCtx c{&p, stk, tid, WIDE ? p.wide.spill + (size_t)linear * p.wide.spillDepth : nullptr};
  const uint32_t pixels = ((uint32_t)p.width / kTile) * ((uint32_t)p.height / kTile) * kTileThreads;
  int gx = 0, gy = 0;
  PtPath s;
  persistent_items(
      pixels, queue,
      [&](uint32_t item) {
        tile_pixel(item, p.width, gx, gy);
        pt_begin(p, s, gx, gy, p.frameSeed);
      },
      [&]() { return pt_bounce<WIDE>(p, c, s); }, [&]() { pt_finish(p, s, gx, gy); });
}

void FreeIntegratorState(IntegratorState* s) {
  for (void** b : {&s->spill, &s->lightCache, &s->scratch, &s->staging, &s->pathCounts, (void**)&s->queue})
    if (*b) { (void)hipFree(*b); *b = nullptr; }
  s->residentBlocks.clear();
  s->spillBytes = s->lightCacheBytes = s->scratchBytes = s->stagingBytes = s->pathCountsBytes = 0;
  s->lastTotal = nullptr;
}

void InvalidateIntegratorCaches(IntegratorState* s) { s->lastTotal = nullptr; }

hipError_t PersistentGrid(IntegratorState* st, const void* kernel, size_t items, uint32_t* blocks,
                          hipStream_t stream, bool resident) {
  if (!st->queue) {
    hipError_t e = hipMalloc((void**)&st->queue, 256);
    if (e != hipSuccess) return e;
  }
  auto it = st->residentBlocks.find(kernel);
  if (it == st->residentBlocks.end()) {
    int perCU = 0, dev = 0, cus = 0;
    hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&perCU, kernel, kTileThreads, 0);
    if (e == hipSuccess) e = hipGetDevice(&dev);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    if (e != hipSuccess) return e;
    it = st->residentBlocks.emplace(kernel, (uint32_t)std::max(1, perCU * cus)).first;
  }
  const size_t chunks = (items + kTileThreads - 1) / kTileThreads;
  *blocks = (uint32_t)(resident ? std::min<size_t>(it->second, chunks) : chunks);
  hipLaunchKernelGGL(k_queue_reset, dim3(1), dim3(64), 0, stream, st->queue);
  return hipGetLastError();
}

static hipError_t DrawPT(IntegratorState* st, const DrawParams& p0, hipStream_t stream) {
  DrawParams p = p0;
  const size_t tiles = (size_t)(p.width / kTile) * (p.height / kTile);  // truncating, like pathtracer.cpp:50
  if (tiles == 0) return hipSuccess;
  const void* fn = p.useWide ? reinterpret_cast<const void*>(&k_pathtracer<true>)
                             : reinterpret_cast<const void*>(&k_pathtracer<false>);
  uint32_t blocks = 0;
  if (hipError_t e = PersistentGrid(st, fn, tiles * kTileThreads, &blocks, stream)) return e;
  if (p.useWide) {
    const size_t threads = (size_t)blocks * kTileThreads;
    const size_t bytes = std::max<size_t>(8, threads * p.wide.spillDepth * sizeof(uint2));
    hipError_t e = EnsureBuffer(&st->spill, &st->spillBytes, bytes);
    if (e != hipSuccess) return e;
    p.wide.spill = (uint2*)st->spill;
    hipLaunchKernelGGL(k_pathtracer<true>, dim3(blocks), dim3(kTileThreads), 0, stream, p, st->queue);
  } else {
    hipLaunchKernelGGL(k_pathtracer<false>, dim3(blocks), dim3(kTileThreads), 0, stream, p, st->queue);
  }
  return hipGetLastError();
}

hipError_t DrawIntegrator(IntegratorState* st, int integrator, const DrawParams& p, hipStream_t stream,
                          std::string* err) {
  switch (integrator) {
    case GPUGI_PT:
      return DrawPT(st, p, stream);
    case GPUGI_LT:
      return DrawLT(st, p, stream);
    case GPUGI_BPT:
      return DrawBPT(st, p, stream, err);
    default:
      *err = "integrator not available in this build";
      return hipSuccess;
  }
}

}  // namespace gpugi
