// Path tracer, light tracer and bidirectional path tracer kernels.
//
// Compiled with -ffp-contract=off.  Each kernel restates the GLSL main() it
// replaces with the same random-number draw order, so that with the same
// FrameSeed and light samples a render matches a CPU restatement pixel for
// pixel when the reference traversal is selected.
#include <hip/hip_runtime.h>

#include "integrator_common.h"

namespace gpugi {

// pathtracer.comp:16-110 for one 8x8 tile per block (the per-pixel body is pt_pixel).
template <bool WIDE>
__global__ __launch_bounds__(kTileThreads) void k_pathtracer(DrawParams p) {
  __shared__ uint2 stk[WIDE ? kWideLdsStack * kTileThreads : 1];
  const int tid = threadIdx.x;
  const int gx = blockIdx.x * kTile + (tid % kTile);
  const int gy = blockIdx.y * kTile + (tid / kTile);
  const uint32_t linear = (blockIdx.y * gridDim.x + blockIdx.x) * kTileThreads + tid;
  TraceCtx c{&p, stk, tid, WIDE ? p.wide.spill + (size_t)linear * p.wide.spillDepth : nullptr};
  pt_pixel<WIDE>(p, c, gx, gy, nullptr);
}

void FreeIntegratorState(IntegratorState* s) {
  for (void** b : {&s->spill, &s->lightCache, &s->scratch, &s->staging, &s->pathCounts})
    if (*b) { (void)hipFree(*b); *b = nullptr; }
  s->spillBytes = s->lightCacheBytes = s->scratchBytes = s->stagingBytes = s->pathCountsBytes = 0;
  s->lastTotal = nullptr;
}

void InvalidateIntegratorCaches(IntegratorState* s) { s->lastTotal = nullptr; }

static hipError_t DrawPT(IntegratorState* st, const DrawParams& p0, hipStream_t stream) {
  DrawParams p = p0;
  dim3 grid(p.width / kTile, p.height / kTile);  // truncating, like pathtracer.cpp:50
  if (grid.x == 0 || grid.y == 0) return hipSuccess;
  if (p.useWide) {
    const size_t threads = (size_t)grid.x * grid.y * kTileThreads;
    const size_t bytes = std::max<size_t>(8, threads * p.wide.spillDepth * sizeof(uint2));
    hipError_t e = EnsureBuffer(&st->spill, &st->spillBytes, bytes);
    if (e != hipSuccess) return e;
    p.wide.spill = (uint2*)st->spill;
    hipLaunchKernelGGL(k_pathtracer<true>, grid, dim3(kTileThreads), 0, stream, p);
  } else {
    hipLaunchKernelGGL(k_pathtracer<false>, grid, dim3(kTileThreads), 0, stream, p);
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
