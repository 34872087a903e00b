// Path tracer, light tracer and bidirectional path tracer kernels.
//
// Compiled with -ffp-contract=off.  Each kernel restates the GLSL main() it
// replaces with the same random-number draw order, so that with the same
// FrameSeed and light samples a render matches a CPU restatement pixel for
// pixel when the reference traversal is selected.
#include <hip/hip_runtime.h>

#include "integrator_common.h"

namespace gpugi {

// pathtracer.comp:16-110 with path regeneration: persistent one-wave blocks
// take 8x8 tiles (raster order) from an atomic counter and give each lane
// whose path has ended the next pixel of the wave's tile, so a wave no longer
// runs as long as its longest path.  Every pixel's path is the same
// computation as in a one-pixel-per-thread launch (the order in which pixels
// run changes nothing), and each pixel is written by one lane.
__global__ void k_pt_queue_reset(uint32_t* q) {
  if (threadIdx.x == 0) q[0] = 0u;
}

template <bool WIDE>
__global__ __launch_bounds__(kTileThreads) void k_pathtracer(DrawParams p, uint32_t* __restrict__ queue) {
  __shared__ uint2 stk[WIDE ? kWideLdsStack * kTileThreads : 1];
  const int tid = threadIdx.x;
  const uint32_t linear = blockIdx.x * kTileThreads + tid;
  TraceCtx c{&p, stk, tid, WIDE ? p.wide.spill + (size_t)linear * p.wide.spillDepth : nullptr};
  const uint32_t tilesX = (uint32_t)p.width / kTile;
  const uint32_t tiles = tilesX * ((uint32_t)p.height / kTile);
  const uint64_t below = (1ull << tid) - 1ull;
  // wave-uniform: the tile being handed out and how many of its pixels are gone
  uint32_t curTile = 0, curPos = kTileThreads;
  bool exhausted = false, live = false;
  int gx = 0, gy = 0;
  PtPath s;
  for (;;) {
    const uint64_t idle = ~__ballot(live);  // the block is one full wave
    const uint32_t nidle = (uint32_t)__popcll(idle);
    if (!exhausted && nidle > 0u) {
      const uint32_t avail = kTileThreads - curPos;
      uint32_t next = tiles;
      if (avail < nidle) {
        uint32_t v = 0;
        if (tid == 0) v = atomicAdd(queue, 1u);
        v = (uint32_t)__shfl((int)v, 0);
        if (v < tiles) next = v;
      }
      if (!live) {
        const uint32_t rank = (uint32_t)__popcll(idle & below);
        uint32_t tile = tiles, k = 0;
        if (rank < avail) { tile = curTile; k = curPos + rank; }
        else if (next < tiles) { tile = next; k = rank - avail; }
        if (tile < tiles) {
          gx = (int)((tile % tilesX) * kTile + k % kTile);
          gy = (int)((tile / tilesX) * kTile + k / kTile);
          pt_begin(p, s, gx, gy);
          live = true;
        }
      }
      if (avail >= nidle) {
        curPos += nidle;
      } else if (next < tiles) {
        curTile = next;
        curPos = nidle - avail;
      } else {
        exhausted = true;
      }
    }
    if (__ballot(live) == 0) {
      if (exhausted) break;
      continue;
    }
    if (live && pt_bounce<WIDE>(p, c, s)) {
      pt_finish(p, s, gx, gy);
      live = false;
    }
  }
}

void FreeIntegratorState(IntegratorState* s) {
  for (void** b : {&s->spill, &s->lightCache, &s->scratch, &s->staging, &s->pathCounts, (void**)&s->queue})
    if (*b) { (void)hipFree(*b); *b = nullptr; }
  s->ptBlocks[0] = s->ptBlocks[1] = 0;
  s->spillBytes = s->lightCacheBytes = s->scratchBytes = s->stagingBytes = s->pathCountsBytes = 0;
  s->lastTotal = nullptr;
}

void InvalidateIntegratorCaches(IntegratorState* s) { s->lastTotal = nullptr; }

static hipError_t DrawPT(IntegratorState* st, const DrawParams& p0, hipStream_t stream) {
  DrawParams p = p0;
  const size_t tiles = (size_t)(p.width / kTile) * (p.height / kTile);  // truncating, like pathtracer.cpp:50
  if (tiles == 0) return hipSuccess;
  if (!st->queue) {
    hipError_t e = hipMalloc((void**)&st->queue, 256);
    if (e != hipSuccess) return e;
  }
  // persistent grid: the resident one-wave blocks, at most one per tile
  const void* fn = p.useWide ? reinterpret_cast<const void*>(&k_pathtracer<true>)
                             : reinterpret_cast<const void*>(&k_pathtracer<false>);
  if (st->ptBlocks[p.useWide ? 1 : 0] == 0) {
    int perCU = 0, dev = 0, cus = 0;
    hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&perCU, fn, kTileThreads, 0);
    if (e == hipSuccess) e = hipGetDevice(&dev);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    if (e != hipSuccess) return e;
    st->ptBlocks[p.useWide ? 1 : 0] = (uint32_t)std::max(1, perCU * cus);
  }
  const uint32_t blocks = (uint32_t)std::min<size_t>(st->ptBlocks[p.useWide ? 1 : 0], tiles);
  hipLaunchKernelGGL(k_pt_queue_reset, dim3(1), dim3(64), 0, stream, st->queue);
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
