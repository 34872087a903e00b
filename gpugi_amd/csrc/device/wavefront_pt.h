// Wavefront schedule of the path tracer (device buffers and the per-bounce driver).
#pragma once
#include <hip/hip_runtime.h>

#include "integrators.h"
#include "wide_bvh.h"

namespace gpugi {

// Per-pixel path state and per-queue-slot buffers, all in one allocation.
struct WavefrontBuffers {
  uint32_t* counters;  // [0], [1]: lengths of the two path queues
  float4* colorSeed;   // per pixel: color.xyz, w = the random state's bits
  float4* tpFlags;     // per pixel: throughput.xyz, w = bits of bounce << 1 | didNEE (PT)
  float2* misState;    // per pixel (BPT eye paths): last connection probability, last |cos|
  uint32_t* queue[2];  // pixel indices of the live paths
  float4* rays[2];     // their rays (gpugi_ray layout), in queue order
  int4* hits;          // closest-hit records of the current queue
  float4* shadowRays;  // NEE shadow rays of the current queue
  int4* occluded;      // their any-hit records
  float4* neeValue;    // NEE value, w: 1 shadow-tested term, 0.5 zero term, 0 none
  float4* neeWeight;   // throughput and light-sample weight of the term
};

// How the passes trace: the wide kernel (with its launch parameters) or the
// reference walk; passBlocks: the grid of the shade and resolve passes.
// With `side` set, each bounce's shadow trace and resolve pass run on that
// stream, beside the next bounce's closest-hit trace on the main stream; the
// shadow traces then use their own batch counters and spill stacks (wsShadow).
// evShade orders a bounce's shade pass before its shadow trace, evResolve its
// resolve pass before the next shade pass (both write the path's color).
struct WavefrontTrace {
  bool wide = false;
  WideScene ws, wsShadow;
  int stack = 10, refill = 8, wait = 16;
  uint32_t passBlocks = 0;
  hipStream_t side = nullptr;
  hipEvent_t evShade = nullptr, evResolve = nullptr;
};

size_t WavefrontBytes(size_t numPixels);
WavefrontBuffers CarveWavefront(void* base, size_t numPixels);
// One PT draw: maxPathLength bounces of passes enqueued on `stream` (and on
// tr.side, forked from and joined back into `stream`), each reading its queue
// length on the device (no host synchronization, so the draw can be captured
// into a graph).
hipError_t DrawPTWavefront(const DrawParams& p, const WavefrontTrace& tr, const WavefrontBuffers& b,
                           hipStream_t stream);
// The eye pass of a BPT draw (pathtracer_bidir.comp) as the same passes, after
// the light pass filled `cache`; numEntries: device word with the number of
// entries written (at most p.bptCapacity are used), as k_bpt_eye reads it.
hipError_t DrawBPTWavefront(const DrawParams& p, const WavefrontTrace& tr, const WavefrontBuffers& b,
                            const LightCacheEntryDev* cache, const uint32_t* numEntries, hipStream_t stream);

}  // namespace gpugi
