// Integrator entry points (device side of gpugi_draw).
#pragma once
#include <map>
#include <hip/hip_runtime.h>

#include <string>

#include "gpugi_rt.h"
#include "shading.h"
#include "trace_ref.h"
#include "wide_bvh.h"

namespace gpugi {

struct DrawParams {
  RefScene ref;
  WideScene wide;
  bool useWide = false;
  const gpugi_material* materials = nullptr;
  const gpugi_light_sample* lights = nullptr;
  uint32_t numLights = 0;
  gpugi_camera camera{};
  int width = 0, height = 0;
  float4* accum = nullptr;
  uint32_t frameSeed = 0;
  int maxPathLength = 16;
  int rrStartDepth = 0;                   // Russian roulette from this bounce on (0: every bounce, as the reference)
  int specificPathLength = -1;            // >= 0: only paths of this many bounces (SHOW_SPECIFIC_PATHLENGTH)
  bool watertight = false;                // watertight triangle test instead of IntersectTriangle
  MaterialTextures tex;                   // material textures (none by default)
  const float4* env = nullptr;            // environment cube map: 6 faces x envSize^2 linear RGB texels, or none (black)
  int envSize = 0;
  uint32_t* counters = nullptr;
  int bptCapacity = -1;                   // LightCacheCapacity (bidirectionalpathtracer.cpp:110-112)
};

// LightCacheEntry (bidirpathtracer/lightcache.glsl:1-16): 64 bytes.
struct LightCacheEntryDev {
  float flux[3];
  int32_t materialIndex;
  float position[3];
  float normal0;
  float texcoord[2];
  float normal1;
  int32_t pathLength;
  float incidentPacked[2];
  float chainedMisN;
  float chainedMisT;
};
static_assert(sizeof(LightCacheEntryDev) == 64, "light cache entry is 64 bytes");

// NumRaysPerLightSample (lightpathtracer.cpp:29-31): every 64-wide block works on one light sample.
inline int RaysPerLightSample(int width, int height, uint32_t numLightSamples) {
  int r = (int)(((uint32_t)(width * height) / numLightSamples / 64u) * 64u);
  return r < 64 ? 64 : r;
}

// Device buffers an integrator keeps between draws (spill stacks, the BPT
// light cache and its capacity estimate).
struct IntegratorState {
  void* spill = nullptr;
  size_t spillBytes = 0;
  void* lightCache = nullptr;
  size_t lightCacheBytes = 0;
  void* scratch = nullptr;
  size_t scratchBytes = 0;
  void* staging = nullptr;          // BPT per-path vertex staging (deterministic append)
  size_t stagingBytes = 0;
  void* pathCounts = nullptr;       // per light path: stored vertices; then exclusive offsets
  size_t pathCountsBytes = 0;
  uint32_t* lastTotal = nullptr;    // device NumLightCacheEntries of the last BPT draw
  uint32_t* queue = nullptr;        // work counter of the persistent integrator kernels
  std::map<const void*, uint32_t> residentBlocks;  // per kernel: resident one-wave blocks on the device
};

// Zeroes st->queue on `stream` and sizes the grid of one-wave blocks for a
// persistent_items kernel: the blocks resident on the device, at most one per
// 64 items (resident), or one per 64 items, so that blocks rather than lanes
// take new work (the light paths, whose splats into the image prefer it).
hipError_t PersistentGrid(IntegratorState* st, const void* kernel, size_t items, uint32_t* blocks,
                          hipStream_t stream, bool resident = true);

// Grows a device buffer to at least `bytes` (contents are not preserved).
inline hipError_t EnsureBuffer(void** buf, size_t* cap, size_t bytes) {
  if (*cap >= bytes) return hipSuccess;
  if (*buf) (void)hipFree(*buf);
  *buf = nullptr;
  *cap = 0;
  hipError_t e = hipMalloc(buf, bytes);
  if (e == hipSuccess) *cap = bytes;
  return e;
}

hipError_t DrawLT(IntegratorState* st, const DrawParams& p, hipStream_t stream);
hipError_t DrawBPT(IntegratorState* st, const DrawParams& p, hipStream_t stream, std::string* err);
// The light pass of a BPT draw (light paths into the cache, bidirectionalpathtracer.cpp:188-194):
// *cache receives the cache, *total the device word with its entry count (both
// null, with *err set, when the cache capacity is not set); p->wide.spill is
// pointed at the integrators' spill stacks.
hipError_t BptLightPass(IntegratorState* st, DrawParams* p, hipStream_t stream, std::string* err,
                        const LightCacheEntryDev** cache, const uint32_t** total);

// Result of a BPT warm-up light pass (the two SSBO counters read back by
// CreateLightCacheWithCapacityEstimate, bidirectionalpathtracer.cpp:75-104).
struct BptWarmupCounts {
  uint64_t numLightCacheEntries;  // NumRaysPerLightSample * N + stored vertices
  uint64_t lightPathLengthSum;    // stored vertices
};
hipError_t BptWarmup(IntegratorState* st, const DrawParams& p, hipStream_t stream, BptWarmupCounts* out);

// ---------------------------------------------------------------- radiance cache

// Query / training records of the cached path tracer.  Feature vectors are the
// 16-float raw cache inputs (mlp.h).
struct CacheParams {
  int query = GPUGI_CACHE_SPREAD;  // where paths query the cache (gpugi_cache_query)
  float spreadC = 0.01f;          // terminate when the path spread exceeds spreadC * a0
  int extraBounces = 2;           // training paths continue this many vertices past termination
  uint32_t trainStride = 32;      // one training path per trainStride pixels
  uint32_t trainOffset = 0;       // which pixel of each stride (varies per frame)
  float bboxMin[3] = {0, 0, 0};
  float bboxInvExtent[3] = {1, 1, 1};
  float* pixelColor = nullptr;    // W*H float4: the short path's own contributions
  float* queryFeat = nullptr;     // W*H x 16: features at the termination vertex
  float* queryWeight = nullptr;   // W*H float4: throughput there (w = 1 when a query exists)
  float* trainVert = nullptr;     // numTrainPaths * maxPathLength x 24: features, T_i, S_i
  uint32_t* trainCount = nullptr; // training vertices per path
  float* trainPath = nullptr;     // numTrainPaths x 8: accumulated C, tail throughput (w = tail valid)
  float* tailFeat = nullptr;      // numTrainPaths x 16
  const float* queryPred = nullptr;  // W*H float4 predictions (resolve)
  const float* tailPred = nullptr;   // numTrainPaths float4 predictions (targets)
};
inline uint32_t CacheTrainPaths(int width, int height, uint32_t stride) {
  return (uint32_t)(((size_t)width * height + stride - 1) / stride);
}
// p[0..n) = value (a kernel, so the call can be recorded into a graph).
hipError_t LaunchFillU32(uint32_t* p, uint32_t value, size_t n, hipStream_t stream);
// PRIMARY_INDIRECT wavefront passes over all W*H pixels (ray-indexed buffers
// in a permuted pixel order): primary rays (2 float4 each)
// and random states; shading of the closest-hit records into pixelColor, the
// cache queries, shadow rays (in place of the primary rays) and their NEE
// contributions; the NEE sum given the any-hit records of the shadow rays.
hipError_t LaunchPrimaryRaygen(const DrawParams& p, float4* rays, uint32_t* seeds, hipStream_t stream);
hipError_t LaunchPrimaryShade(const DrawParams& p, const CacheParams& cp, float4* rays, const int4* hits,
                              const uint32_t* seeds, float4* nee, hipStream_t stream);
hipError_t LaunchPrimaryFinish(const DrawParams& p, const CacheParams& cp, const int4* occluded, const float4* nee,
                               hipStream_t stream);
// train == false: the pixels' paths up to their query vertex (SPREAD only;
// spill for W/8 * H/8 * 64 threads).  train == true: the training paths (spill
// for CacheTrainPaths rounded up to 64 threads).
hipError_t LaunchCachedPaths(const DrawParams& p, const CacheParams& cp, void* spill, bool train,
                             hipStream_t stream);
// accum += pixelColor + queryWeight * max(queryPred, 0).
hipError_t LaunchCacheResolve(const DrawParams& p, const CacheParams& cp, hipStream_t stream);
// Dense training samples (features n x 16, targets n x 4), n = min(*total, maxSamples):
// the vertices at the exclusive offsets `offsets` of the per-path counts, evenly
// thinned to maxSamples and interleaved over `steps` equal training steps.
hipError_t LaunchCacheTargets(const DrawParams& p, const CacheParams& cp, const uint32_t* offsets,
                              const uint32_t* total, uint32_t maxSamples, uint32_t steps, float* feat, float* targets,
                              hipStream_t stream);

void FreeIntegratorState(IntegratorState* s);
void InvalidateIntegratorCaches(IntegratorState* s);
hipError_t DrawIntegrator(IntegratorState* st, int integrator, const DrawParams& p, hipStream_t stream,
                          std::string* err);

}  // namespace gpugi
