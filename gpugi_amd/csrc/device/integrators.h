// Integrator entry points (device side of gpugi_draw).
#pragma once
#include <hip/hip_runtime.h>

#include <string>

#include "gpugi_rt.h"
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
  uint32_t* counters = nullptr;
};

// Device buffers an integrator keeps between draws (spill stacks, the BPT
// light cache and its capacity estimate).
struct IntegratorState {
  void* spill = nullptr;
  size_t spillBytes = 0;
  void* lightCache = nullptr;
  size_t lightCacheBytes = 0;
  void* scratch = nullptr;
  size_t scratchBytes = 0;
  int64_t lightCacheCapacity = -1;  // -1: estimate on the next BPT draw
  int averageLightPathLength = 0;
};

void FreeIntegratorState(IntegratorState* s);
void InvalidateIntegratorCaches(IntegratorState* s);
hipError_t DrawIntegrator(IntegratorState* st, int integrator, const DrawParams& p, hipStream_t stream,
                          std::string* err);

}  // namespace gpugi
