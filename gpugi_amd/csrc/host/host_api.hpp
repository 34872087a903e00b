// Internal declarations shared by the C-ABI translation units.
#pragma once

#include <string>

#include "gpugi_rt.h"
#include "scene.hpp"

struct gpugi_hostscene {
  gpugi::HostScene scene;
};

struct gpugi_lightsampler {
  const gpugi_hostscene* scene;
  uint64_t state;
};

namespace gpugi {

// Thread-local error of calls without a context.
void SetGlobalError(const std::string& msg);
const char* GlobalError();

void GenerateLightSamples(const HostScene& scene, uint64_t* rngState, gpugi_light_sample* dst,
                          uint32_t numSamples, float positionBias);
void ComputeCamera(const float pos[3], const float lookat[3], const float up[3], float vfovDeg,
                   int width, int height, gpugi_camera* out);

}  // namespace gpugi
