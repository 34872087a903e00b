// Host light sampler and camera parameters.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "../common/gmath.h"
#include "scene.hpp"
#include "host_api.hpp"

namespace gpugi {

void GenerateLightSamples(const HostScene& scene, uint64_t* rngState, gpugi_light_sample* dst,
                          uint32_t numSamples, float positionBias) {
  // LightSampler::GenerateRandomSamples (lightsampler.cpp:18-95).
  if (scene.totalAreaLightFlux + scene.totalPointLightFlux == 0.0f) {
    // No light at all: black dummy samples (lightsampler.cpp:21-31).
    for (uint32_t i = 0; i < numSamples; ++i) std::memset(&dst[i], 0, sizeof(gpugi_light_sample));
    return;
  }
  const float pointLightPercentage =
      scene.totalPointLightFlux / (scene.totalAreaLightFlux + scene.totalPointLightFlux);
  const uint32_t numPointSamples = (uint32_t)(pointLightPercentage * (float)numSamples);
  const uint32_t numAreaSamples = numSamples - numPointSamples;
  float sampleWeight = scene.lightAreaSum / (float)numAreaSamples;
  const uint32_t nt = (uint32_t)scene.lightTriangles.size();
  for (uint32_t s = 0; s < numAreaSamples; ++s, ++dst) {
    float randomTriangle = gg::xorshiftf(*rngState);
    const float* it = std::lower_bound(scene.lightSummedArea.data(), scene.lightSummedArea.data() + nt, randomTriangle);
    uint32_t ti = (uint32_t)(it - scene.lightSummedArea.data());
    if (ti >= nt) ti = nt - 1;  // guards a table whose last entry rounds below 1
    // Uniform barycentrics (Osada et al., "Shape distributions", 2002, sec. 4.2).
    float xi0 = gg::xorshiftf(*rngState);
    float xi1 = gg::xorshiftf(*rngState);
    xi0 = std::sqrt(xi0);
    float alpha = 1.0f - xi0;
    float beta = xi0 * (1.0f - xi1);
    float gamma = xi0 * xi1;
    const LightTriangle& lt = scene.lightTriangles[ti];
    float p[3], a[3], b[3], n[3];
    for (int k = 0; k < 3; ++k) {
      p[k] = lt.v0[k] * alpha + lt.v1[k] * beta + lt.v2[k] * gamma;
      a[k] = lt.v0[k] - lt.v1[k];
      b[k] = lt.v0[k] - lt.v2[k];
    }
    n[0] = a[1] * b[2] - a[2] * b[1];
    n[1] = a[2] * b[0] - a[0] * b[2];
    n[2] = a[0] * b[1] - a[1] * b[0];
    float len = std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
    for (int k = 0; k < 3; ++k) {
      n[k] /= len;
      dst->position[k] = p[k] + n[k] * positionBias;
      dst->intensity[k] = lt.luminance[k] * sampleWeight;
    }
    dst->normal_phi = std::atan2(n[1], n[0]);
    dst->normal_theta_cos = n[2];
  }
  const uint32_t np = (uint32_t)scene.pointLights.size();
  for (uint32_t s = 0; s < numPointSamples; ++s, ++dst) {
    float r = gg::xorshiftf(*rngState);
    const float* it = std::lower_bound(scene.pointLightSummedFlux.data(), scene.pointLightSummedFlux.data() + np, r);
    uint32_t li = (uint32_t)(it - scene.pointLightSummedFlux.data());
    if (li >= np) li = np - 1;
    sampleWeight = scene.pointLightSummedFlux[li];
    if (li > 0) sampleWeight -= scene.pointLightSummedFlux[li - 1];
    const PointLight& pl = scene.pointLights[li];
    for (int k = 0; k < 3; ++k) {
      dst->position[k] = pl.position[k];
      dst->intensity[k] = pl.intensity[k] / ((float)numPointSamples * sampleWeight);
    }
    dst->normal_phi = 0.0f;
    dst->normal_theta_cos = 2.0f;  // "wrong" normal marks a point light (lightsource.glsl:25)
  }
}

void ComputeCamera(const float pos[3], const float lookat[3], const float up[3], float vfovDeg,
                   int width, int height, gpugi_camera* out) {
  // Camera::ComputeCameraParams (camera.cpp:76-85) and the PixelArea of
  // RendererSystem::SetCamera (renderersystem.cpp:208).
  auto normalize = [](float* v) {
    float l = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    v[0] /= l; v[1] /= l; v[2] /= l;
  };
  auto cross = [](const float* a, const float* b, float* r) {
    r[0] = a[1] * b[2] - a[2] * b[1];
    r[1] = a[2] * b[0] - a[0] * b[2];
    r[2] = a[0] * b[1] - a[1] * b[0];
  };
  float w[3] = {lookat[0] - pos[0], lookat[1] - pos[1], lookat[2] - pos[2]};
  normalize(w);
  float u[3], v[3];
  cross(w, up, u);
  normalize(u);
  cross(u, w, v);
  normalize(v);
  const float aspect = (float)width / (float)height;
  const float f = std::tan(vfovDeg / 2.0f * (3.14159265358979f / 180.0f));
  for (int k = 0; k < 3; ++k) {
    out->position[k] = pos[k];
    out->u[k] = u[k] * (f * aspect);
    out->v[k] = v[k] * f;
    out->w[k] = w[k];
  }
  float lu = std::sqrt(out->u[0] * out->u[0] + out->u[1] * out->u[1] + out->u[2] * out->u[2]);
  float lv = std::sqrt(out->v[0] * out->v[0] + out->v[1] * out->v[1] + out->v[2] * out->v[2]);
  out->pixel_area = lu * lv * 4.0f / (float)(width * height);
}

}  // namespace gpugi
