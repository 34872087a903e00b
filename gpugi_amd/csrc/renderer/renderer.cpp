// C++ host layer above the C ABI (renderer/ and scene/ interfaces of gpugi).
#include "renderer.hpp"

#include <algorithm>
#include <cctype>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <limits>

namespace gpugi {

void Check(int status, const gpugi_ctx* ctx, const char* call) {
  if (status == GPUGI_OK) return;
  const char* msg = gpugi_last_error(ctx);
  throw Error(status, std::string(call) + " failed (" + std::to_string(status) + "): " + (msg ? msg : ""));
}

// ---------------------------------------------------------------- Scene

std::shared_ptr<Scene> Scene::Load(const std::string& rawscenePath, const std::string& jsonPath) {
  gpugi_hostscene* s = nullptr;
  Check(gpugi_hostscene_load(rawscenePath.c_str(), jsonPath.empty() ? nullptr : jsonPath.c_str(), &s), nullptr,
        "gpugi_hostscene_load");
  return std::shared_ptr<Scene>(new Scene(s));
}

std::shared_ptr<Scene> Scene::Cornell() {
  gpugi_hostscene* s = nullptr;
  Check(gpugi_hostscene_cornell(&s), nullptr, "gpugi_hostscene_cornell");
  return std::shared_ptr<Scene>(new Scene(s));
}

std::shared_ptr<Scene> Scene::Heightfield(uint32_t cells, uint32_t seed) {
  gpugi_hostscene* s = nullptr;
  Check(gpugi_hostscene_heightfield(cells, seed, &s), nullptr, "gpugi_hostscene_heightfield");
  return std::shared_ptr<Scene>(new Scene(s));
}

Scene::~Scene() { gpugi_hostscene_destroy(m_scene); }

void Scene::AddPointLight(const Vec3& position, const Vec3& intensity) {
  const float p[3] = {position.x, position.y, position.z};
  const float i[3] = {intensity.x, intensity.y, intensity.z};
  Check(gpugi_hostscene_add_point_light(m_scene, p, i), nullptr, "gpugi_hostscene_add_point_light");
}

uint32_t Scene::GetNumTriangles() const {
  uint32_t n = 0, depth = 0;
  float lo[3], hi[3];
  Check(gpugi_hostscene_info(m_scene, &n, &depth, lo, hi), nullptr, "gpugi_hostscene_info");
  return n;
}

void Scene::GetBoundingBox(Vec3* min, Vec3* max) const {
  uint32_t n = 0, depth = 0;
  float lo[3], hi[3];
  Check(gpugi_hostscene_info(m_scene, &n, &depth, lo, hi), nullptr, "gpugi_hostscene_info");
  if (min) *min = Vec3{lo[0], lo[1], lo[2]};
  if (max) *max = Vec3{hi[0], hi[1], hi[2]};
}

// ---------------------------------------------------------------- Camera

Camera::Camera(const Vec3& position, const Vec3& lookat, float aspectRatio, float verticalFOV, const Vec3& up)
    : m_position(position), m_lookat(lookat), m_up(up), m_yfov(verticalFOV), m_aspectRatio(aspectRatio) {}

gpugi_camera Camera::ComputeCameraParams(int width, int height) const {
  const float p[3] = {m_position.x, m_position.y, m_position.z};
  const float l[3] = {m_lookat.x, m_lookat.y, m_lookat.z};
  const float u[3] = {m_up.x, m_up.y, m_up.z};
  gpugi_camera c;
  Check(gpugi_camera_compute(p, l, u, m_yfov, width, height, &c), nullptr, "gpugi_camera_compute");
  return c;
}

// ---------------------------------------------------------------- LightSampler

LightSampler::~LightSampler() { gpugi_lightsampler_destroy(m_sampler); }

void LightSampler::SetScene(std::shared_ptr<const Scene> scene) {
  if (!scene) throw Error(GPUGI_ERR_ARG, "LightSampler::SetScene: null scene");
  if (m_sampler) {
    Check(gpugi_lightsampler_set_scene(m_sampler, scene->Handle()), nullptr, "gpugi_lightsampler_set_scene");
  } else {
    Check(gpugi_lightsampler_create(scene->Handle(), 123, &m_sampler), nullptr, "gpugi_lightsampler_create");
  }
  m_scene = std::move(scene);
}

void LightSampler::GenerateRandomSamples(gpugi_light_sample* destination, unsigned int numSamples,
                                         float positionBias) {
  if (!m_sampler) throw Error(GPUGI_ERR_STATE, "LightSampler::GenerateRandomSamples: no scene");
  Check(gpugi_lightsampler_generate(m_sampler, destination, numSamples, positionBias), nullptr,
        "gpugi_lightsampler_generate");
}

// ---------------------------------------------------------------- renderers

namespace {

// "Every block should work with the same initial light sample": rays per light
// sample, a multiple of the 64-invocation block (lightpathtracer.cpp:34-36,
// bidirectionalpathtracer.cpp:47-49).
unsigned int RaysPerLightSample(int width, int height, unsigned int numLightSamples) {
  const unsigned int block = 64;
  const int numPixels = width * height;
  const unsigned int r = (unsigned int)(numPixels / (int)numLightSamples / (int)block) * block;
  return std::max(block, r);
}

void ApplyCompileDefines(RendererSystem& rs) {
  gpugi_ctx* ctx = rs.GetContext();
  Check(gpugi_set_max_path_length(ctx, rs.GetPathLength()), ctx, "gpugi_set_max_path_length");
  Check(gpugi_set_integrator_kernel(ctx, rs.GetTraversalKernel()), ctx, "gpugi_set_integrator_kernel");
}

}  // namespace

Pathtracer::Pathtracer(RendererSystem& rs) : Renderer(rs) { m_rendererSystem.SetNumInitialLightSamples(128); }

void Pathtracer::SetScene(std::shared_ptr<Scene>) { ApplyCompileDefines(m_rendererSystem); }

void Pathtracer::Draw() {
  gpugi_ctx* ctx = m_rendererSystem.GetContext();
  Check(gpugi_draw(ctx, GPUGI_PT, m_rendererSystem.GetFrameSeed()), ctx, "gpugi_draw(PT)");
}

LightPathtracer::LightPathtracer(RendererSystem& rs) : Renderer(rs) { m_rendererSystem.SetNumInitialLightSamples(256); }

void LightPathtracer::SetScene(std::shared_ptr<Scene>) { ApplyCompileDefines(m_rendererSystem); }

void LightPathtracer::SetScreenSize(int width, int height) {
  m_numRaysPerLightSample = RaysPerLightSample(width, height, m_rendererSystem.GetNumInitialLightSamples());
}

void LightPathtracer::Draw() {
  gpugi_ctx* ctx = m_rendererSystem.GetContext();
  // lighttracer.comp followed by showlightcaches.comp (lightpathtracer.cpp:45-51)
  Check(gpugi_draw(ctx, GPUGI_LT, m_rendererSystem.GetFrameSeed()), ctx, "gpugi_draw(LT)");
}

BidirectionalPathtracer::BidirectionalPathtracer(RendererSystem& rs) : Renderer(rs) {
  m_rendererSystem.SetNumInitialLightSamples(128);
}

void BidirectionalPathtracer::SetScene(std::shared_ptr<Scene>) {
  ApplyCompileDefines(m_rendererSystem);
  // The reference only writes LightCacheCapacity = INT_MAX here and keeps the
  // old cache buffer; the device side drops the cache with the scene, so the
  // capacity is estimated again on the next Draw.
  m_needToDetermineNeededLightCacheCapacity = true;
}

void BidirectionalPathtracer::SetScreenSize(int width, int height) {
  m_numRaysPerLightSample = RaysPerLightSample(width, height, m_rendererSystem.GetNumInitialLightSamples());
  m_needToDetermineNeededLightCacheCapacity = true;
}

void BidirectionalPathtracer::CreateLightCacheWithCapacityEstimate() {
  // bidirectionalpathtracer.cpp:65-132: 8 warm-up light passes, each with the
  // next FrameSeed and light-sample batch, then the iteration count restarts.
  gpugi_ctx* ctx = m_rendererSystem.GetContext();
  const unsigned int numWarmupRuns = 8;
  const unsigned int localSize = 64;
  const unsigned int blockCountPerRun =
      m_rendererSystem.GetNumInitialLightSamples() * m_numRaysPerLightSample / localSize;
  uint64_t numCacheEntries = 0, summedPathLength = 0;
  for (unsigned int run = 0; run < numWarmupRuns; ++run) {
    uint64_t entries = 0, lengthSum = 0;
    Check(gpugi_bpt_warmup(ctx, m_rendererSystem.GetFrameSeed(), &entries, &lengthSum), ctx, "gpugi_bpt_warmup");
    // the reference reads both counters as uint32
    numCacheEntries += (uint32_t)entries;
    summedPathLength += (uint32_t)lengthSum;
    m_rendererSystem.PerIterationBufferUpdate();
  }
  m_rendererSystem.ResetIterationCount();
  m_averageLightPathLength = (int)(summedPathLength / numWarmupRuns / blockCountPerRun / localSize + 1);
  numCacheEntries /= numWarmupRuns;
  m_lightCacheCapacity = (int)(unsigned int)(numCacheEntries + numCacheEntries / 10);
  Check(gpugi_bpt_set_light_cache(ctx, m_lightCacheCapacity, m_averageLightPathLength), ctx,
        "gpugi_bpt_set_light_cache");
}

void BidirectionalPathtracer::Draw() {
  if (m_needToDetermineNeededLightCacheCapacity) {
    CreateLightCacheWithCapacityEstimate();
    m_needToDetermineNeededLightCacheCapacity = false;
  }
  gpugi_ctx* ctx = m_rendererSystem.GetContext();
  // reset counter, light pass filling the cache, eye pass (bidirectionalpathtracer.cpp:182-203)
  Check(gpugi_draw(ctx, GPUGI_BPT, m_rendererSystem.GetFrameSeed()), ctx, "gpugi_draw(BPT)");
}

// ---------------------------------------------------------------- RendererSystem

RendererSystem::RendererSystem(int device) {
  Check(gpugi_create(device, &m_ctx), nullptr, "gpugi_create");
  // renderersystem.cpp:23-37: one per-iteration update, then the count restarts at 0.
  PerIterationBufferUpdate();
  m_iterationCount = 0;
  m_renderTimeMs = 0;
}

RendererSystem::~RendererSystem() {
  m_activeRenderer.reset();
  gpugi_destroy(m_ctx);
}

void RendererSystem::ResetIterationCount() {
  m_iterationCount = 0;
  m_renderTimeMs = 0;
  m_frameSeed = gpugi_wanghash((uint32_t)m_iterationCount);
}

void RendererSystem::PerIterationBufferUpdate(bool iterationIncrement) {
  if (iterationIncrement) ++m_iterationCount;
  m_frameSeed = gpugi_wanghash((uint32_t)m_iterationCount);
  // The reference generates a batch whenever the sample buffer exists; it has
  // no scene to sample before the first SetScene, so neither do we.
  if (m_numInitialLightSamples > 0 && m_lightSampler.HasScene()) {
    m_lightSamples.resize(m_numInitialLightSamples);
    m_lightSampler.GenerateRandomSamples(m_lightSamples.data(), m_numInitialLightSamples);
    Check(gpugi_set_light_samples(m_ctx, m_lightSamples.data(), m_numInitialLightSamples), m_ctx,
          "gpugi_set_light_samples");
  }
}

void RendererSystem::SetNumInitialLightSamples(unsigned int numInitialLightSamples) {
  if (numInitialLightSamples == 0) throw Error(GPUGI_ERR_ARG, "SetNumInitialLightSamples: need at least one sample");
  m_numInitialLightSamples = numInitialLightSamples;
  PerIterationBufferUpdate(false);
}

void RendererSystem::SetScene(std::shared_ptr<Scene> scene) {
  if (!scene) throw Error(GPUGI_ERR_ARG, "RendererSystem::SetScene: null scene");
  gpugi_scene_desc desc;
  Check(gpugi_hostscene_desc(scene->Handle(), &desc), nullptr, "gpugi_hostscene_desc");
  Check(gpugi_load_scene(m_ctx, &desc), m_ctx, "gpugi_load_scene");
  // the material textures the scene's json names (Scene::LoadMaterial, scene.cpp:161-178)
  uint32_t numTextures = 0;
  Check(gpugi_hostscene_num_textures(scene->Handle(), &numTextures), nullptr, "gpugi_hostscene_num_textures");
  if (numTextures > 0) {
    std::vector<int> ids(numTextures);
    for (uint32_t t = 0; t < numTextures; ++t) {
      int w = 0, h = 0, format = 0;
      const void* texels = nullptr;
      Check(gpugi_hostscene_texture(scene->Handle(), t, &w, &h, &format, &texels), nullptr, "gpugi_hostscene_texture");
      Check(gpugi_add_texture(m_ctx, w, h, format, texels, &ids[t]), m_ctx, "gpugi_add_texture");
    }
    for (uint32_t m = 0; m < desc.num_materials; ++m) {
      int32_t slots[4];
      Check(gpugi_hostscene_material_textures(scene->Handle(), m, slots), nullptr, "gpugi_hostscene_material_textures");
      for (int k = 0; k < 4; ++k)
        if (slots[k] >= 0) Check(gpugi_set_material_texture(m_ctx, m, k, ids[slots[k]]), m_ctx, "gpugi_set_material_texture");
    }
  }
  m_scene = scene;
  m_lightSampler.SetScene(m_scene);
  PerIterationBufferUpdate(false);
  m_iterationCount = 0;
  m_renderTimeMs = 0;
  if (m_hasBackbuffer) Check(gpugi_reset_accum(m_ctx), m_ctx, "gpugi_reset_accum");
  if (m_activeRenderer) m_activeRenderer->SetScene(m_scene);
}

void RendererSystem::SetCamera(const Camera& camera) {
  m_iterationCount = 0;
  m_renderTimeMs = 0;
  m_camera = camera;
  m_hasCamera = true;
  if (m_hasBackbuffer) {
    const gpugi_camera c = m_camera.ComputeCameraParams(m_width, m_height);
    Check(gpugi_set_camera(m_ctx, &c), m_ctx, "gpugi_set_camera");  // also clears the backbuffer
  }
  if (m_activeRenderer) m_activeRenderer->SetCamera(camera);
}

void RendererSystem::SetScreenSize(int width, int height) {
  Check(gpugi_set_screen(m_ctx, width, height), m_ctx, "gpugi_set_screen");
  m_width = width;
  m_height = height;
  m_hasBackbuffer = true;
  m_iterationCount = 0;
  m_renderTimeMs = 0;
  // The reference application pairs every resolution change with SetCamera
  // (application.cpp:128-132); here the camera is applied after the new
  // backbuffer exists, so PixelArea uses the new size.
  if (m_hasCamera) {
    const gpugi_camera c = m_camera.ComputeCameraParams(m_width, m_height);
    Check(gpugi_set_camera(m_ctx, &c), m_ctx, "gpugi_set_camera");
  }
  if (m_activeRenderer) m_activeRenderer->SetScreenSize(width, height);
}

void RendererSystem::Draw() {
  if (!m_activeRenderer) return;
  Check(gpugi_sync(m_ctx), m_ctx, "gpugi_sync");
  const auto begin = std::chrono::steady_clock::now();
  m_activeRenderer->Draw();
  PerIterationBufferUpdate();
  Check(gpugi_sync(m_ctx), m_ctx, "gpugi_sync");
  const auto end = std::chrono::steady_clock::now();
  m_renderTimeMs += (uint64_t)std::chrono::duration_cast<std::chrono::milliseconds>(end - begin).count();
}

std::vector<float> RendererSystem::GetBackbuffer() const {
  if (!m_hasBackbuffer) throw Error(GPUGI_ERR_STATE, "RendererSystem::GetBackbuffer: no screen size set");
  std::vector<float> rgba((size_t)m_width * m_height * 4);
  Check(gpugi_readback(m_ctx, rgba.data(), rgba.size() * sizeof(float), nullptr), m_ctx, "gpugi_readback");
  return rgba;
}

void RendererSystem::WritePfm(const std::string& filename) const {
  const std::vector<float> rgba = GetBackbuffer();
  FILE* f = std::fopen(filename.c_str(), "wb");
  if (!f) throw Error(GPUGI_ERR_IO, "WritePfm: cannot open " + filename);
  const uint32_t divisor = (uint32_t)m_iterationCount;
  std::fprintf(f, "PF\n%d %d\n-1.000000\n", m_width, m_height);
  std::vector<float> row((size_t)m_width * 3);
  bool ok = true;
  for (int y = 0; y < m_height && ok; ++y) {
    for (int x = 0; x < m_width; ++x) {
      const float* v = &rgba[((size_t)y * m_width + x) * 4];
      row[3 * x + 0] = v[0] / divisor;
      row[3 * x + 1] = v[1] / divisor;
      row[3 * x + 2] = v[2] / divisor;
    }
    ok = std::fwrite(row.data(), sizeof(float), row.size(), f) == row.size();
  }
  if (std::fclose(f) != 0 || !ok) throw Error(GPUGI_ERR_IO, "WritePfm: write failed for " + filename);
}

// ---------------------------------------------------------------- environment map

namespace {

// Binary PPM (P6): magic, width, height, maxval (whitespace separated, '#'
// comments allowed), one whitespace character, then width * height RGB bytes.
bool LoadPpmRgb8(const std::string& filename, int size, uint8_t* out) {
  std::ifstream f(filename, std::ios::binary);
  if (!f) return false;
  auto token = [&f](std::string* t) {
    t->clear();
    int c;
    while ((c = f.get()) != EOF) {
      if (c == '#') {
        while ((c = f.get()) != EOF && c != '\n') {
        }
      } else if (!std::isspace(c)) {
        break;
      }
    }
    while (c != EOF && !std::isspace(c)) {
      t->push_back((char)c);
      c = f.get();
    }
    return !t->empty();
  };
  std::string magic, w, h, maxval;
  if (!token(&magic) || magic != "P6" || !token(&w) || !token(&h) || !token(&maxval)) return false;
  if (std::atoi(w.c_str()) != size || std::atoi(h.c_str()) != size || std::atoi(maxval.c_str()) != 255) return false;
  f.read(reinterpret_cast<char*>(out), (std::streamsize)size * size * 3);
  return (bool)f;
}

}  // namespace

void RendererSystem::SetEnvironmentMap(int size, const std::string& xneg, const std::string& xpos,
                                       const std::string& yneg, const std::string& ypos, const std::string& zneg,
                                       const std::string& zpos) {
  if (size <= 0) throw Error(GPUGI_ERR_ARG, "SetEnvironmentMap: size must be positive");
  const std::string* files[6] = {&xneg, &xpos, &yneg, &ypos, &zneg, &zpos};
  const size_t faceBytes = (size_t)size * size * 3;
  auto env = std::make_shared<EnvironmentMap>();
  env->size = size;
  env->faces.resize(6 * faceBytes);
  for (int k = 0; k < 6; ++k)
    if (!LoadPpmRgb8(*files[k], size, env->faces.data() + k * faceBytes))
      throw Error(GPUGI_ERR_IO, "SetEnvironmentMap: cannot read a " + std::to_string(size) + "x" +
                                    std::to_string(size) + " binary PPM from " + *files[k]);
  Check(gpugi_set_environment_map(m_ctx, size, env->faces.data()), m_ctx, "gpugi_set_environment_map");
  m_envMap = env;
  if (m_activeRenderer) m_activeRenderer->SetEnvironmentMap(m_envMap);
}

// ---------------------------------------------------------------- MSE check

bool LoadPfm(const std::string& filename, std::vector<float>* rgb, int* width, int* height) {
  std::ifstream f(filename, std::ios::binary);
  if (!f) return false;
  char magic[3] = {0, 0, 0};
  f.read(magic, 3);
  if (!f || magic[0] != 'P' || magic[1] != 'F' || (magic[2] != '\n' && magic[2] != ' ' && magic[2] != '\r'))
    return false;
  long w = 0, h = 0;
  double scale = 0.0;
  f >> w >> h >> scale;
  if (!f || w <= 0 || h <= 0 || scale == 0.0 || w > (1L << 20) || h > (1L << 20)) return false;
  f.get();  // the single whitespace character that ends the header
  std::vector<float> data((size_t)w * h * 3);
  f.read(reinterpret_cast<char*>(data.data()), (std::streamsize)(data.size() * sizeof(float)));
  if (!f) return false;
  if (scale > 0.0) {  // big-endian file
    for (float& v : data) {
      uint32_t u;
      std::memcpy(&u, &v, 4);
      u = (u >> 24) | ((u >> 8) & 0xff00u) | ((u << 8) & 0xff0000u) | (u << 24);
      std::memcpy(&v, &u, 4);
    }
  }
  *rgb = std::move(data);
  *width = (int)w;
  *height = (int)h;
  return true;
}

void TextureMSE::LoadReferenceFromPFM(const std::string& filename) {
  std::vector<float> rgb;
  int w = 0, h = 0;
  if (!LoadPfm(filename, &rgb, &w, &h)) return;
  gpugi_ctx* ctx = m_rs.GetContext();
  Check(gpugi_set_reference_image(ctx, rgb.data(), w, h), ctx, "gpugi_set_reference_image");
  m_hasReference = true;
}

float TextureMSE::ComputeCurrentMSE(unsigned int divisor) {
  if (!m_hasReference) return std::numeric_limits<float>::infinity();
  gpugi_ctx* ctx = m_rs.GetContext();
  float mse = 0.0f;
  Check(gpugi_compute_mse(ctx, divisor, &mse), ctx, "gpugi_compute_mse");
  return mse;
}

}  // namespace gpugi
