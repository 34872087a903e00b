// C++ host layer above the C ABI: gpugi's renderer/ and scene/ interfaces.
//
// A maintainer of gpugi swaps the GL shader path for the HIP path by keeping
// these class names and call sequences: RendererSystem owns the device
// context, the accumulation buffer ("backbuffer"), the per-iteration
// FrameSeed and the light-sample batches; a Renderer subclass issues one
// iteration per Draw().  Everything below talks to the GPU only through
// include/gpugi_rt.h.
//
// Errors: C-ABI status codes become gpugi::Error exceptions carrying the
// status and gpugi_last_error(), as the reference's host code throws
// std::invalid_argument / std::exception (globalconfig.cpp:19-76).
#pragma once

#include <chrono>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "gpugi_rt.h"

namespace gpugi {

class Error : public std::runtime_error {
 public:
  Error(int status, const std::string& what) : std::runtime_error(what), m_status(status) {}
  int Status() const { return m_status; }

 private:
  int m_status;
};

// Throws gpugi::Error when status != GPUGI_OK (message from gpugi_last_error(ctx)).
void Check(int status, const gpugi_ctx* ctx, const char* call);

struct Vec3 {
  float x = 0, y = 0, z = 0;
};

// scene/scene.hpp: geometry, BVH, materials and light sources on the host.
// The device copy is made by RendererSystem::SetScene.
class Scene {
 public:
  // Scene::Scene(file) (scene.cpp:15-55): a .rawscene and its material json.
  static std::shared_ptr<Scene> Load(const std::string& rawscenePath, const std::string& jsonPath);
  // Synthetic scenes: the Cornell-style room and the fBm heightfield.
  static std::shared_ptr<Scene> Cornell();
  static std::shared_ptr<Scene> Heightfield(uint32_t cells, uint32_t seed);
  ~Scene();
  Scene(const Scene&) = delete;
  Scene& operator=(const Scene&) = delete;

  // Scene::AddPointLight (scene.hpp:144).
  void AddPointLight(const Vec3& position, const Vec3& intensity);
  uint32_t GetNumTriangles() const;
  void GetBoundingBox(Vec3* min, Vec3* max) const;
  const gpugi_hostscene* Handle() const { return m_scene; }

 private:
  explicit Scene(gpugi_hostscene* s) : m_scene(s) {}
  gpugi_hostscene* m_scene;
};

// camera/camera.hpp: pinhole camera (position, look-at, up, vertical FOV in degrees).
class Camera {
 public:
  Camera() = default;
  Camera(const Vec3& position, const Vec3& lookat, float aspectRatio, float verticalFOV = 60.0f,
         const Vec3& up = Vec3{0.0f, 1.0f, 0.0f});
  void SetAspectRatio(float a) { m_aspectRatio = a; }
  void SetVFov(float v) { m_yfov = v; }
  void SetPosition(const Vec3& p) { m_position = p; }
  void SetLookAt(const Vec3& l) { m_lookat = l; }
  const Vec3& GetPosition() const { return m_position; }
  const Vec3& GetLookAt() const { return m_lookat; }
  const Vec3& GetUp() const { return m_up; }
  float GetVFov() const { return m_yfov; }
  float GetAspectRatio() const { return m_aspectRatio; }
  // Camera::ComputeCameraParams (camera.cpp:76-85) and the PixelArea of
  // RendererSystem::SetCamera (renderersystem.cpp:208) for a width x height backbuffer.
  gpugi_camera ComputeCameraParams(int width, int height) const;

 private:
  Vec3 m_position{0.0f, 0.0f, 0.0f};
  Vec3 m_lookat{0.0f, 0.0f, 1.0f};
  Vec3 m_up{0.0f, 1.0f, 0.0f};
  float m_yfov = 60.0f;
  float m_aspectRatio = 1.0f;
};

// scene/lightsampler.hpp: its random state (seeded 123, lightsampler.cpp:8)
// survives scene changes, as the reference's member does.
class LightSampler {
 public:
  LightSampler() = default;
  ~LightSampler();
  LightSampler(const LightSampler&) = delete;
  LightSampler& operator=(const LightSampler&) = delete;
  void SetScene(std::shared_ptr<const Scene> scene);
  bool HasScene() const { return m_sampler != nullptr; }
  void GenerateRandomSamples(gpugi_light_sample* destination, unsigned int numSamples,
                             float positionBias = 1.0e-4f);

 private:
  std::shared_ptr<const Scene> m_scene;
  gpugi_lightsampler* m_sampler = nullptr;
};

class RendererSystem;

// The cube map RendererSystem::SetEnvironmentMap loaded (the reference's
// gl::TextureCubemap, renderersystem.cpp:173-186): six size x size sRGB RGB8
// faces in the order x-, x+, y-, y+, z-, z+.  The context holds the decoded
// copy every integrator samples on a miss.
struct EnvironmentMap {
  int size = 0;
  std::vector<uint8_t> faces;
};

// renderer/renderer.hpp:25-50.
class Renderer {
 public:
  virtual ~Renderer() = default;
  virtual std::string GetName() const = 0;
  virtual void SetScene(std::shared_ptr<Scene> /*scene*/) {}
  virtual void SetCamera(const Camera& /*camera*/) {}
  virtual void SetScreenSize(int width, int height) = 0;
  // renderer.hpp:40; called with the map RendererSystem holds (null: none).
  virtual void SetEnvironmentMap(std::shared_ptr<const EnvironmentMap> /*envMap*/) {}
  virtual void Draw() = 0;
  const RendererSystem& GetRendererSystem() const { return m_rendererSystem; }

 protected:
  explicit Renderer(RendererSystem& rs) : m_rendererSystem(rs) {}
  RendererSystem& m_rendererSystem;
};

// renderer/pathtracer.cpp: 128 initial light samples, one pathtracer.comp dispatch per Draw.
class Pathtracer : public Renderer {
 public:
  explicit Pathtracer(RendererSystem& rs);
  std::string GetName() const override { return "Pathtracer"; }
  void SetScene(std::shared_ptr<Scene> scene) override;
  void SetScreenSize(int, int) override {}
  void Draw() override;
};

// renderer/lightpathtracer.cpp: 256 initial light samples; lighttracer.comp + showlightcaches.comp.
class LightPathtracer : public Renderer {
 public:
  explicit LightPathtracer(RendererSystem& rs);
  std::string GetName() const override { return "LightPathtracer"; }
  void SetScene(std::shared_ptr<Scene> scene) override;
  void SetScreenSize(int width, int height) override;
  void Draw() override;
  unsigned int GetNumRaysPerLightSample() const { return m_numRaysPerLightSample; }

 private:
  unsigned int m_numRaysPerLightSample = 0;
};

// renderer/bidirectionalpathtracer.cpp: 128 initial light samples; the light
// cache is sized on the first Draw after a screen (or scene) change by 8
// warm-up light passes.
class BidirectionalPathtracer : public Renderer {
 public:
  explicit BidirectionalPathtracer(RendererSystem& rs);
  std::string GetName() const override { return "BidirectionalPathtracer"; }
  void SetScene(std::shared_ptr<Scene> scene) override;
  void SetScreenSize(int width, int height) override;
  void Draw() override;
  int GetLightCacheCapacity() const { return m_lightCacheCapacity; }
  int GetAverageLightPathLength() const { return m_averageLightPathLength; }

 private:
  void CreateLightCacheWithCapacityEstimate();
  unsigned int m_numRaysPerLightSample = 0;
  bool m_needToDetermineNeededLightCacheCapacity = false;
  int m_lightCacheCapacity = 0;
  int m_averageLightPathLength = 0;
};

// renderer/renderersystem.{hpp,cpp}.
class RendererSystem {
 public:
  explicit RendererSystem(int device = 0);
  ~RendererSystem();
  RendererSystem(const RendererSystem&) = delete;
  RendererSystem& operator=(const RendererSystem&) = delete;

  void SetScene(std::shared_ptr<Scene> scene);
  void SetCamera(const Camera& camera);
  void SetScreenSize(int width, int height);

  // Destroys the old renderer, resets the iteration count, clears the
  // backbuffer and activates a new one (renderersystem.hpp:73-86).
  template <typename RendererType>
  void SetRenderer() {
    m_activeRenderer.reset();
    ResetIterationCount();
    if (m_hasBackbuffer) Check(gpugi_reset_accum(m_ctx), m_ctx, "gpugi_reset_accum");
    m_activeRenderer.reset(new RendererType(*this));
    if (m_hasBackbuffer) m_activeRenderer->SetScreenSize(m_width, m_height);
    m_activeRenderer->SetCamera(m_camera);
    if (m_scene) m_activeRenderer->SetScene(m_scene);
    m_activeRenderer->SetEnvironmentMap(m_envMap);
  }
  Renderer* GetActiveRenderer() { return m_activeRenderer.get(); }

  // One iteration: the renderer's Draw, then PerIterationBufferUpdate.
  void Draw();

  unsigned long GetIterationCount() const { return m_iterationCount; }
  // Wall time of the draws since the last reset, in milliseconds.
  uint64_t GetRenderTime() const { return m_renderTimeMs; }

  void SetNumInitialLightSamples(unsigned int numInitialLightSamples);
  unsigned int GetNumInitialLightSamples() const { return m_numInitialLightSamples; }
  void ResetIterationCount();
  // ++iteration (optional), FrameSeed = WangHash(iteration), next light-sample batch.
  void PerIterationBufferUpdate(bool iterationIncrement = true);

  // GlobalConfig "pathLength" (application.cpp:123, default 16); read when a renderer
  // (re)compiles, i.e. on its SetScene.
  void SetPathLength(int pathLength) { m_pathLength = pathLength; }
  int GetPathLength() const { return m_pathLength; }
  // Traversal the integrators use (GPUGI_KERNEL_*); applied on the next SetScene of a renderer.
  void SetTraversalKernel(int kernel) { m_traversalKernel = kernel; }
  int GetTraversalKernel() const { return m_traversalKernel; }
  // GlobalConfig "bvhType" (application.cpp:135): GPUGI_BVH_AABOX or GPUGI_BVH_OBOX,
  // used by the next SetScene, as the reference reads it when a scene loads.
  void SetBvhType(int bvhType) { Check(gpugi_set_bvh_type(m_ctx, bvhType), m_ctx, "gpugi_set_bvh_type"); }

  // RendererSystem::SetEnvironmentMap (renderersystem.cpp:173-186): six
  // size x size faces in the order x-, x+, y-, y+, z-, z+, here binary PPM
  // (P6, maxval 255) files holding sRGB RGB8 texels, the first file row being
  // face row t = 0.5 / size.  Throws Error(GPUGI_ERR_IO) for a file that cannot
  // be read or has another size.
  void SetEnvironmentMap(int size, const std::string& xneg, const std::string& xpos, const std::string& yneg,
                         const std::string& ypos, const std::string& zneg, const std::string& zpos);

  int GetWidth() const { return m_width; }
  int GetHeight() const { return m_height; }
  uint32_t GetFrameSeed() const { return m_frameSeed; }
  gpugi_ctx* GetContext() const { return m_ctx; }
  const std::shared_ptr<Scene>& GetScene() const { return m_scene; }

  // RGBA32F accumulation buffer (sum over iterations), row-major, bottom row first
  // as GL textures are read.
  std::vector<float> GetBackbuffer() const;
  // Application::SaveImage + WritePfm (application.cpp:394-414, hdrimage.cpp:6-33):
  // RGB of the backbuffer divided by the iteration count.
  void WritePfm(const std::string& filename) const;

 private:
  gpugi_ctx* m_ctx = nullptr;
  std::shared_ptr<Scene> m_scene;
  Camera m_camera;
  LightSampler m_lightSampler;
  std::unique_ptr<Renderer> m_activeRenderer;
  std::shared_ptr<const EnvironmentMap> m_envMap;
  std::vector<gpugi_light_sample> m_lightSamples;
  unsigned long m_iterationCount = 0;
  uint64_t m_renderTimeMs = 0;
  unsigned int m_numInitialLightSamples = 0;
  uint32_t m_frameSeed = 0;
  int m_width = 0, m_height = 0;
  bool m_hasBackbuffer = false;
  bool m_hasCamera = false;
  int m_pathLength = 16;
  int m_traversalKernel = GPUGI_KERNEL_AUTO;
};

// LoadPfm (hdrimage.cpp:35-63): "PF", width and height, the scale line, then
// width*height RGB float triples in file order (a positive scale marks
// big-endian floats, which are swapped).  Returns false, leaving the outputs
// untouched, when the file cannot be read or is not an RGB PFM image.
bool LoadPfm(const std::string& filename, std::vector<float>* rgb, int* width, int* height);

// TextureMSE (imageut/texturemse.hpp): mean squared error of a RendererSystem's
// backbuffer against a reference image, computed on its device.
class TextureMSE {
 public:
  explicit TextureMSE(RendererSystem& rs) : m_rs(rs) {}
  bool HasValidReferenceImage() const { return m_hasReference; }
  // texturemse.cpp:26-37: a file that cannot be loaded changes nothing.
  void LoadReferenceFromPFM(const std::string& filename);
  // texturemse.cpp:39-66: mean over the reference's pixels of
  // |backbuffer / divisor - reference|^2; +inf without a reference image.
  float ComputeCurrentMSE(unsigned int divisor);

 private:
  RendererSystem& m_rs;
  bool m_hasReference = false;
};

}  // namespace gpugi
