// C ABI of the host-side utilities: scenes, camera, light sampler, RNG.
#include <cstring>
#include <new>

#include "../common/gmath.h"
#include "host_api.hpp"

namespace gpugi {
namespace {
thread_local std::string g_error;
}
void SetGlobalError(const std::string& msg) { g_error = msg; }
const char* GlobalError() { return g_error.c_str(); }
}  // namespace gpugi

using gpugi::SetGlobalError;

#define GPUGI_GUARD_BEGIN try {
#define GPUGI_GUARD_END                                   \
  }                                                       \
  catch (const std::bad_alloc&) {                         \
    SetGlobalError("host allocation failed");             \
    return GPUGI_ERR_NOMEM;                               \
  }                                                       \
  catch (const std::exception& e) {                       \
    SetGlobalError(e.what());                             \
    return GPUGI_ERR_ARG;                                 \
  }                                                       \
  catch (...) {                                           \
    SetGlobalError("unknown exception");                  \
    return GPUGI_ERR_ARG;                                 \
  }

static int FinishScene(gpugi_hostscene* hs, const std::string& buildErr, gpugi_hostscene** out) {
  if (!buildErr.empty()) {
    delete hs;
    SetGlobalError(buildErr);
    return GPUGI_ERR_ARG;
  }
  std::string err = hs->scene.Finalize();
  if (!err.empty()) {
    delete hs;
    SetGlobalError(err);
    return GPUGI_ERR_FORMAT;
  }
  *out = hs;
  return GPUGI_OK;
}

extern "C" {

uint32_t gpugi_wanghash(uint32_t x) { return gg::wanghash(x); }

int gpugi_hostscene_build(const float* positions, const float* normals, const float* texcoords,
                          uint32_t num_vertices, const uint32_t* indices, const uint32_t* material_ids,
                          uint32_t num_triangles, const gpugi_material_params* materials,
                          uint32_t num_materials, uint32_t triangles_per_leaf, gpugi_hostscene** out) {
  GPUGI_GUARD_BEGIN
  if (!out || !positions || !indices || num_vertices == 0 || num_triangles == 0) {
    SetGlobalError("gpugi_hostscene_build: null array or empty mesh");
    return GPUGI_ERR_ARG;
  }
  if (num_materials > 256) {
    SetGlobalError("gpugi_hostscene_build: more than 256 materials");
    return GPUGI_ERR_ARG;
  }
  std::vector<gpugi::RawVertex> verts(num_vertices);
  for (uint32_t i = 0; i < num_vertices; ++i) {
    for (int k = 0; k < 3; ++k) {
      verts[i].position[k] = positions[(size_t)i * 3 + k];
      verts[i].normal[k] = normals ? normals[(size_t)i * 3 + k] : 0.0f;
    }
    verts[i].texcoord[0] = texcoords ? texcoords[(size_t)i * 2] : 0.0f;
    verts[i].texcoord[1] = texcoords ? texcoords[(size_t)i * 2 + 1] : 0.0f;
  }
  std::vector<gpugi::RawTriangle> tris(num_triangles);
  for (uint32_t i = 0; i < num_triangles; ++i) {
    for (int k = 0; k < 3; ++k) tris[i].vertices[k] = indices[(size_t)i * 3 + k];
    tris[i].material = material_ids ? material_ids[i] : 0;
    if (tris[i].material >= (num_materials ? num_materials : 1)) {
      SetGlobalError("gpugi_hostscene_build: material id out of range");
      return GPUGI_ERR_ARG;
    }
  }
  auto* hs = new gpugi_hostscene();
  for (uint32_t i = 0; i < num_materials; ++i) hs->scene.materials.push_back(materials[i]);
  std::string err = gpugi::BuildSweepBVH(verts, tris, triangles_per_leaf, &hs->scene);
  return FinishScene(hs, err, out);
  GPUGI_GUARD_END
}

int gpugi_hostscene_cornell(gpugi_hostscene** out) {
  GPUGI_GUARD_BEGIN
  if (!out) return GPUGI_ERR_ARG;
  std::vector<gpugi::RawVertex> v;
  std::vector<gpugi::RawTriangle> t;
  auto* hs = new gpugi_hostscene();
  gpugi::MakeCornellMesh(&v, &t, &hs->scene.materials);
  std::string err = gpugi::BuildSweepBVH(v, t, 2, &hs->scene);
  return FinishScene(hs, err, out);
  GPUGI_GUARD_END
}

int gpugi_hostscene_heightfield(uint32_t cells, uint32_t seed, gpugi_hostscene** out) {
  GPUGI_GUARD_BEGIN
  if (!out || cells == 0 || cells > 8192) {
    SetGlobalError("gpugi_hostscene_heightfield: cells must be in [1, 8192]");
    return GPUGI_ERR_ARG;
  }
  std::vector<gpugi::RawVertex> v;
  std::vector<gpugi::RawTriangle> t;
  auto* hs = new gpugi_hostscene();
  gpugi::MakeHeightfieldMesh(cells, seed, &v, &t, &hs->scene.materials);
  std::string err = gpugi::BuildSweepBVH(v, t, 2, &hs->scene);
  return FinishScene(hs, err, out);
  GPUGI_GUARD_END
}

int gpugi_hostscene_clutter(uint32_t spheres, uint32_t seed, gpugi_hostscene** out) {
  GPUGI_GUARD_BEGIN
  if (!out || spheres > 100000) {
    SetGlobalError("gpugi_hostscene_clutter: spheres must be in [0, 100000]");
    return GPUGI_ERR_ARG;
  }
  std::vector<gpugi::RawVertex> v;
  std::vector<gpugi::RawTriangle> t;
  std::vector<gpugi::PointLight> lights;
  auto* hs = new gpugi_hostscene();
  gpugi::MakeClutterMesh(spheres, seed, &v, &t, &hs->scene.materials, &lights);
  std::string err = gpugi::BuildSweepBVH(v, t, 2, &hs->scene);
  int rc = FinishScene(hs, err, out);
  if (rc == GPUGI_OK) {
    (*out)->scene.pointLights = lights;
    (*out)->scene.ComputePointLightTable();
  }
  return rc;
  GPUGI_GUARD_END
}

int gpugi_hostscene_load(const char* rawscene_path, const char* json_path, gpugi_hostscene** out) {
  GPUGI_GUARD_BEGIN
  if (!rawscene_path || !out) return GPUGI_ERR_ARG;
  auto* hs = new gpugi_hostscene();
  int status = GPUGI_OK;
  std::string err = gpugi::LoadRawScene(rawscene_path, json_path ? json_path : "", &hs->scene, &status);
  if (!err.empty()) {
    delete hs;
    SetGlobalError(err);
    return status == GPUGI_OK ? GPUGI_ERR_FORMAT : status;
  }
  *out = hs;
  return GPUGI_OK;
  GPUGI_GUARD_END
}

int gpugi_hostscene_save(const gpugi_hostscene* scene, const char* rawscene_path, const char* json_path) {
  GPUGI_GUARD_BEGIN
  if (!scene || !rawscene_path) return GPUGI_ERR_ARG;
  std::string err = gpugi::SaveRawScene(scene->scene, rawscene_path, json_path ? json_path : "");
  if (!err.empty()) {
    SetGlobalError(err);
    return GPUGI_ERR_IO;
  }
  return GPUGI_OK;
  GPUGI_GUARD_END
}

int gpugi_hostscene_desc(const gpugi_hostscene* scene, gpugi_scene_desc* out) {
  if (!scene || !out) {
    SetGlobalError("gpugi_hostscene_desc: null argument");
    return GPUGI_ERR_ARG;
  }
  scene->scene.Describe(out);
  return GPUGI_OK;
}

int gpugi_hostscene_info(const gpugi_hostscene* scene, uint32_t* num_triangles, uint32_t* max_depth,
                         float bbox_min[3], float bbox_max[3]) {
  if (!scene) {
    SetGlobalError("gpugi_hostscene_info: null scene");
    return GPUGI_ERR_ARG;
  }
  if (num_triangles) *num_triangles = scene->scene.numValidTriangles;
  if (max_depth) *max_depth = scene->scene.maxDepth;
  for (int k = 0; k < 3; ++k) {
    if (bbox_min) bbox_min[k] = scene->scene.bboxMin[k];
    if (bbox_max) bbox_max[k] = scene->scene.bboxMax[k];
  }
  return GPUGI_OK;
}

int gpugi_hostscene_obox_hierarchy(const gpugi_hostscene* scene, float* out, size_t num_floats) {
  GPUGI_GUARD_BEGIN
  if (!scene || !out) return GPUGI_ERR_ARG;
  gpugi_scene_desc d;
  scene->scene.Describe(&d);
  if (num_floats < (size_t)d.num_nodes * 12) {
    SetGlobalError("gpugi_hostscene_obox_hierarchy: need num_nodes * 12 floats");
    return GPUGI_ERR_ARG;
  }
  const std::vector<float> h =
      gpugi::FitOBoxHierarchy(d.hierarchy, d.num_nodes, d.triangles, d.num_leaves, d.positions, d.num_vertices);
  std::memcpy(out, h.data(), h.size() * sizeof(float));
  return GPUGI_OK;
  GPUGI_GUARD_END
}

int gpugi_hostscene_num_textures(const gpugi_hostscene* scene, uint32_t* n) {
  if (!scene || !n) {
    SetGlobalError("gpugi_hostscene_num_textures: null argument");
    return GPUGI_ERR_ARG;
  }
  *n = (uint32_t)scene->scene.textures.size();
  return GPUGI_OK;
}

int gpugi_hostscene_texture(const gpugi_hostscene* scene, uint32_t i, int* width, int* height, int* format,
                            const void** texels) {
  if (!scene || !width || !height || !format || !texels) {
    SetGlobalError("gpugi_hostscene_texture: null argument");
    return GPUGI_ERR_ARG;
  }
  if (i >= scene->scene.textures.size()) {
    SetGlobalError("gpugi_hostscene_texture: index out of range");
    return GPUGI_ERR_ARG;
  }
  const gpugi::HostTexture& t = scene->scene.textures[i];
  *width = t.width;
  *height = t.height;
  *format = t.format;
  *texels = t.data.data();
  return GPUGI_OK;
}

int gpugi_hostscene_material_textures(const gpugi_hostscene* scene, uint32_t material, int32_t slots[4]) {
  if (!scene || !slots) {
    SetGlobalError("gpugi_hostscene_material_textures: null argument");
    return GPUGI_ERR_ARG;
  }
  if (material >= scene->scene.materials.size()) {
    SetGlobalError("gpugi_hostscene_material_textures: material out of range");
    return GPUGI_ERR_ARG;
  }
  for (int k = 0; k < 4; ++k)
    slots[k] = material < scene->scene.materialTextures.size() ? scene->scene.materialTextures[material][k] : -1;
  return GPUGI_OK;
}

int gpugi_hostscene_add_point_light(gpugi_hostscene* scene, const float position[3], const float intensity[3]) {
  GPUGI_GUARD_BEGIN
  if (!scene || !position || !intensity) return GPUGI_ERR_ARG;
  gpugi::PointLight pl;
  std::memcpy(pl.position, position, 12);
  std::memcpy(pl.intensity, intensity, 12);
  scene->scene.pointLights.push_back(pl);
  scene->scene.ComputePointLightTable();
  return GPUGI_OK;
  GPUGI_GUARD_END
}

void gpugi_hostscene_destroy(gpugi_hostscene* scene) { delete scene; }

int gpugi_camera_compute(const float position[3], const float lookat[3], const float up[3],
                         float vfov_degrees, int width, int height, gpugi_camera* out) {
  if (!position || !lookat || !up || !out || width <= 0 || height <= 0) {
    SetGlobalError("gpugi_camera_compute: invalid argument");
    return GPUGI_ERR_ARG;
  }
  gpugi::ComputeCamera(position, lookat, up, vfov_degrees, width, height, out);
  return GPUGI_OK;
}

int gpugi_lightsampler_create(const gpugi_hostscene* scene, uint64_t seed, gpugi_lightsampler** out) {
  GPUGI_GUARD_BEGIN
  if (!scene || !out) return GPUGI_ERR_ARG;
  *out = new gpugi_lightsampler{scene, seed};
  return GPUGI_OK;
  GPUGI_GUARD_END
}

int gpugi_lightsampler_generate(gpugi_lightsampler* sampler, gpugi_light_sample* out, uint32_t n,
                                float position_bias) {
  if (!sampler || !out || n == 0) {
    SetGlobalError("gpugi_lightsampler_generate: invalid argument");
    return GPUGI_ERR_ARG;
  }
  gpugi::GenerateLightSamples(sampler->scene->scene, &sampler->state, out, n, position_bias);
  return GPUGI_OK;
}

int gpugi_lightsampler_set_scene(gpugi_lightsampler* sampler, const gpugi_hostscene* scene) {
  if (!sampler || !scene) {
    SetGlobalError("gpugi_lightsampler_set_scene: invalid argument");
    return GPUGI_ERR_ARG;
  }
  sampler->scene = scene;
  return GPUGI_OK;
}

void gpugi_lightsampler_destroy(gpugi_lightsampler* sampler) { delete sampler; }

}  // extern "C"
