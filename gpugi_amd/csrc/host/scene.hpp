// Host-side scene: the content of a .rawscene file plus its material json, the
// SAH sweep builder that produces it, and the reference GPU layout derived
// from it.
#pragma once

#include <cstdint>
#include <array>
#include <string>
#include <vector>

#include "gpugi_rt.h"

namespace gpugi {

// FileDecl::Vertex (bvhmake/filedef.hpp:19-24): 32 bytes.
struct RawVertex {
  float position[3];
  float normal[3];
  float texcoord[2];
};
// FileDecl::Triangle (bvhmake/filedef.hpp:27-39): 16 bytes; invalid slot = all zero.
struct RawTriangle {
  uint32_t vertices[3];
  uint32_t material;
};
// FileDecl::Node (bvhmake/filedef.hpp:64-69): 12 bytes.
struct RawNode {
  uint32_t parent;
  uint32_t firstChild;  // MSB set: the remaining bits index a leaf block
  uint32_t escape;      // next node when this one is done; 0 ends the walk
};
// ε::Box as stored in the "bounding_aabox" chunk: 24 bytes.
struct RawBox {
  float min[3];
  float max[3];
};
struct PointLight {
  float position[3];
  float intensity[3];
};
struct LightTriangle {  // Scene::LightTriangle (scene.hpp:95-100)
  float v0[3], v1[3], v2[3];
  float luminance[3];
};

inline bool IsValidSlot(const RawTriangle& t) { return t.vertices[0] != t.vertices[1]; }

// A material texture as loaded from a file (GPUGI_TEXTURE_* format, rows
// bottom first: row 0 is texture coordinate t = 0.5 / height).
struct HostTexture {
  int width = 0, height = 0;
  int format = GPUGI_TEXTURE_RGBA8_SRGB;
  std::vector<uint8_t> data;  // 4 bytes (RGBA8) or 16 bytes (RGBA32F) per texel
  std::string file;  // path it was read from
  std::string name;  // the material json's string (relative to the scene's directory)
};
// The oriented-box form of a reference-layout hierarchy (8 floats per node):
// 12 floats per node {centre, firstChild}{half sides, escape}{rotationInv
// quaternion x, y, z, w} (TreeNode<OBox>, scene.hpp:39-49).
std::vector<float> FitOBoxHierarchy(const float* hier8, uint32_t numNodes, const int32_t* slots, uint32_t numLeaves,
                                    const float* positions, uint32_t numVertices);

// Reads a TGA, binary PPM or PFM file.  Returns an error message or "".
std::string LoadTextureFile(const std::string& path, HostTexture* out);

class HostScene {
 public:
  std::vector<RawVertex> vertices;
  uint32_t trianglesPerLeaf = 2;
  std::vector<RawTriangle> leafSlots;  // numLeaves * trianglesPerLeaf
  std::vector<RawNode> nodes;
  std::vector<RawBox> boxes;
  std::vector<gpugi_material_params> materials;
  std::vector<PointLight> pointLights;
  // Textures the materials reference (Scene::GetBindlessHandle, scene.cpp:189-203)
  // and per material the texture of each slot (diffuse, opacity, reflectiveness,
  // emissivity; -1: constant).  Empty when no material has a texture.
  std::vector<HostTexture> textures;
  std::vector<std::array<int32_t, 4>> materialTextures;

  // Derived data, rebuilt by Finalize().
  std::vector<float> gpuHierarchy;      // 8 floats per node
  std::vector<int32_t> gpuTriangles;    // 4 ints per slot
  std::vector<float> gpuPositions;      // 3 floats per vertex
  std::vector<float> gpuVertexInfo;     // 4 floats per vertex
  std::vector<gpugi_material> gpuMaterials;
  std::vector<LightTriangle> lightTriangles;
  std::vector<float> lightSummedArea;   // normalized to [0,1]
  std::vector<float> pointLightSummedFlux;
  float lightAreaSum = 0.0f;
  float totalAreaLightFlux = 0.0f;
  float totalPointLightFlux = 0.0f;
  uint32_t numValidTriangles = 0;
  uint32_t maxDepth = 0;
  float bboxMin[3] = {0, 0, 0}, bboxMax[3] = {0, 0, 0};

  uint32_t NumLeaves() const { return trianglesPerLeaf ? uint32_t(leafSlots.size() / trianglesPerLeaf) : 0; }
  // Validates references and derives the GPU layout, material values and light tables.
  // Returns an empty string on success, an error message otherwise.
  std::string Finalize();
  void ComputePointLightTable();
  void Describe(gpugi_scene_desc* out) const;
};

// Scene::LoadMaterial (scene.cpp:132-187) for a constant-valued material.
gpugi_material ResolveMaterial(const gpugi_material_params& p);
// Default parameters of Scene::LoadMaterial when a json key is absent.
gpugi_material_params DefaultMaterialParams(const char* name);
// UNORM8 round trip of an RGB8 texel (the `_hdr == false` textures of scene.cpp:213).
float QuantizeUnorm8(float x);

// SAH sweep build (bvhmake/buildmethods/sweep.cpp:14-274).
std::string BuildSweepBVH(const std::vector<RawVertex>& vertices,
                          const std::vector<RawTriangle>& triangles, uint32_t trianglesPerLeaf,
                          HostScene* out);

// .rawscene + material json I/O.
std::string LoadRawScene(const std::string& path, const std::string& jsonPath, HostScene* out, int* status);
std::string SaveRawScene(const HostScene& scene, const std::string& path, const std::string& jsonPath);

// Generators.
// 20 x 20 x 8 room with `spheres` icospheres (subdivision 3, 1280 triangles each),
// radii log-uniform in [0.05, 0.8]; 60% Lambertian / 40% Phong-glossy materials in
// 240 bins; a 4-triangle ceiling area light and 3 point lights (returned in `lights`).
void MakeClutterMesh(uint32_t spheres, uint32_t seed, std::vector<RawVertex>* v, std::vector<RawTriangle>* t,
                     std::vector<gpugi_material_params>* m, std::vector<PointLight>* lights);
void MakeCornellMesh(std::vector<RawVertex>* v, std::vector<RawTriangle>* t,
                     std::vector<gpugi_material_params>* m);
void MakeHeightfieldMesh(uint32_t cells, uint32_t seed, std::vector<RawVertex>* v,
                         std::vector<RawTriangle>* t, std::vector<gpugi_material_params>* m);

}  // namespace gpugi
