// HostScene: validation, reference GPU layout, materials and light tables.
#include "scene.hpp"

#include <cmath>
#include <cstring>

namespace gpugi {

float QuantizeUnorm8(float x) {
  float c = x < 0.0f ? 0.0f : (x > 1.0f ? 1.0f : x);
  if (!(c == c)) c = 0.0f;
  int q = (int)std::floor(c * 255.0f + 0.5f);
  return (float)q / 255.0f;
}

gpugi_material_params DefaultMaterialParams(const char* name) {
  // Defaults of Scene::LoadMaterial (scene.cpp:152-178).
  gpugi_material_params p;
  std::memset(&p, 0, sizeof(p));
  if (name) std::strncpy(p.name, name, sizeof(p.name) - 1);
  for (int i = 0; i < 3; ++i) {
    p.albedo[i] = 0.5f;
    p.specular_color[i] = 0.03f;
    p.emissivity[i] = 0.0f;
  }
  p.roughness = 1.0f;
  p.opacity = 1.0f;
  p.refraction_index = 1.5f;
  for (int i = 0; i < 4; ++i) p.reflectiveness[i] = 0.0f;
  p.has_reflectiveness = 0;
  return p;
}

gpugi_material ResolveMaterial(const gpugi_material_params& p) {
  // Scene::LoadMaterial (scene.cpp:150-180): fresnel1 = 1 - fresnel0,
  // exponent = 1 / (roughness^2 + 1e-20), RGB8 diffuse/opacity, RGBA32F
  // reflectiveness (1,1,1,exponent) and RGB32F emissivity.
  gpugi_material m;
  std::memset(&m, 0, sizeof(m));
  float exponent = 1.0f / (p.roughness * p.roughness + 1e-20f);
  for (int i = 0; i < 3; ++i) {
    m.diffuse[i] = QuantizeUnorm8(p.albedo[i]);
    m.opacity[i] = QuantizeUnorm8(p.opacity);
    m.reflectiveness[i] = 1.0f;
    m.emissivity[i] = p.emissivity[i];
    m.fresnel0[i] = p.specular_color[i];
    m.fresnel1[i] = 1.0f - p.specular_color[i];
  }
  m.reflectiveness[3] = exponent;
  if (p.has_reflectiveness)  // an RGBA32F texture of the given value
    for (int i = 0; i < 4; ++i) m.reflectiveness[i] = p.reflectiveness[i];
  m.refraction_index_avg = p.refraction_index;
  return m;
}

static float TriangleArea(const float* a, const float* b, const float* c) {
  float e0[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
  float e1[3] = {c[0] - a[0], c[1] - a[1], c[2] - a[2]};
  float x = e0[1] * e1[2] - e0[2] * e1[1];
  float y = e0[2] * e1[0] - e0[0] * e1[2];
  float z = e0[0] * e1[1] - e0[1] * e1[0];
  return 0.5f * std::sqrt(x * x + y * y + z * z);
}

void HostScene::ComputePointLightTable() {
  // Scene::ComputePointLightTable (scene.cpp:295-307): CDF over log(1 + flux).
  totalPointLightFlux = 0.0f;
  pointLightSummedFlux.clear();
  for (const PointLight& l : pointLights) {
    float flux = (0.2126f * l.intensity[0] + 0.7152f * l.intensity[1] + 0.0722f * l.intensity[2]) *
                 4.0f * 3.14159265358979f;
    totalPointLightFlux += std::log(1.0f + flux);
    pointLightSummedFlux.push_back(totalPointLightFlux);
  }
  for (float& f : pointLightSummedFlux) f /= totalPointLightFlux;
}

std::string HostScene::Finalize() {
  if (trianglesPerLeaf == 0) return "trianglesPerLeaf must be positive";
  if (leafSlots.size() % trianglesPerLeaf != 0) return "leaf slot count is not a multiple of the leaf size";
  if (nodes.size() != boxes.size()) return "hierarchy and bounding_aabox sizes differ";
  if (nodes.empty()) return "scene has no hierarchy";
  if (materials.empty()) materials.push_back(DefaultMaterialParams("DefaultMaterial"));
  if (materials.size() > 256) return "more than 256 materials (globaluniforms.glsl:23-27)";
  const uint32_t nv = (uint32_t)vertices.size();
  const uint32_t nn = (uint32_t)nodes.size();
  const uint32_t nl = NumLeaves();
  // Sanity check of every reference (Scene::SanityCheck is declared at
  // scene.hpp:192 but never defined in the reference; this is its role).
  numValidTriangles = 0;
  for (uint32_t l = 0; l < nl; ++l) {
    bool seenInvalid = false;
    for (uint32_t k = 0; k < trianglesPerLeaf; ++k) {
      const RawTriangle& t = leafSlots[l * trianglesPerLeaf + k];
      if (!IsValidSlot(t)) { seenInvalid = true; continue; }
      if (seenInvalid) return "valid triangle after an invalid slot in leaf " + std::to_string(l);
      for (int j = 0; j < 3; ++j)
        if (t.vertices[j] >= nv) return "triangle references vertex out of range";
      if (t.material >= materials.size()) return "triangle references material out of range";
      ++numValidTriangles;
    }
  }
  for (uint32_t i = 0; i < nn; ++i) {
    const RawNode& n = nodes[i];
    if (n.firstChild & 0x80000000u) {
      if ((n.firstChild & 0x7fffffffu) >= nl) return "node references leaf out of range";
    } else if (n.firstChild >= nn || n.firstChild <= i) {
      return "node " + std::to_string(i) + " has an invalid firstChild (must be a later node)";
    }
    if (n.escape >= nn || (n.escape != 0 && n.escape <= i))
      return "node " + std::to_string(i) + " has an invalid escape pointer";
  }
  // Max depth of the tree; nodes are in preorder, so a parent precedes its children.
  {
    std::vector<uint32_t> depth(nn, 1);
    maxDepth = 1;
    for (uint32_t i = 1; i < nn; ++i) {
      if (nodes[i].parent >= i) return "node " + std::to_string(i) + " has an invalid parent";
      depth[i] = depth[nodes[i].parent] + 1;
      if (depth[i] > maxDepth) maxDepth = depth[i];
    }
  }
  // Reference GPU layout (Scene::UploadGeometry / UploadHierarchy).
  gpuHierarchy.resize((size_t)nn * 8);
  for (uint32_t i = 0; i < nn; ++i) {
    float* h = &gpuHierarchy[(size_t)i * 8];
    h[0] = boxes[i].min[0]; h[1] = boxes[i].min[1]; h[2] = boxes[i].min[2];
    std::memcpy(&h[3], &nodes[i].firstChild, 4);
    h[4] = boxes[i].max[0]; h[5] = boxes[i].max[1]; h[6] = boxes[i].max[2];
    std::memcpy(&h[7], &nodes[i].escape, 4);
  }
  gpuTriangles.resize(leafSlots.size() * 4);
  for (size_t s = 0; s < leafSlots.size(); ++s) {
    for (int j = 0; j < 3; ++j) gpuTriangles[s * 4 + j] = (int32_t)leafSlots[s].vertices[j];
    gpuTriangles[s * 4 + 3] = (int32_t)leafSlots[s].material;
  }
  gpuPositions.resize((size_t)nv * 3);
  gpuVertexInfo.resize((size_t)nv * 4);
  for (uint32_t v = 0; v < nv; ++v) {
    const RawVertex& rv = vertices[v];
    for (int j = 0; j < 3; ++j) gpuPositions[(size_t)v * 3 + j] = rv.position[j];
    // VertexInfo = {atan2(n.y, n.x), n.z, u, v} (scene.cpp:75-83)
    gpuVertexInfo[(size_t)v * 4 + 0] = std::atan2(rv.normal[1], rv.normal[0]);
    gpuVertexInfo[(size_t)v * 4 + 1] = rv.normal[2];
    gpuVertexInfo[(size_t)v * 4 + 2] = rv.texcoord[0];
    gpuVertexInfo[(size_t)v * 4 + 3] = rv.texcoord[1];
  }
  gpuMaterials.clear();
  for (const auto& p : materials) gpuMaterials.push_back(ResolveMaterial(p));
  for (int j = 0; j < 3; ++j) { bboxMin[j] = boxes[0].min[j]; bboxMax[j] = boxes[0].max[j]; }

  // Scene::LoadLightSources (scene.cpp:244-293), over slots in leaf order.
  lightTriangles.clear();
  lightSummedArea.clear();
  totalAreaLightFlux = 0.0f;
  float sum = 0.0f;
  for (const RawTriangle& t : leafSlots) {
    if (!IsValidSlot(t)) continue;
    const float* e = materials[t.material].emissivity;
    if (e[0] == 0.0f && e[1] == 0.0f && e[2] == 0.0f) continue;
    LightTriangle lt;
    std::memcpy(lt.v0, vertices[t.vertices[0]].position, 12);
    std::memcpy(lt.v1, vertices[t.vertices[1]].position, 12);
    std::memcpy(lt.v2, vertices[t.vertices[2]].position, 12);
    std::memcpy(lt.luminance, e, 12);
    lightTriangles.push_back(lt);
    float area = TriangleArea(lt.v0, lt.v1, lt.v2);
    totalAreaLightFlux += (0.2126f * e[0] + 0.7152f * e[1] + 0.0722f * e[2]) * area * 3.14159265358979f;
    sum += area;
    lightSummedArea.push_back(sum);
  }
  lightAreaSum = 0.0f;
  if (!lightSummedArea.empty()) {
    lightAreaSum = lightSummedArea.back();
    for (float& a : lightSummedArea) a /= lightAreaSum;
  }
  ComputePointLightTable();
  return std::string();
}

void HostScene::Describe(gpugi_scene_desc* out) const {
  out->hierarchy = gpuHierarchy.data();
  out->num_nodes = (uint32_t)nodes.size();
  out->triangles = gpuTriangles.data();
  out->num_leaves = NumLeaves();
  out->triangles_per_leaf = trianglesPerLeaf;
  out->positions = gpuPositions.data();
  out->vertex_info = gpuVertexInfo.data();
  out->num_vertices = (uint32_t)vertices.size();
  out->materials = gpuMaterials.data();
  out->num_materials = (uint32_t)gpuMaterials.size();
}

}  // namespace gpugi
