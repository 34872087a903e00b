// .rawscene reader/writer (documentation/fileformat.tex:37-168) and the
// material json next to it (the bim-era material files in scenes/*/*.json and
// the keys Scene::LoadMaterial reads, scene.cpp:138-180).
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <map>
#include <memory>
#include <sstream>

#include "scene.hpp"

namespace gpugi {
namespace {

// FileDecl::NamedArray (bvhmake/filedef.hpp:11-16): 40 bytes.
struct NamedArray {
  char name[32];
  uint32_t numElements;
  uint32_t elementSize;
};
static_assert(sizeof(NamedArray) == 40, "chunk header is 40 bytes");
static_assert(sizeof(RawVertex) == 32, "vertex is 32 bytes");
static_assert(sizeof(RawTriangle) == 16, "triangle is 16 bytes");
static_assert(sizeof(RawNode) == 12, "node is 12 bytes");
static_assert(sizeof(RawBox) == 24, "box is 24 bytes");

// ---- a small JSON reader: objects, arrays, numbers, strings, literals ----
struct JValue {
  enum Kind { kNull, kBool, kNumber, kString, kArray, kObject } kind = kNull;
  double number = 0;
  bool boolean = false;
  std::string str;
  std::vector<JValue> items;
  std::vector<std::pair<std::string, JValue>> members;
  const JValue* Get(const std::string& key) const {
    for (const auto& m : members)
      if (m.first == key) return &m.second;
    return nullptr;
  }
};

class JParser {
 public:
  explicit JParser(const std::string& s) : s_(s) {}
  bool Parse(JValue* out, std::string* err) {
    pos_ = 0;
    SkipWs();
    if (!Value(out, 0)) { *err = err_; return false; }
    SkipWs();
    if (pos_ != s_.size()) { *err = "trailing characters in json"; return false; }
    return true;
  }

 private:
  const std::string& s_;
  size_t pos_ = 0;
  std::string err_;
  void SkipWs() {
    for (;;) {
      while (pos_ < s_.size() && std::isspace((unsigned char)s_[pos_])) ++pos_;
      // tolerate // line comments, as hand-edited scene files carry them
      if (pos_ + 1 < s_.size() && s_[pos_] == '/' && s_[pos_ + 1] == '/') {
        while (pos_ < s_.size() && s_[pos_] != '\n') ++pos_;
        continue;
      }
      break;
    }
  }
  bool Fail(const char* m) { err_ = std::string("json: ") + m + " at offset " + std::to_string(pos_); return false; }
  bool String(std::string* out) {
    if (pos_ >= s_.size() || s_[pos_] != '"') return Fail("expected string");
    ++pos_;
    out->clear();
    while (pos_ < s_.size() && s_[pos_] != '"') {
      char c = s_[pos_++];
      if (c == '\\' && pos_ < s_.size()) {
        char e = s_[pos_++];
        switch (e) {
          case 'n': c = '\n'; break;
          case 't': c = '\t'; break;
          case 'r': c = '\r'; break;
          case 'b': c = '\b'; break;
          case 'f': c = '\f'; break;
          case 'u': c = '?'; pos_ = std::min(s_.size(), pos_ + 4); break;
          default: c = e; break;
        }
      }
      out->push_back(c);
    }
    if (pos_ >= s_.size()) return Fail("unterminated string");
    ++pos_;
    return true;
  }
  bool Value(JValue* v, int depth) {
    if (depth > 64) return Fail("nesting too deep");
    SkipWs();
    if (pos_ >= s_.size()) return Fail("unexpected end");
    char c = s_[pos_];
    if (c == '{') {
      v->kind = JValue::kObject;
      ++pos_;
      SkipWs();
      if (pos_ < s_.size() && s_[pos_] == '}') { ++pos_; return true; }
      for (;;) {
        SkipWs();
        std::string key;
        if (!String(&key)) return false;
        SkipWs();
        if (pos_ >= s_.size() || s_[pos_] != ':') return Fail("expected ':'");
        ++pos_;
        JValue child;
        if (!Value(&child, depth + 1)) return false;
        v->members.emplace_back(key, std::move(child));
        SkipWs();
        if (pos_ < s_.size() && s_[pos_] == ',') { ++pos_; continue; }
        if (pos_ < s_.size() && s_[pos_] == '}') { ++pos_; return true; }
        return Fail("expected ',' or '}'");
      }
    }
    if (c == '[') {
      v->kind = JValue::kArray;
      ++pos_;
      SkipWs();
      if (pos_ < s_.size() && s_[pos_] == ']') { ++pos_; return true; }
      for (;;) {
        JValue child;
        if (!Value(&child, depth + 1)) return false;
        v->items.push_back(std::move(child));
        SkipWs();
        if (pos_ < s_.size() && s_[pos_] == ',') { ++pos_; continue; }
        if (pos_ < s_.size() && s_[pos_] == ']') { ++pos_; return true; }
        return Fail("expected ',' or ']'");
      }
    }
    if (c == '"') { v->kind = JValue::kString; return String(&v->str); }
    if (s_.compare(pos_, 4, "true") == 0) { v->kind = JValue::kBool; v->boolean = true; pos_ += 4; return true; }
    if (s_.compare(pos_, 5, "false") == 0) { v->kind = JValue::kBool; pos_ += 5; return true; }
    if (s_.compare(pos_, 4, "null") == 0) { v->kind = JValue::kNull; pos_ += 4; return true; }
    const char* begin = s_.c_str() + pos_;
    char* end = nullptr;
    double d = std::strtod(begin, &end);
    if (end == begin) return Fail("unexpected character");
    pos_ += (size_t)(end - begin);
    v->kind = JValue::kNumber;
    v->number = d;
    return true;
  }
};

// Reads a float or an array of 1 or 3 floats into dst[0..n).
bool ReadFloats(const JValue* v, float* dst, int n) {
  if (!v) return false;
  if (v->kind == JValue::kNumber) {
    for (int i = 0; i < n; ++i) dst[i] = (float)v->number;
    return true;
  }
  if (v->kind != JValue::kArray || v->items.empty()) return false;
  for (const auto& it : v->items)
    if (it.kind != JValue::kNumber) return false;
  if ((int)v->items.size() == n) {
    for (int i = 0; i < n; ++i) dst[i] = (float)v->items[i].number;
    return true;
  }
  if (v->items.size() == 1) {
    for (int i = 0; i < n; ++i) dst[i] = (float)v->items[0].number;
    return true;
  }
  return false;
}

std::string FmtFloat(float f) {
  char buf[64];
  std::snprintf(buf, sizeof(buf), "%.9g", (double)f);
  return buf;
}

}  // namespace

std::string LoadRawScene(const std::string& path, const std::string& jsonPath, HostScene* out, int* status) {
  *status = GPUGI_ERR_FORMAT;
  std::ifstream f(path, std::ios::binary);
  if (!f) { *status = GPUGI_ERR_IO; return "cannot open " + path; }
  // texture files are relative to the scene file's directory (Scene::m_sourceDirectory)
  const size_t slash = path.find_last_of('/');
  const std::string sceneDir = slash == std::string::npos ? std::string() : path.substr(0, slash + 1);
  out->textures.clear();
  out->materialTextures.clear();
  std::vector<std::string> materialNames;
  bool haveVertices = false, haveTriangles = false, haveHierarchy = false, haveBoxes = false;
  for (;;) {
    NamedArray h;
    f.read(reinterpret_cast<char*>(&h), sizeof(h));
    if (f.gcount() == 0 && f.eof()) break;
    if (f.gcount() != (std::streamsize)sizeof(h)) return "truncated chunk header";
    h.name[31] = '\0';
    const uint64_t bytes = (uint64_t)h.numElements * h.elementSize;
    if (bytes > (1ull << 36)) return "chunk too large";
    std::string name(h.name);
    auto readInto = [&](void* dst) -> bool {
      f.read(reinterpret_cast<char*>(dst), (std::streamsize)bytes);
      return (uint64_t)f.gcount() == bytes;
    };
    if (name == "vertices") {
      if (h.elementSize != sizeof(RawVertex)) return "vertices: unexpected element size";
      out->vertices.resize(h.numElements);
      if (!readInto(out->vertices.data())) return "vertices: truncated";
      haveVertices = true;
    } else if (name == "triangles") {
      if (h.elementSize == 0 || h.elementSize % sizeof(RawTriangle) != 0)
        return "triangles: element size is not a multiple of 16";
      out->trianglesPerLeaf = h.elementSize / (uint32_t)sizeof(RawTriangle);
      out->leafSlots.resize((size_t)h.numElements * out->trianglesPerLeaf);
      if (!readInto(out->leafSlots.data())) return "triangles: truncated";
      haveTriangles = true;
    } else if (name == "hierarchy") {
      if (h.elementSize != sizeof(RawNode)) return "hierarchy: unexpected element size";
      out->nodes.resize(h.numElements);
      if (!readInto(out->nodes.data())) return "hierarchy: truncated";
      haveHierarchy = true;
    } else if (name == "bounding_aabox") {
      if (h.elementSize != sizeof(RawBox)) return "bounding_aabox: unexpected element size";
      out->boxes.resize(h.numElements);
      if (!readInto(out->boxes.data())) return "bounding_aabox: truncated";
      haveBoxes = true;
    } else if (name == "materialref") {
      if (h.elementSize != 32) return "materialref: unexpected element size";
      std::vector<char> buf(bytes);
      if (!readInto(buf.data())) return "materialref: truncated";
      for (uint32_t i = 0; i < h.numElements; ++i) {
        char nm[33];
        std::memcpy(nm, &buf[(size_t)i * 32], 32);
        nm[32] = '\0';
        materialNames.emplace_back(nm);
      }
    } else {
      // Optional chunk this loader does not use (tangents, qormals, texcoords,
      // approx_sggx, surfels ...): skip numElements * elementSize bytes.
      f.seekg((std::streamoff)bytes, std::ios::cur);
      if (!f) return "chunk '" + name + "': truncated";
    }
  }
  if (!haveVertices) return "missing 'vertices' chunk";
  if (!haveTriangles) return "missing 'triangles' chunk";
  if (!haveHierarchy || !haveBoxes) return "missing 'hierarchy' or 'bounding_aabox' chunk";
  if (out->boxes.size() != out->nodes.size()) return "hierarchy and bounding_aabox sizes differ";

  // Materials: names from materialref, values from the json (defaults if absent).
  out->materials.clear();
  for (const auto& nm : materialNames) out->materials.push_back(DefaultMaterialParams(nm.c_str()));
  if (!jsonPath.empty()) {
    std::ifstream jf(jsonPath);
    if (!jf) { *status = GPUGI_ERR_IO; return "cannot open " + jsonPath; }
    std::stringstream ss;
    ss << jf.rdbuf();
    std::string text = ss.str();
    JValue root;
    std::string err;
    if (!JParser(text).Parse(&root, &err)) return err;
    const JValue* mats = root.Get("materials");
    if (mats && mats->kind == JValue::kObject) {
      for (const auto& m : mats->members) {
        size_t idx = out->materials.size();
        for (size_t i = 0; i < out->materials.size(); ++i)
          if (m.first == out->materials[i].name) idx = i;
        if (idx == out->materials.size()) {
          if (!materialNames.empty()) continue;  // not referenced by the geometry
          out->materials.push_back(DefaultMaterialParams(m.first.c_str()));
        }
        gpugi_material_params& p = out->materials[idx];
        // Scene::LoadMaterial (scene.cpp:155-178): albedo, opacity, reflectiveness
        // and emissivity are either values or texture files (a string, relative
        // to the scene's directory), each file loaded once (GetBindlessHandle).
        static const char* kSlotKeys[4] = {"albedo", "opacity", "reflectiveness", "emissivity"};
        for (int slot = 0; slot < 4; ++slot) {
          const JValue* v = m.second.Get(kSlotKeys[slot]);
          if (!v || v->kind != JValue::kString) continue;
          std::string rel = v->str;
          for (char& ch : rel)
            if (ch == '\\') ch = '/';  // the reference's scene files use Windows separators
          const std::string file = sceneDir + rel;
          int tex = -1;
          for (size_t t = 0; t < out->textures.size(); ++t)
            if (out->textures[t].file == file) tex = (int)t;
          if (tex < 0) {
            HostTexture ht;
            const std::string terr = LoadTextureFile(file, &ht);
            if (!terr.empty()) { *status = GPUGI_ERR_IO; return terr; }
            ht.name = v->str;
            out->textures.push_back(std::move(ht));
            tex = (int)out->textures.size() - 1;
          }
          if (out->materialTextures.size() < out->materials.size())
            out->materialTextures.resize(out->materials.size(), {-1, -1, -1, -1});
          out->materialTextures[idx][slot] = tex;
        }
        ReadFloats(m.second.Get("albedo"), p.albedo, 3);
        ReadFloats(m.second.Get("specularColor"), p.specular_color, 3);
        ReadFloats(m.second.Get("roughness"), &p.roughness, 1);
        ReadFloats(m.second.Get("opacity"), &p.opacity, 1);
        ReadFloats(m.second.Get("emissivity"), p.emissivity, 3);
        ReadFloats(m.second.Get("refractionIndex"), &p.refraction_index, 1);
        if (ReadFloats(m.second.Get("reflectiveness"), p.reflectiveness, 4)) p.has_reflectiveness = 1;
      }
    }
    if (!out->materialTextures.empty()) out->materialTextures.resize(out->materials.size(), {-1, -1, -1, -1});
    const JValue* lights = root.Get("lights");
    if (lights && lights->kind == JValue::kObject) {
      for (const auto& l : lights->members) {
        PointLight pl;
        if (ReadFloats(l.second.Get("position"), pl.position, 3) &&
            ReadFloats(l.second.Get("intensity"), pl.intensity, 3))
          out->pointLights.push_back(pl);
      }
    }
  }
  std::string err = out->Finalize();
  if (!err.empty()) return err;
  *status = GPUGI_OK;
  return std::string();
}

std::string SaveRawScene(const HostScene& s, const std::string& path, const std::string& jsonPath) {
  std::ofstream f(path, std::ios::binary);
  if (!f) return "cannot open " + path + " for writing";
  auto header = [&](const char* name, uint32_t num, uint32_t size) {
    NamedArray h;
    std::memset(&h, 0, sizeof(h));
    std::strncpy(h.name, name, 31);
    h.numElements = num;
    h.elementSize = size;
    f.write(reinterpret_cast<const char*>(&h), sizeof(h));
  };
  // Chunk order of bvhmake's main (bvhmake/main.cpp:129-141): materialref,
  // vertices, bounding_aabox, hierarchy, triangles.
  header("materialref", (uint32_t)s.materials.size(), 32);
  for (const auto& m : s.materials) {
    char nm[32];
    std::memset(nm, 0, 32);
    std::strncpy(nm, m.name, 31);
    f.write(nm, 32);
  }
  header("vertices", (uint32_t)s.vertices.size(), sizeof(RawVertex));
  f.write(reinterpret_cast<const char*>(s.vertices.data()), (std::streamsize)(s.vertices.size() * sizeof(RawVertex)));
  header("bounding_aabox", (uint32_t)s.boxes.size(), sizeof(RawBox));
  f.write(reinterpret_cast<const char*>(s.boxes.data()), (std::streamsize)(s.boxes.size() * sizeof(RawBox)));
  header("hierarchy", (uint32_t)s.nodes.size(), sizeof(RawNode));
  f.write(reinterpret_cast<const char*>(s.nodes.data()), (std::streamsize)(s.nodes.size() * sizeof(RawNode)));
  header("triangles", s.NumLeaves(), (uint32_t)(sizeof(RawTriangle) * s.trianglesPerLeaf));
  f.write(reinterpret_cast<const char*>(s.leafSlots.data()), (std::streamsize)(s.leafSlots.size() * sizeof(RawTriangle)));
  if (!f) return "write failed for " + path;
  if (jsonPath.empty()) return std::string();
  std::ofstream j(jsonPath);
  if (!j) return "cannot open " + jsonPath + " for writing";
  auto vec = [](const float* v, int n) {
    std::string r = "[";
    for (int i = 0; i < n; ++i) r += (i ? ", " : "") + FmtFloat(v[i]);
    return r + "]";
  };
  j << "{\n\t\"materials\": {\n";
  for (size_t i = 0; i < s.materials.size(); ++i) {
    const auto& m = s.materials[i];
    // a textured slot is written as its texture file, as it was read
    auto slot = [&](int k, const std::string& value) {
      if (i < s.materialTextures.size() && s.materialTextures[i][k] >= 0) {
        std::string name = s.textures[s.materialTextures[i][k]].name;
        std::string esc;
        for (char ch : name) {
          if (ch == '\\' || ch == '"') esc.push_back('\\');
          esc.push_back(ch);
        }
        return "\"" + esc + "\"";
      }
      return value;
    };
    j << "\t\t\"" << m.name << "\": {\n"
      << "\t\t\t\"type\": \"legacy\",\n"
      << "\t\t\t\"albedo\": " << slot(0, vec(m.albedo, 3)) << ",\n"
      << "\t\t\t\"specularColor\": " << vec(m.specular_color, 3) << ",\n"
      << "\t\t\t\"roughness\": [" << FmtFloat(m.roughness) << "],\n"
      << "\t\t\t\"opacity\": " << slot(1, "[" + FmtFloat(m.opacity) + "]") << ",\n";
    if (m.has_reflectiveness || (i < s.materialTextures.size() && s.materialTextures[i][2] >= 0))
      j << "\t\t\t\"reflectiveness\": " << slot(2, vec(m.reflectiveness, 4)) << ",\n";
    j << "\t\t\t\"emissivity\": " << slot(3, vec(m.emissivity, 3)) << ",\n"
      << "\t\t\t\"refractionIndex\": [" << FmtFloat(m.refraction_index) << "]\n"
      << "\t\t}" << (i + 1 < s.materials.size() ? "," : "") << "\n";
  }
  j << "\t},\n\t\"lights\": {\n";
  for (size_t i = 0; i < s.pointLights.size(); ++i) {
    j << "\t\t\"point" << i << "\": {\"position\": " << vec(s.pointLights[i].position, 3)
      << ", \"intensity\": " << vec(s.pointLights[i].intensity, 3) << "}"
      << (i + 1 < s.pointLights.size() ? "," : "") << "\n";
  }
  j << "\t}\n}\n";
  if (!j) return "write failed for " + jsonPath;
  return std::string();
}

}  // namespace gpugi
