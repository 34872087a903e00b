// Deterministic synthetic scenes: the Cornell-style room (36 triangles), the
// fBm heightfield of the traversal microbenchmark and the clutter room of
// icospheres (5,120,016 triangles at 4000 spheres).
#include <cmath>
#include <cstdio>
#include <cstring>
#include <utility>

#include "../common/gmath.h"
#include "scene.hpp"

namespace gpugi {
namespace {

struct P3 {
  double x, y, z;
};

void AddQuad(std::vector<RawVertex>* v, std::vector<RawTriangle>* t, P3 a, P3 b, P3 c, P3 d, P3 n,
             uint32_t mat) {
  // Triangles (a, b, c) and (a, c, d) with their own four vertices (flat shading).
  uint32_t base = (uint32_t)v->size();
  for (const P3& p : {a, b, c, d}) {
    RawVertex rv;
    rv.position[0] = (float)p.x; rv.position[1] = (float)p.y; rv.position[2] = (float)p.z;
    rv.normal[0] = (float)n.x; rv.normal[1] = (float)n.y; rv.normal[2] = (float)n.z;
    rv.texcoord[0] = 0.0f; rv.texcoord[1] = 0.0f;
    v->push_back(rv);
  }
  t->push_back(RawTriangle{{base, base + 1, base + 2}, mat});
  t->push_back(RawTriangle{{base, base + 2, base + 3}, mat});
}

gpugi_material_params Mat(const char* name, float r, float g, float b) {
  gpugi_material_params p = DefaultMaterialParams(name);
  p.albedo[0] = r; p.albedo[1] = g; p.albedo[2] = b;
  for (int i = 0; i < 3; ++i) p.specular_color[i] = 0.0f;
  p.roughness = 1.0f;
  p.opacity = 1.0f;
  return p;
}

void AddBox(std::vector<RawVertex>* v, std::vector<RawTriangle>* t, double cx, double cz, double sx,
            double sy, double sz, double angleDeg, uint32_t mat) {
  const double a = angleDeg * 3.14159265358979323846 / 180.0;
  const double ca = std::cos(a), sa = std::sin(a);
  auto rot = [&](double x, double y, double z) { return P3{cx + ca * x + sa * z, y, cz - sa * x + ca * z}; };
  auto rn = [&](double x, double y, double z) { return P3{ca * x + sa * z, y, -sa * x + ca * z}; };
  const double hx = sx / 2, hz = sz / 2;
  P3 c000 = rot(-hx, 0, -hz), c100 = rot(hx, 0, -hz), c110 = rot(hx, sy, -hz), c010 = rot(-hx, sy, -hz);
  P3 c001 = rot(-hx, 0, hz), c101 = rot(hx, 0, hz), c111 = rot(hx, sy, hz), c011 = rot(-hx, sy, hz);
  AddQuad(v, t, c001, c101, c111, c011, rn(0, 0, 1), mat);   // +z
  AddQuad(v, t, c100, c000, c010, c110, rn(0, 0, -1), mat);  // -z
  AddQuad(v, t, c101, c100, c110, c111, rn(1, 0, 0), mat);   // +x
  AddQuad(v, t, c000, c001, c011, c010, rn(-1, 0, 0), mat);  // -x
  AddQuad(v, t, c011, c111, c110, c010, rn(0, 1, 0), mat);   // top
  AddQuad(v, t, c000, c100, c101, c001, rn(0, -1, 0), mat);  // bottom
}

// Value noise on an integer lattice hashed with wanghash, smoothstep blend.
double LatticeValue(int32_t ix, int32_t iz, uint32_t octave, uint32_t seed) {
  uint32_t h = gg::wanghash((uint32_t)iz + 0x85ebca77u * (octave + 1u));
  h = gg::wanghash((uint32_t)ix * 0x9e3779b1u ^ h);
  h = gg::wanghash(seed ^ h);
  return (double)h / 4294967295.0;
}
double ValueNoise(double x, double z, uint32_t octave, uint32_t seed) {
  double fx = std::floor(x), fz = std::floor(z);
  int32_t ix = (int32_t)fx, iz = (int32_t)fz;
  double tx = x - fx, tz = z - fz;
  tx = tx * tx * (3.0 - 2.0 * tx);
  tz = tz * tz * (3.0 - 2.0 * tz);
  double v00 = LatticeValue(ix, iz, octave, seed), v10 = LatticeValue(ix + 1, iz, octave, seed);
  double v01 = LatticeValue(ix, iz + 1, octave, seed), v11 = LatticeValue(ix + 1, iz + 1, octave, seed);
  double a = v00 + (v10 - v00) * tx, b = v01 + (v11 - v01) * tx;
  return a + (b - a) * tz;
}
double Fbm(double x, double z, uint32_t seed) {
  double sum = 0, norm = 0, amp = 1, freq = 4.0;
  for (uint32_t o = 0; o < 6; ++o) {
    sum += amp * ValueNoise(x * freq, z * freq, o, seed);
    norm += amp;
    amp *= 0.5;   // gain
    freq *= 2.0;  // lacunarity
  }
  return sum / norm;
}

// Unit icosphere: the icosahedron subdivided `levels` times (20 * 4^levels
// triangles), vertices projected onto the sphere; midpoints are shared.
void Icosphere(int levels, std::vector<P3>* verts, std::vector<uint32_t>* idx) {
  const double g = (1.0 + std::sqrt(5.0)) / 2.0;
  *verts = {{-1, g, 0}, {1, g, 0}, {-1, -g, 0}, {1, -g, 0}, {0, -1, g}, {0, 1, g},
            {0, -1, -g}, {0, 1, -g}, {g, 0, -1}, {g, 0, 1}, {-g, 0, -1}, {-g, 0, 1}};
  *idx = {0, 11, 5, 0, 5, 1, 0, 1, 7, 0, 7, 10, 0, 10, 11, 1, 5, 9, 5, 11, 4, 11, 10, 2, 10, 7, 6, 7, 1, 8,
          3, 9, 4, 3, 4, 2, 3, 2, 6, 3, 6, 8, 3, 8, 9, 4, 9, 5, 2, 4, 11, 6, 2, 10, 8, 6, 7, 9, 8, 1};
  auto norm = [](P3 p) {
    const double l = std::sqrt(p.x * p.x + p.y * p.y + p.z * p.z);
    return P3{p.x / l, p.y / l, p.z / l};
  };
  for (P3& p : *verts) p = norm(p);
  for (int l = 0; l < levels; ++l) {
    std::vector<uint32_t> next;
    next.reserve(idx->size() * 4);
    std::vector<std::pair<uint64_t, uint32_t>> mids;  // (edge key, vertex), small per level
    auto midpoint = [&](uint32_t a, uint32_t b) {
      const uint64_t key = a < b ? ((uint64_t)a << 32 | b) : ((uint64_t)b << 32 | a);
      for (const auto& m : mids)
        if (m.first == key) return m.second;
      const P3 pa = (*verts)[a], pb = (*verts)[b];
      verts->push_back(norm(P3{(pa.x + pb.x) / 2, (pa.y + pb.y) / 2, (pa.z + pb.z) / 2}));
      mids.push_back({key, (uint32_t)verts->size() - 1});
      return (uint32_t)verts->size() - 1;
    };
    for (size_t i = 0; i < idx->size(); i += 3) {
      const uint32_t a = (*idx)[i], b = (*idx)[i + 1], c = (*idx)[i + 2];
      const uint32_t ab = midpoint(a, b), bc = midpoint(b, c), ca = midpoint(c, a);
      next.insert(next.end(), {a, ab, ca, b, bc, ab, c, ca, bc, ab, bc, ca});
    }
    *idx = std::move(next);
  }
}

// Deterministic uniform doubles in [0, 1) from a xorshift64* stream.
struct Rng {
  uint64_t s;
  double next() {
    s ^= s >> 12;
    s ^= s << 25;
    s ^= s >> 27;
    return (double)((s * 0x2545F4914F6CDD1Dull) >> 11) * (1.0 / 9007199254740992.0);
  }
};

}  // namespace

void MakeClutterMesh(uint32_t spheres, uint32_t seed, std::vector<RawVertex>* v, std::vector<RawTriangle>* t,
                     std::vector<gpugi_material_params>* m, std::vector<PointLight>* lights) {
  // Room [-10,10] x [0,8] x [-10,10]; materials: 0 room, 1 light, 2..121 Lambertian bins
  // (albedo uniform in [0.2, 0.9] per channel), 122..241 Phong-glossy bins (exponent
  // log-uniform in [10, 1000]).  Spheres pick a Lambertian bin with probability 0.6.
  const int kBins = 120;
  Rng rng{0x9E3779B97F4A7C15ull ^ ((uint64_t)seed << 1 | 1)};
  m->clear();
  m->push_back(Mat("Room", 0.7f, 0.7f, 0.7f));
  gpugi_material_params light = Mat("Light", 0.0f, 0.0f, 0.0f);
  light.emissivity[0] = light.emissivity[1] = light.emissivity[2] = 40.0f;
  m->push_back(light);
  for (int i = 0; i < kBins; ++i) {
    char name[32];
    std::snprintf(name, sizeof(name), "Lambert%03d", i);
    m->push_back(Mat(name, (float)(0.2 + 0.7 * rng.next()), (float)(0.2 + 0.7 * rng.next()),
                     (float)(0.2 + 0.7 * rng.next())));
  }
  for (int i = 0; i < kBins; ++i) {
    char name[32];
    std::snprintf(name, sizeof(name), "Glossy%03d", i);
    const double albedo = 0.2 + 0.7 * rng.next();
    gpugi_material_params g = Mat(name, (float)(0.3 * albedo), (float)(0.3 * albedo), (float)(0.3 * albedo));
    const double exponent = std::exp(std::log(10.0) + (std::log(1000.0) - std::log(10.0)) * rng.next());
    g.roughness = (float)(1.0 / std::sqrt(exponent));  // exponent = 1 / roughness^2 (scene.cpp:160,172)
    for (int c = 0; c < 3; ++c) g.specular_color[c] = (float)(0.5 * albedo + 0.3);
    m->push_back(g);
  }
  v->clear();
  t->clear();
  AddQuad(v, t, {-10, 0, 10}, {10, 0, 10}, {10, 0, -10}, {-10, 0, -10}, {0, 1, 0}, 0);    // floor
  AddQuad(v, t, {-10, 8, -10}, {10, 8, -10}, {10, 8, 10}, {-10, 8, 10}, {0, -1, 0}, 0);   // ceiling
  AddQuad(v, t, {-10, 0, -10}, {10, 0, -10}, {10, 8, -10}, {-10, 8, -10}, {0, 0, 1}, 0);  // back
  AddQuad(v, t, {10, 0, 10}, {-10, 0, 10}, {-10, 8, 10}, {10, 8, 10}, {0, 0, -1}, 0);     // front
  AddQuad(v, t, {-10, 0, 10}, {-10, 0, -10}, {-10, 8, -10}, {-10, 8, 10}, {1, 0, 0}, 0);  // left
  AddQuad(v, t, {10, 0, -10}, {10, 0, 10}, {10, 8, 10}, {10, 8, -10}, {-1, 0, 0}, 0);     // right
  // Area light: a 2 x 2 quad just below the ceiling as a 4-triangle fan around its centre,
  // wound so that cross(v0 - v1, v0 - v2) points down into the room.
  {
    const uint32_t base = (uint32_t)v->size();
    const P3 c[5] = {{0, 7.99, 0}, {-1, 7.99, -1}, {1, 7.99, -1}, {1, 7.99, 1}, {-1, 7.99, 1}};
    for (const P3& p : c) {
      RawVertex rv;
      rv.position[0] = (float)p.x; rv.position[1] = (float)p.y; rv.position[2] = (float)p.z;
      rv.normal[0] = 0.0f; rv.normal[1] = -1.0f; rv.normal[2] = 0.0f;
      rv.texcoord[0] = rv.texcoord[1] = 0.0f;
      v->push_back(rv);
    }
    for (uint32_t k = 0; k < 4; ++k)
      t->push_back(RawTriangle{{base, base + 1 + (k + 1) % 4, base + 1 + k}, 1});
  }
  std::vector<P3> sv;
  std::vector<uint32_t> si;
  Icosphere(3, &sv, &si);
  v->reserve(v->size() + (size_t)spheres * sv.size());
  t->reserve(t->size() + (size_t)spheres * si.size() / 3);
  for (uint32_t s = 0; s < spheres; ++s) {
    const double r = std::exp(std::log(0.05) + (std::log(0.8) - std::log(0.05)) * rng.next());
    const double cx = -10.0 + r + (20.0 - 2 * r) * rng.next();
    const double cy = r + (8.0 - 2 * r) * rng.next();
    const double cz = -10.0 + r + (20.0 - 2 * r) * rng.next();
    const bool lambert = rng.next() < 0.6;
    const uint32_t bin = (uint32_t)(rng.next() * kBins) % kBins;
    const uint32_t mat = (lambert ? 2u : 2u + kBins) + bin;
    const uint32_t base = (uint32_t)v->size();
    for (const P3& p : sv) {
      RawVertex rv;
      rv.position[0] = (float)(cx + r * p.x); rv.position[1] = (float)(cy + r * p.y);
      rv.position[2] = (float)(cz + r * p.z);
      rv.normal[0] = (float)p.x; rv.normal[1] = (float)p.y; rv.normal[2] = (float)p.z;
      rv.texcoord[0] = rv.texcoord[1] = 0.0f;
      v->push_back(rv);
    }
    for (size_t i = 0; i < si.size(); i += 3)
      t->push_back(RawTriangle{{base + si[i], base + si[i + 1], base + si[i + 2]}, mat});
  }
  lights->clear();
  const float lp[3][3] = {{-6.0f, 6.5f, -5.0f}, {5.0f, 5.5f, 6.0f}, {6.0f, 7.0f, -6.0f}};
  for (int k = 0; k < 3; ++k) {
    PointLight pl;
    std::memcpy(pl.position, lp[k], 12);
    pl.intensity[0] = pl.intensity[1] = pl.intensity[2] = 30.0f;
    lights->push_back(pl);
  }
}

void MakeCornellMesh(std::vector<RawVertex>* v, std::vector<RawTriangle>* t,
                     std::vector<gpugi_material_params>* m) {
  // Room [-1,1] x [0,2] x [-1,1], open towards +z where the camera looks in from.
  m->clear();
  m->push_back(Mat("White", 0.75f, 0.75f, 0.75f));
  m->push_back(Mat("Red", 0.63f, 0.06f, 0.04f));
  m->push_back(Mat("Green", 0.14f, 0.45f, 0.09f));
  gpugi_material_params light = Mat("Light", 0.75f, 0.75f, 0.75f);
  light.emissivity[0] = 17.0f; light.emissivity[1] = 12.0f; light.emissivity[2] = 4.0f;
  m->push_back(light);
  v->clear();
  t->clear();
  // 5 walls, normals into the room.
  AddQuad(v, t, {-1, 0, 1}, {1, 0, 1}, {1, 0, -1}, {-1, 0, -1}, {0, 1, 0}, 0);    // floor
  AddQuad(v, t, {-1, 2, -1}, {1, 2, -1}, {1, 2, 1}, {-1, 2, 1}, {0, -1, 0}, 0);   // ceiling
  AddQuad(v, t, {-1, 0, -1}, {1, 0, -1}, {1, 2, -1}, {-1, 2, -1}, {0, 0, 1}, 0);  // back
  AddQuad(v, t, {-1, 0, 1}, {-1, 0, -1}, {-1, 2, -1}, {-1, 2, 1}, {1, 0, 0}, 1);  // left (red)
  AddQuad(v, t, {1, 0, -1}, {1, 0, 1}, {1, 2, 1}, {1, 2, -1}, {-1, 0, 0}, 2);     // right (green)
  // Ceiling light, wound so that cross(v0 - v1, v0 - v2) points down into the
  // room: that normal sets the emission side of the light sampler
  // (lightsampler.cpp:98).
  AddQuad(v, t, {-0.3, 1.99, -0.3}, {0.3, 1.99, -0.3}, {0.3, 1.99, 0.3}, {-0.3, 1.99, 0.3}, {0, -1, 0}, 3);
  AddBox(v, t, 0.35, 0.30, 0.6, 0.6, 0.6, -17.0, 0);   // short box
  AddBox(v, t, -0.35, -0.35, 0.6, 1.2, 0.6, 20.0, 0);  // tall box
}

void MakeHeightfieldMesh(uint32_t cells, uint32_t seed, std::vector<RawVertex>* v,
                         std::vector<RawTriangle>* t, std::vector<gpugi_material_params>* m) {
  m->clear();
  m->push_back(Mat("Terrain", 0.75f, 0.75f, 0.75f));
  const uint32_t n1 = cells + 1;
  std::vector<double> h((size_t)n1 * n1);
  for (uint32_t j = 0; j < n1; ++j)
    for (uint32_t i = 0; i < n1; ++i)
      h[(size_t)j * n1 + i] = 0.2 * Fbm((double)i / cells, (double)j / cells, seed);
  v->resize((size_t)n1 * n1);
  for (uint32_t j = 0; j < n1; ++j) {
    for (uint32_t i = 0; i < n1; ++i) {
      RawVertex& rv = (*v)[(size_t)j * n1 + i];
      rv.position[0] = (float)((double)i / cells);
      rv.position[1] = (float)h[(size_t)j * n1 + i];
      rv.position[2] = (float)((double)j / cells);
      uint32_t il = i ? i - 1 : i, ir = i + 1 < n1 ? i + 1 : i;
      uint32_t jl = j ? j - 1 : j, jr = j + 1 < n1 ? j + 1 : j;
      double dx = (h[(size_t)j * n1 + ir] - h[(size_t)j * n1 + il]) / ((double)(ir - il) / cells);
      double dz = (h[(size_t)jr * n1 + i] - h[(size_t)jl * n1 + i]) / ((double)(jr - jl) / cells);
      double nx = -dx, ny = 1.0, nz = -dz;
      double len = std::sqrt(nx * nx + ny * ny + nz * nz);
      rv.normal[0] = (float)(nx / len);
      rv.normal[1] = (float)(ny / len);
      rv.normal[2] = (float)(nz / len);
      rv.texcoord[0] = (float)((double)i / cells);
      rv.texcoord[1] = (float)((double)j / cells);
    }
  }
  t->clear();
  t->reserve((size_t)cells * cells * 2);
  for (uint32_t j = 0; j < cells; ++j) {
    for (uint32_t i = 0; i < cells; ++i) {
      uint32_t v00 = j * n1 + i, v10 = v00 + 1, v01 = v00 + n1, v11 = v01 + 1;
      t->push_back(RawTriangle{{v00, v10, v11}, 0});
      t->push_back(RawTriangle{{v00, v11, v01}, 0});
    }
  }
}

}  // namespace gpugi
