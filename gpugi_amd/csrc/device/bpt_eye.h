// The eye paths of the bidirectional path tracer (pathtracer_bidir.comp:16-251
// without CHAINED_MIS), one bounce split into the steps around its two rays:
//   bpt_eye_hit        the vertex at the closest hit and its MIS-weighted emission
//   bpt_connect_setup  the connection to one random light-cache vertex (its ray)
//   bpt_connect_value  what a visible connection adds to the color
//   bpt_eye_continue   BSDF sample, roulette and the MIS state of the next vertex
// The one-kernel eye pass runs them in order with its own traces; the
// wavefront schedule traces the two rays in separate passes.  Both add the
// same terms in the same order with the same arithmetic.
#pragma once
#include <hip/hip_runtime.h>

#include "integrator_common.h"

namespace gpugi {

__device__ __forceinline__ float mis_heuristic(float p) { return fminf(1e18f, p); }     // lightcache.glsl:61-73
__device__ __forceinline__ float distance_mis(float d, float c) { return mis_heuristic(d * d); }  // :84-89

struct BptEyePath {
  gg::V3 color, tp, o, d;
  uint32_t seed;
  int bounce;
  float lastConnectionProbability, lastCosAbs, mis;
};

// The vertex a bounce's closest hit makes.
struct BptEyeVertex {
  gg::V3 n;
  MaterialData m;
  float rayHit;
};

// The connection of a vertex to light-cache entry e: a shadow ray from lo along
// ld of length tmax, traced only when `test` (the combined path is not too long).
struct BptConnection {
  bool test;
  gg::V3 lo, ld;
  float tmax, lightDist, lightDistSq;
  const LightCacheEntryDev* e;
};

__device__ __forceinline__ void bpt_eye_begin(const DrawParams& p, BptEyePath& s, int gx, int gy) {
  s.seed = gg::init_random_seed(p.frameSeed, (uint32_t)(gx + gy * p.width));
  const float j0 = gg::random01(s.seed);
  const float j1 = gg::random01(s.seed);
  const gg::V3 U = gg::v3(p.camera.u[0], p.camera.u[1], p.camera.u[2]);
  const gg::V3 V = gg::v3(p.camera.v[0], p.camera.v[1], p.camera.v[2]);
  const gg::V3 W = gg::v3(p.camera.w[0], p.camera.w[1], p.camera.w[2]);
  s.o = gg::v3(p.camera.position[0], p.camera.position[1], p.camera.position[2]);
  s.d = gg::primary_direction(j0, j1, gx, gy, p.width, p.height, U, V, W);
  s.color = gg::v3s(0.0f);
  s.tp = gg::v3s(1.0f);
  s.lastConnectionProbability = 1.0f;  // InitialMIS
  s.lastCosAbs = 1.0f;
  s.mis = 0.0f;
  s.bounce = 0;
}

// The ray (s.o, s.d) hit h: moves to the hit and adds its emission, MIS
// weighted (pathtracer_bidir.comp:57-92).
__device__ __forceinline__ void bpt_eye_hit(const DrawParams& p, BptEyePath& s, const ClosestHit& h,
                                            BptEyeVertex& v) {
  const int i = s.bounce;
  v.rayHit = h.t;
  s.o = s.o + v.rayHit * s.d;
  v.n = shading_normal(p.ref.vinfo, h.tri, h.bary);
  v.m = load_material_hit(p.materials, p.tex, p.ref.vinfo, h.tri, h.bary);
  const float incidentSurfaceCosSat = gg::saturate(-gg::dot(v.n, s.d));
  const float connectionLightToEyePath =
      distance_mis(v.rayHit, s.lastCosAbs) * mis_heuristic(incidentSurfaceCosSat / gg::kPI);
  const int spl = p.specificPathLength;
  if (spl < 0 || i == spl) {  // SHOW_SPECIFIC_PATHLENGTH (pathtracer_bidir.comp:75-92)
    s.mis = (i == 0) ? 1.0f
                     : (s.lastConnectionProbability /
                        fmaxf(s.lastConnectionProbability + connectionLightToEyePath, gg::kDivisorEpsilon));
    s.color = s.color + s.mis * s.tp * v.m.emissivity;
  }
}

// Connection to one random light-cache vertex (pathtracer_bidir.comp:99-116);
// actualN: light-cache entries in use (> 0).
__device__ __forceinline__ BptConnection bpt_connect_setup(const DrawParams& p, BptEyePath& s,
                                                           const LightCacheEntryDev* __restrict__ cache,
                                                           uint32_t actualN) {
  BptConnection c;
  const uint32_t idx = gg::xorshift32(s.seed) % actualN;
  c.e = &cache[idx];
  const int pathLength = c.e->pathLength + s.bounce + 1;
  const int spl = p.specificPathLength;
  // a connection closes a path of pathLength bounces (pathtracer_bidir.comp:101-108)
  c.test = pathLength <= p.maxPathLength && (spl < 0 || pathLength == spl);
  if (c.test) {
    const gg::V3 epos = gg::v3(c.e->position[0], c.e->position[1], c.e->position[2]);
    gg::V3 ld = epos - s.o;
    c.lightDistSq = gg::dot(ld, ld);
    c.lightDist = sqrtf(c.lightDistSq);
    c.ld = ld / fmaxf(c.lightDist, gg::kDivisorEpsilon);
    c.lo = gg::kRayHitEpsilon * c.ld + s.o;
    c.tmax = c.lightDist - gg::kRayHitEpsilon * 2.0f;
  }
  return c;
}

// What a visible connection adds to the color (pathtracer_bidir.comp:117-205).
// The MIS weight it computes is its own: the next use of s.mis assigns it anew.
__device__ __forceinline__ gg::V3 bpt_connect_value(const DrawParams& p, const BptEyePath& s, const BptEyeVertex& v,
                                                    const BptConnection& c, float lightSampleWeight) {
  const LightCacheEntryDev& e = *c.e;
  const gg::V3 ld = c.ld;
  const int lmat = e.materialIndex;
  const gg::V3 flux = gg::v3(e.flux[0], e.flux[1], e.flux[2]);
  gg::V3 lnormal = gg::v3(0.0f, 1.0f, 0.0f);
  float cosLight = 0.0f;
  gg::V3 lintensity;
  const float cosCamera = fabsf(gg::dot(ld, v.n));
  if (e.normal1 < 1.5f) {
    lnormal = gg::unpack_normal_raw(e.normal0, e.normal1);
    cosLight = fabsf(gg::dot(ld, lnormal));
  }
  float pdf;
  const gg::V3 bsdf = bsdf_eye(s.d, ld, v.m, v.n, pdf);
  float eyeToLight = mis_heuristic(pdf);
  float mis;
  if (s.bounce > 0) {
    const float distMIS = distance_mis(v.rayHit, s.lastCosAbs);
    eyeToLight = eyeToLight * distMIS;
    mis = s.lastConnectionProbability / fmaxf(s.lastConnectionProbability + eyeToLight, gg::kDivisorEpsilon);
    if (v.m.refl[3] >= 100.0f) mis = 1.0f;
  } else {
    mis = eyeToLight / fmaxf(s.lastConnectionProbability + eyeToLight, gg::kDivisorEpsilon);
  }
  float misLight = 1.0f;
  if (e.normal1 < 1.5f) {
    float lightToEye;
    if (lmat != -1) {
      // the light vertex's material at its stored texture coordinate (lightcache.glsl:1-16)
      const MaterialData lm = load_material_uv(p.materials, p.tex, lmat, make_float2(e.texcoord[0], e.texcoord[1]));
      const gg::V3 incidentDir = gg::unpack_normal(e.incidentPacked[0], e.incidentPacked[1]);
      const gg::V3 lbsdf = bsdf_adjoint(ld, -incidentDir, lm, lnormal, pdf);
      lintensity = flux * lbsdf * cosLight;
      lightToEye = mis_heuristic(pdf) * distance_mis(c.lightDist, cosCamera);
    } else {
      lintensity = flux * (cosLight * gg::kPI_2);
      lightToEye = mis_heuristic(cosLight / gg::kPI) * distance_mis(c.lightDist, cosCamera);
    }
    misLight = lightToEye / fmaxf(lightToEye + eyeToLight, gg::kDivisorEpsilon);
  } else {
    lintensity = flux * gg::kPI_2;
    misLight = 1.0f;
  }
  mis = mis * misLight;
  const gg::V3 irradiance = (mis * lightSampleWeight * cosCamera / c.lightDistSq) * lintensity;
  return irradiance * s.tp * bsdf;
}

// A connection whose value is exactly zero (for example to a light vertex
// behind an opaque surface, where the BSDF is zero) leaves the color as it is
// whether or not its ray is blocked: the color starts at +0 and only gains
// non-negative terms, so adding a zero never changes its bits.  Its ray is not
// traced.  (A NaN value fails the test and is traced and added as before.)
__device__ __forceinline__ bool bpt_value_is_zero(gg::V3 v) { return v.x == 0.0f && v.y == 0.0f && v.z == 0.0f; }

// The rest of the bounce (pathtracer_bidir.comp:207-247): the next direction,
// roulette from rrStartDepth on, the MIS state; true when the path has ended.
__device__ __forceinline__ bool bpt_eye_continue(const DrawParams& p, BptEyePath& s, const BptEyeVertex& v) {
  const int i = s.bounce;
  float samplePDF;
  gg::V3 t1 = gg::v3s(1.0f);
  bool diffuse;
  s.d = sample_bsdf(s.d, v.m, s.seed, v.n, t1, false, samplePDF, diffuse);
  s.o = s.o + s.d * gg::kRayHitEpsilon;
  s.tp = s.tp * t1;
  if (i >= p.rrStartDepth) {  // Russian roulette from rrStartDepth on
    const float q = gg::saturate(gg::luminance(t1));
    if (gg::random01(s.seed) >= q) return true;
    s.tp = s.tp / q;
  }
  float reverse = mis_heuristic(samplePDF);
  if (i == 0) {
    s.mis = reverse / fmaxf(reverse + s.lastConnectionProbability, gg::kDivisorEpsilon);
  } else {
    const float distMIS = distance_mis(v.rayHit, s.lastCosAbs);
    reverse = reverse * distMIS;
    s.mis = s.lastConnectionProbability / fmaxf(s.lastConnectionProbability + distMIS * reverse, gg::kDivisorEpsilon);
    if (v.m.refl[3] >= 100.0f) s.mis = 1.0f;
  }
  s.tp = s.tp * s.mis;
  s.lastConnectionProbability = reverse;
  s.lastCosAbs = fabsf(gg::dot(s.d, v.n));
  ++s.bounce;
  return s.bounce >= p.maxPathLength;
}

}  // namespace gpugi
