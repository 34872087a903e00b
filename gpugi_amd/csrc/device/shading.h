// Materials, BSDF and next-event estimation of the reference renderer (device
// code).  Every expression keeps the operand order of the GLSL it restates,
// and the translation unit is compiled with -ffp-contract=off, so a render
// with matched random streams reproduces a CPU restatement bit for bit.
#pragma once
#include <hip/hip_runtime.h>

#include "../common/gshade.h"
#include "gpugi_rt.h"

namespace gpugi {

// MaterialData (materials.glsl:9-18) for a constant material.
struct MaterialData {
  float refl[4];  // Reflectiveness, .w = Phong exponent
  gg::V3 opacity;
  gg::V3 diffuse;
  gg::V3 emissivity;
  gg::V3 fresnel0;
  float refractionIndexAvg;
  gg::V3 fresnel1;
};

// Material textures (the bindless handles of the reference's Material,
// globaluniforms.glsl:23-27): per material the texture of its diffuse,
// opacity, reflectiveness and emissivity slots (-1: the constant of the
// material record); per texture its linear RGBA float texels (row-major, row
// 0 at t = 0.5 / height) and size.  matTex == nullptr: no textures at all.
struct TextureDesc {
  const float4* texels;
  int width, height;
};
struct MaterialTextures {
  const int4* __restrict__ matTex = nullptr;
  const TextureDesc* __restrict__ desc = nullptr;
};

// textureLod(sampler, uv, 0) with the scene's sampler (scene.cpp:16-20: LINEAR
// magnification, REPEAT): bilinear over texel centres, coordinates wrapped.
__device__ __forceinline__ float4 sample_texture(const MaterialTextures& tx, int id, float2 uv) {
  const TextureDesc d = tx.desc[id];
  const float u = uv.x * (float)d.width - 0.5f, v = uv.y * (float)d.height - 0.5f;
  const float fu = floorf(u), fv = floorf(v);
  const float a = u - fu, b = v - fv;
  if (!(a >= 0.0f && a <= 1.0f && b >= 0.0f && b <= 1.0f)) return make_float4(0.0f, 0.0f, 0.0f, 0.0f);  // NaN / inf uv
  // REPEAT: fu, fv are integers (fmodf is exact on them); the clamp only
  // guards the index against what rounding cannot produce
  float wx = fmodf(fu, (float)d.width), wy = fmodf(fv, (float)d.height);
  wx = wx < 0.0f ? wx + (float)d.width : wx;
  wy = wy < 0.0f ? wy + (float)d.height : wy;
  const int x0 = min(max((int)wx, 0), d.width - 1), y0 = min(max((int)wy, 0), d.height - 1);
  const int x1 = x0 + 1 == d.width ? 0 : x0 + 1, y1 = y0 + 1 == d.height ? 0 : y0 + 1;
  const float4* t = d.texels;
  const float4 t00 = t[(size_t)y0 * d.width + x0], t10 = t[(size_t)y0 * d.width + x1];
  const float4 t01 = t[(size_t)y1 * d.width + x0], t11 = t[(size_t)y1 * d.width + x1];
  const float ia = 1.0f - a, ib = 1.0f - b;
  return make_float4((t00.x * ia + t10.x * a) * ib + (t01.x * ia + t11.x * a) * b,
                     (t00.y * ia + t10.y * a) * ib + (t01.y * ia + t11.y * a) * b,
                     (t00.z * ia + t10.z * a) * ib + (t01.z * ia + t11.z * a) * b,
                     (t00.w * ia + t10.w * a) * ib + (t01.w * ia + t11.w * a) * b);
}

__device__ __forceinline__ MaterialData load_material(const gpugi_material* __restrict__ m, int id) {
  const gpugi_material& s = m[id];
  MaterialData d;
  d.refl[0] = s.reflectiveness[0]; d.refl[1] = s.reflectiveness[1];
  d.refl[2] = s.reflectiveness[2]; d.refl[3] = s.reflectiveness[3];
  d.opacity = gg::v3(s.opacity[0], s.opacity[1], s.opacity[2]);
  d.diffuse = gg::v3(s.diffuse[0], s.diffuse[1], s.diffuse[2]);
  d.emissivity = gg::v3(s.emissivity[0], s.emissivity[1], s.emissivity[2]);
  d.fresnel0 = gg::v3(s.fresnel0[0], s.fresnel0[1], s.fresnel0[2]);
  d.refractionIndexAvg = s.refraction_index_avg;
  d.fresnel1 = gg::v3(s.fresnel1[0], s.fresnel1[1], s.fresnel1[2]);
  return d;
}

// textureLod(EnvironmentMap, direction, 0) of pathtracer.comp:57.  Face and
// (s, t) follow the OpenGL cube-map selection table (major axis; ties go to x,
// then y); faces are stored x-, x+, y-, y+, z-, z+, each size x size linear RGB
// texels, row j holding t = (j + 0.5) / size.  Bilinear filtering with the
// texel coordinates clamped to the face.  Directions that select no texel
// (zero or NaN components) give black.
__device__ __forceinline__ gg::V3 env_radiance(const float4* __restrict__ tex, int size, gg::V3 d) {
  const float ax = fabsf(d.x), ay = fabsf(d.y), az = fabsf(d.z);
  int face;
  float sc, tc, ma;
  if (ax >= ay && ax >= az) {
    ma = ax;
    face = d.x >= 0.0f ? 1 : 0;
    sc = d.x >= 0.0f ? -d.z : d.z;
    tc = -d.y;
  } else if (ay >= az) {
    ma = ay;
    face = d.y >= 0.0f ? 3 : 2;
    sc = d.x;
    tc = d.y >= 0.0f ? d.z : -d.z;
  } else {
    ma = az;
    face = d.z >= 0.0f ? 5 : 4;
    sc = d.z >= 0.0f ? d.x : -d.x;
    tc = -d.y;
  }
  const float s = (sc / ma + 1.0f) * 0.5f;
  const float t = (tc / ma + 1.0f) * 0.5f;
  if (!(ma > 0.0f) || !(s >= 0.0f && s <= 1.0f && t >= 0.0f && t <= 1.0f)) return gg::v3s(0.0f);
  const float u = s * (float)size - 0.5f, v = t * (float)size - 0.5f;
  const float fu = floorf(u), fv = floorf(v);
  const float a = u - fu, b = v - fv;
  const int i0 = (int)fu, j0 = (int)fv;
  const int x0 = max(i0, 0), x1 = min(i0 + 1, size - 1), y0 = max(j0, 0), y1 = min(j0 + 1, size - 1);
  const float4* f = tex + (size_t)face * size * size;
  const float4 t00 = f[(size_t)y0 * size + x0], t10 = f[(size_t)y0 * size + x1];
  const float4 t01 = f[(size_t)y1 * size + x0], t11 = f[(size_t)y1 * size + x1];
  const float ia = 1.0f - a, ib = 1.0f - b;
  return gg::v3((t00.x * ia + t10.x * a) * ib + (t01.x * ia + t11.x * a) * b,
                (t00.y * ia + t10.y * a) * ib + (t01.y * ia + t11.y * a) * b,
                (t00.z * ia + t10.z * a) * ib + (t01.z * ia + t11.z * a) * b);
}

// AvgProbability / Avg (materials.glsl:33-41).
__device__ __forceinline__ float avg_probability(gg::V3 c) {
  return (c.x + c.y + c.z + gg::kDivisorEpsilon) / (3.0f + gg::kDivisorEpsilon);
}
__device__ __forceinline__ float avg3(gg::V3 c) { return (c.x + c.y + c.z) / 3.0f; }
__device__ __forceinline__ float signf_(float x) { return x > 0.0f ? 1.0f : (x < 0.0f ? -1.0f : 0.0f); }

// GetBSDFDecisionPropabilities with USE_CONSTANT_REFLECTION (materials.glsl:43-63).
__device__ __forceinline__ void decision_probabilities(const MaterialData& m, gg::V3& preflect, gg::V3& prefract,
                                                       gg::V3& pdiffuse) {
  preflect = m.fresnel0;
  prefract = gg::v3s(1.0f) - gg::saturate(preflect);
  pdiffuse = prefract * m.opacity;
  prefract = prefract - pdiffuse;
}

// GetDiffuseProbability (materials.glsl:96-101).
__device__ __forceinline__ gg::V3 diffuse_probability(const MaterialData& m) {
  return (gg::v3s(1.0f) - gg::saturate(m.fresnel0)) * m.opacity;
}

// __SampleBSDF (materials.glsl:118-277).  Draws Random2 then Random.
__device__ __forceinline__ gg::V3 sample_bsdf(gg::V3 inc, const MaterialData& m, uint32_t& seed, gg::V3 N,
                                              gg::V3& tp, bool adjoint, float& pdf, bool& reflectedDiffuse) {
  const float cosTheta = gg::dot(N, inc);
  const float cosThetaAbs = gg::saturate(fabsf(cosTheta));
  gg::V3 rrd = inc - (2.0f * cosTheta) * N;
  const float r0 = gg::random01(seed);
  const float r1 = gg::random01(seed);
  gg::V3 preflect, prefract, pdiffuse;
  decision_probabilities(m, preflect, prefract, pdiffuse);
  const float avgPReflect = avg_probability(preflect);
  const float avgPRefract = avg_probability(prefract);
  const float avgPDiffuse = avg3(pdiffuse);
  gg::V3 pphong = preflect;
  float avgPPhong = avgPReflect;
  const float pathDecisionVar = gg::random01(seed);
  if (avgPDiffuse > pathDecisionVar) {
    reflectedDiffuse = true;
    tp = tp * ((m.diffuse * pdiffuse) / avgPDiffuse);
    if (cosTheta > 0.0f) N = -N;
    gg::V3 U, V;
    gg::create_onb(N, U, V);
    gg::V3 outDir = gg::sample_hemisphere_cosine(r0, r1, U, V, N);
    const float cosThetaOut = gg::dot(N, outDir);
    pdf = gg::saturate(cosThetaOut) / gg::kPI;
    return outDir;
  } else if (avgPDiffuse + avgPRefract > pathDecisionVar) {
    pphong = prefract;
    avgPPhong = avgPRefract;
    const float eta = cosTheta < 0.0f ? 1.0f / m.refractionIndexAvg : m.refractionIndexAvg;
    const float etaSq = eta * eta;
    const float sinTheta2Sq = etaSq * (1.0f - cosTheta * cosTheta);
    if (sinTheta2Sq < 1.0f) {
      rrd = eta * inc + (signf_(cosTheta) * (sqrtf(gg::saturate(1.0f - sinTheta2Sq)) - eta * cosThetaAbs)) * N;
      if (!adjoint) tp = tp / etaSq;
    }
  }
  reflectedDiffuse = false;
  rrd = gg::normalize(rrd);
  gg::V3 RU, RV;
  gg::create_onb(rrd, RU, RV);
  // The rejection loop of materials.glsl:224-240 runs exactly once (plausible = true).
  gg::V3 outDir = gg::sample_phong_lobe(r0, r1, m.refl[3], RU, RV, rrd);
  if (!adjoint) tp = tp * (fabsf(gg::dot(N, outDir)) / fabsf(gg::dot(N, rrd)));
  const float cosThetaOutAbs = fabsf(gg::dot(outDir, rrd));
  pdf = (m.refl[3] + 1.0f) / gg::kPI_2 * gg::powf_(cosThetaOutAbs, m.refl[3]);
  pdf = pdf * 1.0f;
  tp = tp * (pphong / avgPPhong);
  return outDir;
}

// __BSDF (materials.glsl:306-391).
__device__ __forceinline__ gg::V3 eval_bsdf(gg::V3 inc, gg::V3 exc, const MaterialData& m, gg::V3 N, float& pdf,
                                            bool adjoint, bool specularOnly) {
  gg::V3 bsdf = gg::v3s(0.0f);
  const float cosTheta = gg::dot(N, inc);
  const float cosThetaAbs = gg::saturate(fabsf(cosTheta));
  const float cosThetaOut = gg::dot(N, exc);
  gg::V3 preflect, prefract, pdiffuse;
  decision_probabilities(m, preflect, prefract, pdiffuse);
  const float phongFactorBrdf = (m.refl[3] + 1.0f) / (2.0f * gg::kPI_2 * fabsf(cosThetaOut) + gg::kDivisorEpsilon);
  const float phongNormalizationPdf = (m.refl[3] + 1.0f) / gg::kPI_2;
  gg::V3 rrd = gg::normalize(inc - (2.0f * cosTheta) * N);
  const float reflPow = gg::powf_(gg::saturate(gg::dot(rrd, exc)), m.refl[3]);
  float brdfFactor = phongFactorBrdf * reflPow;
  float pdfFactor = phongNormalizationPdf * reflPow;
  bsdf = preflect * brdfFactor;
  bsdf = bsdf * (fabsf(cosThetaOut) / cosThetaAbs);
  pdf = avg_probability(preflect) * pdfFactor;
  if (!specularOnly && cosTheta * cosThetaOut < 0.0f) {
    bsdf = bsdf + pdiffuse * m.diffuse / gg::kPI;
    pdf = pdf + avg3(pdiffuse) * gg::saturate(cosThetaOut) / gg::kPI;
  }
  const float eta = cosTheta < 0.0f ? 1.0f / m.refractionIndexAvg : m.refractionIndexAvg;
  const float etaSq = eta * eta;
  const float sinTheta2Sq = etaSq * (1.0f - cosTheta * cosTheta);
  if (sinTheta2Sq < 1.0f) {
    rrd = gg::normalize(eta * inc + (signf_(cosTheta) * (sqrtf(gg::saturate(1.0f - sinTheta2Sq)) - eta * cosThetaAbs)) * N);
    const float refrPow = gg::powf_(gg::saturate(gg::dot(rrd, exc)), m.refl[3]);
    brdfFactor = phongFactorBrdf * refrPow;
    pdfFactor = phongNormalizationPdf * refrPow;
    if (!adjoint) brdfFactor = brdfFactor / etaSq;
    brdfFactor = brdfFactor * (fabsf(gg::dot(N, exc)) / fabsf(gg::dot(N, rrd)));
  }
  bsdf = bsdf + prefract * brdfFactor;
  pdf = pdf + avg_probability(prefract) * pdfFactor;
  return bsdf;
}

__device__ __forceinline__ gg::V3 bsdf_eye(gg::V3 inc, gg::V3 exc, const MaterialData& m, gg::V3 N, float& pdf) {
  return eval_bsdf(inc, exc, m, N, pdf, false, false);
}
// AdjointBSDF (materials.glsl:405-408).
__device__ __forceinline__ gg::V3 bsdf_adjoint(gg::V3 inc, gg::V3 exc, const MaterialData& m, gg::V3 N, float& pdf) {
  return eval_bsdf(-exc, -inc, m, N, pdf, true, false);
}

// AdjointBSDFShadingNormalCorrectedOutDotN (materials.glsl:415-426).
__device__ __forceinline__ float adjoint_corrected_out_dot_n(gg::V3 incomingLight, gg::V3 toCamera, gg::V3 gN,
                                                             gg::V3 sN) {
  if (gg::dot(incomingLight, gN) * gg::dot(incomingLight, sN) < 0.0f ||
      gg::dot(toCamera, gN) * gg::dot(toCamera, sN) < 0.0f)
    return 0.0f;
  return gg::saturate(fabsf(gg::dot(incomingLight, sN) * gg::dot(toCamera, gN)) /
                      fmaxf(fabsf(gg::dot(incomingLight, gN)), gg::kDivisorEpsilon));
}
// AdjointBSDFShadingNormalCorrection (materials.glsl:427-436).
__device__ __forceinline__ float adjoint_correction(gg::V3 inDir, gg::V3 outDir, gg::V3 gN, gg::V3 sN) {
  if (gg::dot(inDir, gN) * gg::dot(inDir, sN) < 0.0f || gg::dot(outDir, gN) * gg::dot(outDir, sN) < 0.0f)
    return 0.0f;
  return gg::saturate(fabsf(gg::dot(inDir, sN) * gg::dot(outDir, gN)) /
                      fmaxf(fabsf(gg::dot(inDir, gN) * gg::dot(outDir, sN)), gg::kDivisorEpsilon));
}

// GetTriangleHitInfo (scenedata.glsl:94-105): interpolated shading normal.
__device__ __forceinline__ gg::V3 shading_normal(const float4* __restrict__ vinfo, int4 tri, gg::V3 b) {
  const float4 a = vinfo[tri.x], c = vinfo[tri.y], e = vinfo[tri.z];
  return gg::normalize(gg::unpack_normal(a.x, a.y) * b.x + gg::unpack_normal(c.x, c.y) * b.y +
                       gg::unpack_normal(e.x, e.y) * b.z);
}
__device__ __forceinline__ float2 hit_texcoord(const float4* __restrict__ vinfo, int4 tri, gg::V3 b) {
  const float4 a = vinfo[tri.x], c = vinfo[tri.y], e = vinfo[tri.z];
  return make_float2(a.z * b.x + c.z * b.y + e.z * b.z, a.w * b.x + c.w * b.y + e.w * b.z);
}

// One texture sample of a textured material slot, kept out of line (only
// scenes with textures call it); a float4 comes back in registers, so the
// caller's MaterialData never has to live in memory.
__device__ __noinline__ float4 sample_material_slot(MaterialTextures tx, int texture, float2 uv) {
  return sample_texture(tx, texture, uv);
}

// SampleMaterialData (materials.glsl:19-30) at texture coordinate uv for a
// scene with textures: the material record's constants, with each textured
// slot sampled instead.
__device__ __forceinline__ MaterialData load_material_textured(const gpugi_material* __restrict__ m,
                                                               const MaterialTextures& tx, int id, float2 uv) {
  MaterialData d = load_material(m, id);
  const int4 t = tx.matTex[id];
  if (t.x >= 0) {
    const float4 c = sample_material_slot(tx, t.x, uv);
    d.diffuse = gg::v3(c.x, c.y, c.z);
  }
  if (t.y >= 0) {
    const float4 c = sample_material_slot(tx, t.y, uv);
    d.opacity = gg::v3(c.x, c.y, c.z);
  }
  if (t.z >= 0) {
    const float4 c = sample_material_slot(tx, t.z, uv);
    d.refl[0] = c.x; d.refl[1] = c.y; d.refl[2] = c.z; d.refl[3] = c.w;
  }
  if (t.w >= 0) {
    const float4 c = sample_material_slot(tx, t.w, uv);
    d.emissivity = gg::v3(c.x, c.y, c.z);
  }
  return d;
}

// SampleMaterialData (materials.glsl:19-30) at texture coordinate uv.
__device__ __forceinline__ MaterialData load_material_uv(const gpugi_material* __restrict__ m,
                                                         const MaterialTextures& tx, int id, float2 uv) {
  if (!tx.matTex) return load_material(m, id);
  return load_material_textured(m, tx, id, uv);
}

// The same at a hit: the texture coordinate of GetTriangleHitInfo
// (scenedata.glsl:94-105), computed only when the scene has textures.
__device__ __forceinline__ MaterialData load_material_hit(const gpugi_material* __restrict__ m,
                                                          const MaterialTextures& tx,
                                                          const float4* __restrict__ vinfo, int4 tri, gg::V3 b) {
  if (!tx.matTex) return load_material(m, tri.w);
  return load_material_uv(m, tx, tri.w, hit_texcoord(vinfo, tri, b));
}

}  // namespace gpugi
