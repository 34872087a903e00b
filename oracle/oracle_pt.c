/*
 * CPU oracle: materials, BSDF, next-event estimation and the path tracer.
 * TEST INFRASTRUCTURE ONLY (see oracle.h).  Restates materials.glsl,
 * lightsource.glsl, scenedata.glsl:94-105 and pathtracer.comp.
 */
#include <string.h>

#include "oracle.h"
#include "oracle_vec.h"

or_material or_mat_of(const or_scene* s, int id) {
  const float* m = s->materials + (size_t)id * 20;
  or_material r;
  r.diffuse = v3(m[0], m[1], m[2]);
  r.opacity = v3(m[3], m[4], m[5]);
  r.refl[0] = m[6]; r.refl[1] = m[7]; r.refl[2] = m[8]; r.refl[3] = m[9];
  r.emissivity = v3(m[10], m[11], m[12]);
  r.fresnel0 = v3(m[13], m[14], m[15]);
  r.rIndex = m[16];
  r.fresnel1 = v3(m[17], m[18], m[19]);
  return r;
}

/* Material textures (the bindless handles of Material, globaluniforms.glsl:
 * 23-27): per material the texture of its diffuse, opacity, reflectiveness and
 * emissivity slots (-1: the record's constant); per texture its linear RGBA
 * texels and size.  NULL: no textures. */
static const int32_t* g_matTex = NULL;
static uint32_t g_numMatTex = 0;
static const float* const* g_texels = NULL;
static const int32_t* g_texSize = NULL;

void or_set_textures(const int32_t* matTex, uint32_t numMaterials, const float* const* texels, const int32_t* sizes) {
  g_matTex = matTex;
  g_numMatTex = numMaterials;
  g_texels = texels;
  g_texSize = sizes;
}

/* textureLod(uv, 0) with the scene's sampler (scene.cpp:16-20): bilinear over
 * texel centres, REPEAT; a NaN or infinite coordinate gives 0. */
static void tex_sample(int id, float uu, float vv, float out[4]) {
  const int w = g_texSize[2 * id], h = g_texSize[2 * id + 1];
  float u = uu * (float)w - 0.5f, v = vv * (float)h - 0.5f;
  float fu = floorf(u), fv = floorf(v);
  float a = u - fu, b = v - fv;
  if (!(a >= 0.0f && a <= 1.0f && b >= 0.0f && b <= 1.0f)) {
    out[0] = out[1] = out[2] = out[3] = 0.0f;
    return;
  }
  float wx = fmodf(fu, (float)w), wy = fmodf(fv, (float)h);
  if (wx < 0.0f) wx += (float)w;
  if (wy < 0.0f) wy += (float)h;
  int x0 = (int)wx, y0 = (int)wy;
  x0 = x0 < 0 ? 0 : (x0 > w - 1 ? w - 1 : x0);
  y0 = y0 < 0 ? 0 : (y0 > h - 1 ? h - 1 : y0);
  int x1 = x0 + 1 == w ? 0 : x0 + 1, y1 = y0 + 1 == h ? 0 : y0 + 1;
  const float* t = g_texels[id];
  const float* t00 = t + ((size_t)y0 * w + x0) * 4;
  const float* t10 = t + ((size_t)y0 * w + x1) * 4;
  const float* t01 = t + ((size_t)y1 * w + x0) * 4;
  const float* t11 = t + ((size_t)y1 * w + x1) * 4;
  float ia = 1.0f - a, ib = 1.0f - b;
  for (int c = 0; c < 4; ++c) out[c] = (t00[c] * ia + t10[c] * a) * ib + (t01[c] * ia + t11[c] * a) * b;
}

/* SampleMaterialData (materials.glsl:19-30) at texture coordinate (u, v). */
or_material or_mat_at(const or_scene* s, int id, float u, float v) {
  or_material r = or_mat_of(s, id);
  if (!g_matTex || (uint32_t)id >= g_numMatTex) return r;
  const int32_t* t = g_matTex + (size_t)id * 4;
  float c[4];
  if (t[0] >= 0) { tex_sample(t[0], u, v, c); r.diffuse = v3(c[0], c[1], c[2]); }
  if (t[1] >= 0) { tex_sample(t[1], u, v, c); r.opacity = v3(c[0], c[1], c[2]); }
  if (t[2] >= 0) { tex_sample(t[2], u, v, c); r.refl[0] = c[0]; r.refl[1] = c[1]; r.refl[2] = c[2]; r.refl[3] = c[3]; }
  if (t[3] >= 0) { tex_sample(t[3], u, v, c); r.emissivity = v3(c[0], c[1], c[2]); }
  return r;
}

/* The texture coordinate of GetTriangleHitInfo (scenedata.glsl:94-105). */
void or_hit_texcoord(const or_scene* s, const int32_t* tri, vec3 b, float* u, float* v) {
  const float* va = s->vertex_info + (size_t)tri[0] * 4;
  const float* vb = s->vertex_info + (size_t)tri[1] * 4;
  const float* vc = s->vertex_info + (size_t)tri[2] * 4;
  *u = va[2] * b.x + vb[2] * b.y + vc[2] * b.z;
  *v = va[3] * b.x + vb[3] * b.y + vc[3] * b.z;
}

/* AvgProbability / Avg (materials.glsl:33-41) */
static float avg_p(vec3 c) { return (c.x + c.y + c.z + 1e-5f) / (3.0f + 1e-5f); }
static float avg(vec3 c) { return (c.x + c.y + c.z) / 3.0f; }
static float sgn(float x) { return x > 0.0f ? 1.0f : (x < 0.0f ? -1.0f : 0.0f); }

/* GetBSDFDecisionPropabilities, USE_CONSTANT_REFLECTION (materials.glsl:43-63) */
static void decision(const or_material* m, vec3* pr, vec3* pt, vec3* pd) {
  *pr = m->fresnel0;
  vec3 t = vsub(v3(1.0f, 1.0f, 1.0f), vsat(*pr));
  *pd = vmul(t, m->opacity);
  *pt = vsub(t, *pd);
}

/* __SampleBSDF (materials.glsl:118-277) */
vec3 or_sample_bsdf(vec3 inc, const or_material* m, uint32_t* seed, vec3 N, vec3* tp, int adjoint, float* pdf,
                    int* diffuseOut) {
  float cosTheta = vdot(N, inc);
  float cosThetaAbs = fsat(fabsf(cosTheta));
  vec3 rrd = vsub(inc, vscale(2.0f * cosTheta, N));
  float r0 = or_random01(seed);
  float r1 = or_random01(seed);
  vec3 pr, pt, pd;
  decision(m, &pr, &pt, &pd);
  float aR = avg_p(pr), aT = avg_p(pt), aD = avg(pd);
  vec3 pphong = pr;
  float aPhong = aR;
  float xi = or_random01(seed);
  if (aD > xi) {
    *diffuseOut = 1;
    *tp = vmul(*tp, vdiv(vmul(m->diffuse, pd), aD));
    if (cosTheta > 0.0f) N = vneg(N);
    vec3 U, V;
    or_create_onb(N, &U, &V);
    vec3 out = or_hemisphere_cosine(r0, r1, U, V, N);
    *pdf = fsat(vdot(N, out)) / OR_PI;
    return out;
  } else if (aD + aT > xi) {
    pphong = pt;
    aPhong = aT;
    float eta = cosTheta < 0.0f ? 1.0f / m->rIndex : m->rIndex;
    float etaSq = eta * eta;
    float s2 = etaSq * (1.0f - cosTheta * cosTheta);
    if (s2 < 1.0f) {
      rrd = vadd(vscale(eta, inc), vscale(sgn(cosTheta) * (sqrtf(fsat(1.0f - s2)) - eta * cosThetaAbs), N));
      if (!adjoint) *tp = vdiv(*tp, etaSq);
    }
  }
  *diffuseOut = 0;
  rrd = vnormalize(rrd);
  vec3 RU, RV;
  or_create_onb(rrd, &RU, &RV);
  vec3 out = or_phong_lobe(r0, r1, m->refl[3], RU, RV, rrd);
  if (!adjoint) *tp = vscalev(*tp, fabsf(vdot(N, out)) / fabsf(vdot(N, rrd)));
  float cAbs = fabsf(vdot(out, rrd));
  *pdf = (m->refl[3] + 1.0f) / OR_PI_2 * or_powf(cAbs, m->refl[3]);
  *pdf = *pdf * 1.0f;
  *tp = vmul(*tp, vdiv(pphong, aPhong));
  return out;
}

/* __BSDF (materials.glsl:306-391) */
vec3 or_bsdf(vec3 inc, vec3 exc, const or_material* m, vec3 N, float* pdf, int adjoint, int specularOnly) {
  float cosTheta = vdot(N, inc);
  float cosThetaAbs = fsat(fabsf(cosTheta));
  float cosThetaOut = vdot(N, exc);
  vec3 pr, pt, pd;
  decision(m, &pr, &pt, &pd);
  float fB = (m->refl[3] + 1.0f) / (2.0f * OR_PI_2 * fabsf(cosThetaOut) + 1e-5f);
  float fP = (m->refl[3] + 1.0f) / OR_PI_2;
  vec3 rrd = vnormalize(vsub(inc, vscale(2.0f * cosTheta, N)));
  float pw = or_powf(fsat(vdot(rrd, exc)), m->refl[3]);
  float bF = fB * pw, pF = fP * pw;
  vec3 bsdf = vscalev(pr, bF);
  bsdf = vscalev(bsdf, fabsf(cosThetaOut) / cosThetaAbs);
  *pdf = avg_p(pr) * pF;
  if (!specularOnly && cosTheta * cosThetaOut < 0.0f) {
    bsdf = vadd(bsdf, vdiv(vmul(pd, m->diffuse), OR_PI));
    *pdf = *pdf + avg(pd) * fsat(cosThetaOut) / OR_PI;
  }
  float eta = cosTheta < 0.0f ? 1.0f / m->rIndex : m->rIndex;
  float etaSq = eta * eta;
  float s2 = etaSq * (1.0f - cosTheta * cosTheta);
  if (s2 < 1.0f) {
    rrd = vnormalize(vadd(vscale(eta, inc), vscale(sgn(cosTheta) * (sqrtf(fsat(1.0f - s2)) - eta * cosThetaAbs), N)));
    float pw2 = or_powf(fsat(vdot(rrd, exc)), m->refl[3]);
    bF = fB * pw2;
    pF = fP * pw2;
    if (!adjoint) bF = bF / etaSq;
    bF = bF * (fabsf(vdot(N, exc)) / fabsf(vdot(N, rrd)));
  }
  bsdf = vadd(bsdf, vscalev(pt, bF));
  *pdf = *pdf + avg_p(pt) * pF;
  return bsdf;
}

/* GetTriangleHitInfo (scenedata.glsl:94-105) */
vec3 or_shading_normal(const or_scene* s, const int32_t* tri, vec3 b) {
  const float* a = s->vertex_info + (size_t)tri[0] * 4;
  const float* c = s->vertex_info + (size_t)tri[1] * 4;
  const float* e = s->vertex_info + (size_t)tri[2] * 4;
  vec3 n = vadd(vadd(vscalev(or_unpack_normal(a[0], a[1]), b.x), vscalev(or_unpack_normal(c[0], c[1]), b.y)),
                vscalev(or_unpack_normal(e[0], e[1]), b.z));
  return vnormalize(n);
}

/* Environment cube map of the path tracer (pathtracer.comp:56-58,
 * renderersystem.cpp:39-47, 173-186): NULL means the black default.  Texels are
 * linear RGB, 4 floats each, faces x-, x+, y-, y+, z-, z+. */
/* Path options shared by every integrator of the oracle (or_set_path_options). */
int or_spl = -1;
int or_rr_start = 0;

void or_set_path_options(int specificPathLength, int rrStartDepth) {
  or_spl = specificPathLength < 0 ? -1 : specificPathLength;
  or_rr_start = rrStartDepth < 0 ? 0 : rrStartDepth;
}

static const float* g_env = NULL;
static int g_envSize = 0;

void or_set_environment(const float* texels, int size) {
  g_env = size > 0 ? texels : NULL;
  g_envSize = size;
}

/* The OpenGL cube-map selection (major axis, ties to x then y; per-face
 * (sc, tc, ma) of the specification's table), s = (sc / |ma| + 1) / 2,
 * t = (tc / |ma| + 1) / 2, then bilinear filtering over texel centres with the
 * coordinates clamped to the face. */
vec3 or_env_lookup(const float* tex, int size, vec3 d) {
  float ax = fabsf(d.x), ay = fabsf(d.y), az = fabsf(d.z), sc, tc, ma;
  int face;
  if (ax >= ay && ax >= az) {
    ma = ax;
    if (d.x >= 0.0f) { face = 1; sc = -d.z; } else { face = 0; sc = d.z; }
    tc = -d.y;
  } else if (ay >= az) {
    ma = ay;
    sc = d.x;
    if (d.y >= 0.0f) { face = 3; tc = d.z; } else { face = 2; tc = -d.z; }
  } else {
    ma = az;
    if (d.z >= 0.0f) { face = 5; sc = d.x; } else { face = 4; sc = -d.x; }
    tc = -d.y;
  }
  float s = (sc / ma + 1.0f) * 0.5f, t = (tc / ma + 1.0f) * 0.5f;
  if (!(ma > 0.0f) || !(s >= 0.0f && s <= 1.0f && t >= 0.0f && t <= 1.0f)) return v3(0, 0, 0);
  float u = s * (float)size - 0.5f, v = t * (float)size - 0.5f;
  float fu = floorf(u), fv = floorf(v);
  float a = u - fu, b = v - fv, ia = 1.0f - a, ib = 1.0f - b;
  int i0 = (int)fu, j0 = (int)fv;
  int x0 = i0 < 0 ? 0 : i0, x1 = i0 + 1 > size - 1 ? size - 1 : i0 + 1;
  int y0 = j0 < 0 ? 0 : j0, y1 = j0 + 1 > size - 1 ? size - 1 : j0 + 1;
  const float* f = tex + (size_t)face * size * size * 4;
  const float* t00 = f + ((size_t)y0 * size + x0) * 4;
  const float* t10 = f + ((size_t)y0 * size + x1) * 4;
  const float* t01 = f + ((size_t)y1 * size + x0) * 4;
  const float* t11 = f + ((size_t)y1 * size + x1) * 4;
  float r[3];
  for (int c = 0; c < 3; ++c) r[c] = (t00[c] * ia + t10[c] * a) * ib + (t01[c] * ia + t11[c] * a) * b;
  return v3(r[0], r[1], r[2]);
}

int or_any_hit(const or_scene* s, vec3 o, vec3 d, float len) {
  int32_t slot;
  vec3 b;
  return or_trace_one(s, 1, o, d, &len, &slot, &b, NULL, NULL, NULL, NULL);
}

/* EstimateDirectLight (lightsource.glsl:3-93) */
static vec3 or_direct_light(const or_scene* s, const float* lights, vec3 pos, vec3 normal, int idx, vec3 view,
                            const or_material* m, int sampleArea) {
  const float* ls = lights + (size_t)idx * 8;
  vec3 dir = vsub(v3(ls[0], ls[1], ls[2]), pos);
  float distSq = vdot(dir, dir) + 1e-5f;
  float dist = sqrtf(distSq);
  dir = vdiv(dir, dist);
  float surfaceCos = fsat(vdot(dir, normal));
  if (surfaceCos <= 0.0f) return v3(0, 0, 0);
  vec3 I = v3(ls[4], ls[5], ls[6]);
  if (ls[7] > 1.5f) {
    vec3 o = vadd(vscale(0.0001f, dir), pos);
    if (!or_any_hit(s, o, dir, dist - 0.0001f * 2.0f)) {
      vec3 irr = vscale(surfaceCos / distSq, I);
      float pdf;
      vec3 f = or_bsdf(view, dir, m, normal, &pdf, 0, 0);
      return vmul(irr, f);
    }
  } else if (sampleArea) {
    vec3 ln = or_unpack_normal(ls[3], ls[7]);
    float k = fsat(vdot(ln, vneg(dir)));
    if (k > 0.0f) {
      vec3 o = vadd(vscale(0.0001f, dir), pos);
      if (!or_any_hit(s, o, dir, dist - 0.0001f * 2.0f)) {
        float pdf;
        vec3 li = vscalev(I, k);
        vec3 f = or_bsdf(view, dir, m, normal, &pdf, 0, 0);
        vec3 irr = vscale(surfaceCos / distSq, li);
        return vmul(irr, f);
      }
    }
  }
  return v3(0, 0, 0);
}

/* pathtracer.comp:16-110, rows [row_begin, row_end) of the truncated (W/8*8) x (H/8*8) grid */
void or_pathtracer(const or_scene* s, const float* lights, uint32_t numLights, const or_camera* c, int W, int H,
                   uint32_t frameSeed, int maxLen, float* accum, int row_begin, int row_end) {
  const int gw = (W / 8) * 8, gh = (H / 8) * 8;
  vec3 U = v3(c->u[0], c->u[1], c->u[2]), V = v3(c->v[0], c->v[1], c->v[2]), Wd = v3(c->w[0], c->w[1], c->w[2]);
  const float lsw = (float)numLights / 1.0f;
  if (row_end > gh) row_end = gh;
  for (int y = row_begin; y < row_end; ++y) {
    for (int x = 0; x < gw; ++x) {
      uint32_t seed = or_wanghash(frameSeed + (uint32_t)(x + y * W));
      vec3 color = v3(0, 0, 0), thr = v3(1, 1, 1);
      int didNEE = 0;
      float j0 = or_random01(&seed), j1 = or_random01(&seed);
      vec3 o = v3(c->position[0], c->position[1], c->position[2]);
      vec3 d = or_primary_dir(j0, j1, x, y, W, H, U, V, Wd);
      for (int i = 0; i < maxLen; ++i) {
        float t = 3.402823466e+38f;
        int32_t slot, tri[4];
        vec3 b, gn;
        or_trace_one(s, 0, o, d, &t, &slot, &b, &gn, tri, NULL, NULL);
        if (t == 3.402823466e+38f) {
          color = vadd(color, vmul(thr, g_env ? or_env_lookup(g_env, g_envSize, d) : v3(0, 0, 0)));
          break;
        }
        o = vadd(o, vscale(t, d));
        vec3 n = or_shading_normal(s, tri, b);
        float tu, tv;
        or_hit_texcoord(s, tri, b, &tu, &tv);
        or_material m = or_mat_at(s, tri[3], tu, tv);
        /* SHOW_SPECIFIC_PATHLENGTH (pathtracer.comp:74-84): emission at vertex i
         * closes a path of length i, the light sample of vertex i one of length i+1 */
        if (or_spl < 0 || i == or_spl)
          if (!didNEE) color = vadd(color, vmul(thr, m.emissivity));
        didNEE = m.refl[3] < 500.0f;
        if (or_spl < 0 || i + 1 == or_spl) {
          int idx = (int)(or_xorshift32(&seed) % numLights);
          vec3 rad = or_direct_light(s, lights, o, n, idx, d, &m, didNEE);
          color = vadd(color, vscalev(vmul(rad, thr), lsw));
        }
        vec3 tp = v3(1, 1, 1);
        float pdf;
        int dif;
        d = or_sample_bsdf(d, &m, &seed, n, &tp, 0, &pdf, &dif);
        o = vadd(o, vscalev(d, 0.0001f));
        thr = vmul(thr, tp);
        if (i >= or_rr_start) { /* Russian roulette (pathtracer.comp:99-104) from or_rr_start on */
          float q = fsat(or_luminance(tp));
          if (or_random01(&seed) >= q) break;
          thr = vdivv(thr, v3(q, q, q));
        }
      }
      float* px = accum + ((size_t)y * W + x) * 4;
      px[0] = color.x + px[0];
      px[1] = color.y + px[1];
      px[2] = color.z + px[2];
      px[3] = 0.0f + px[3];
    }
  }
}

/* Batched probes of the two BSDF routines for the consistency tests: mat20 is
 * one 20-float material record (the scene layout), N the shading normal.
 * or_bsdf_samples: n draws of or_sample_bsdf for the incident direction `inc`
 * from the stream seed, writing direction (3), throughput (3), pdf and the
 * diffuse flag as 8 floats per draw.  or_bsdf_evals: or_bsdf for n excident
 * directions (3 floats each), writing the BSDF (3) and pdf as 4 floats. */
void or_bsdf_samples(const float* mat20, const float* inc, const float* N, uint32_t seed, int adjoint, size_t n,
                     float* out) {
  or_material m;
  m.diffuse = v3(mat20[0], mat20[1], mat20[2]);
  m.opacity = v3(mat20[3], mat20[4], mat20[5]);
  m.refl[0] = mat20[6]; m.refl[1] = mat20[7]; m.refl[2] = mat20[8]; m.refl[3] = mat20[9];
  m.emissivity = v3(mat20[10], mat20[11], mat20[12]);
  m.fresnel0 = v3(mat20[13], mat20[14], mat20[15]);
  m.rIndex = mat20[16];
  m.fresnel1 = v3(mat20[17], mat20[18], mat20[19]);
  const vec3 vi = v3(inc[0], inc[1], inc[2]), vn = v3(N[0], N[1], N[2]);
  for (size_t i = 0; i < n; ++i) {
    vec3 tp = v3(1, 1, 1);
    float pdf;
    int dif;
    const vec3 d = or_sample_bsdf(vi, &m, &seed, vn, &tp, adjoint, &pdf, &dif);
    float* o = out + i * 8;
    o[0] = d.x; o[1] = d.y; o[2] = d.z;
    o[3] = tp.x; o[4] = tp.y; o[5] = tp.z;
    o[6] = pdf;
    o[7] = (float)dif;
  }
}

void or_bsdf_evals(const float* mat20, const float* inc, const float* N, int adjoint, const float* exc, size_t n,
                   float* out) {
  or_material m;
  m.diffuse = v3(mat20[0], mat20[1], mat20[2]);
  m.opacity = v3(mat20[3], mat20[4], mat20[5]);
  m.refl[0] = mat20[6]; m.refl[1] = mat20[7]; m.refl[2] = mat20[8]; m.refl[3] = mat20[9];
  m.emissivity = v3(mat20[10], mat20[11], mat20[12]);
  m.fresnel0 = v3(mat20[13], mat20[14], mat20[15]);
  m.rIndex = mat20[16];
  m.fresnel1 = v3(mat20[17], mat20[18], mat20[19]);
  const vec3 vi = v3(inc[0], inc[1], inc[2]), vn = v3(N[0], N[1], N[2]);
  for (size_t i = 0; i < n; ++i) {
    float pdf;
    const vec3 f = or_bsdf(vi, v3(exc[3 * i], exc[3 * i + 1], exc[3 * i + 2]), &m, vn, &pdf, adjoint, 0);
    float* o = out + i * 4;
    o[0] = f.x; o[1] = f.y; o[2] = f.z;
    o[3] = pdf;
  }
}

void or_env_lookups(const float* tex, int size, const float* dirs, size_t n, float* out) {
  for (size_t i = 0; i < n; ++i) {
    vec3 r = or_env_lookup(tex, size, v3(dirs[3 * i], dirs[3 * i + 1], dirs[3 * i + 2]));
    out[3 * i] = r.x;
    out[3 * i + 1] = r.y;
    out[3 * i + 2] = r.z;
  }
}

void or_texture_samples(int id, const float* uvs, size_t n, float* out) {
  for (size_t i = 0; i < n; ++i) tex_sample(id, uvs[2 * i], uvs[2 * i + 1], out + 4 * i);
}
