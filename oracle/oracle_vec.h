/* FP32 vector helpers of the CPU oracle (TEST INFRASTRUCTURE ONLY).
 * GLSL semantics with a fixed evaluation order: dot sums left to right,
 * normalize divides by the length, min/max return the non-NaN operand. */
#ifndef GPUGI_ORACLE_VEC_H
#define GPUGI_ORACLE_VEC_H

#include <math.h>
#include <stdint.h>
#include <stdlib.h>

#include "oracle.h"

typedef struct { float x, y, z; } vec3;

static inline vec3 v3(float x, float y, float z) { vec3 r = {x, y, z}; return r; }
static inline vec3 vadd(vec3 a, vec3 b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline vec3 vsub(vec3 a, vec3 b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline vec3 vmul(vec3 a, vec3 b) { return v3(a.x * b.x, a.y * b.y, a.z * b.z); }
static inline vec3 vneg(vec3 a) { return v3(-a.x, -a.y, -a.z); }
static inline vec3 vscale(float s, vec3 a) { return v3(s * a.x, s * a.y, s * a.z); }
static inline vec3 vscalev(vec3 a, float s) { return v3(a.x * s, a.y * s, a.z * s); }
static inline vec3 vdiv(vec3 a, float s) { return v3(a.x / s, a.y / s, a.z / s); }
static inline vec3 vdivv(vec3 a, vec3 b) { return v3(a.x / b.x, a.y / b.y, a.z / b.z); }
static inline float vdot(vec3 a, vec3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline vec3 vcross(vec3 a, vec3 b) {
  return v3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
static inline vec3 vnormalize(vec3 a) { float l = sqrtf(vdot(a, a)); return vdiv(a, l); }
static inline vec3 vmin3(vec3 a, vec3 b) { return v3(fminf(a.x, b.x), fminf(a.y, b.y), fminf(a.z, b.z)); }
static inline vec3 vmax3(vec3 a, vec3 b) { return v3(fmaxf(a.x, b.x), fmaxf(a.y, b.y), fmaxf(a.z, b.z)); }
static inline float fsat(float x) { return fminf(fmaxf(x, 0.0f), 1.0f); }
static inline vec3 vsat(vec3 a) { return v3(fsat(a.x), fsat(a.y), fsat(a.z)); }

#define OR_PI 3.14159265358979f
#define OR_PI_2 6.28318530717958f

/* CreateONB (helper.glsl:22-29) */
static inline void or_create_onb(vec3 n, vec3* U, vec3* V) {
  vec3 u = vcross(n, v3(0.0f, 1.0f, 0.0f));
  if (fabsf(u.x) < 0.0001f && fabsf(u.y) < 0.0001f && fabsf(u.z) < 0.0001f) u = vcross(n, v3(1.0f, 0.0f, 0.0f));
  *U = vnormalize(u);
  *V = vcross(n, *U);
}
/* SampleHemisphereCosine (random.glsl:62-71) */
static inline vec3 or_hemisphere_cosine(float r0, float r1, vec3 U, vec3 V, vec3 W) {
  float phi = OR_PI_2 * r0;
  float sinTheta = sqrtf(r1);
  float x = sinTheta * or_cosf(phi);
  float y = sinTheta * or_sinf(phi);
  float z = sqrtf(1.0f - r1);
  return vadd(vadd(vscale(x, U), vscale(y, V)), vscale(z, W));
}
/* SamplePhongLobe (random.glsl:75-87) */
static inline vec3 or_phong_lobe(float r0, float r1, float exponent, vec3 U, vec3 V, vec3 W) {
  float phi = OR_PI_2 * r0;
  float power = or_powf(r1, 1.0f / (exponent + 1.0f));
  float scale = sqrtf(1.0f - power * power);
  float x = or_cosf(phi) * scale;
  float y = or_sinf(phi) * scale;
  float z = power;
  return vadd(vadd(vscale(x, U), vscale(y, V)), vscale(z, W));
}
/* SampleDirection (random.glsl:52-58) */
static inline vec3 or_sample_direction(float r0, float r1) {
  float cosTheta = r0 * 2.0f - 1.0f;
  float sinTheta = sqrtf((1.0f - cosTheta) * (1.0f + cosTheta));
  float phi = r1 * OR_PI_2;
  return vnormalize(v3(sinTheta * or_sinf(phi), sinTheta * or_cosf(phi), cosTheta));
}
/* UnpackNormal(vec2) (helper.glsl:38-42) */
static inline vec3 or_unpack_normal(float phi, float z) {
  float sinPhi = sqrtf(fsat((1.0f - z) * (1.0f + z)));
  return v3(or_cosf(phi) * sinPhi, or_sinf(phi) * sinPhi, z);
}
/* UnpackNormal(float, float) (helper.glsl:43-47) */
static inline vec3 or_unpack_normal_raw(float phi, float z) {
  float sinPhi = sqrtf(1.0f - z * z);
  return v3(or_cosf(phi) * sinPhi, or_sinf(phi) * sinPhi, z);
}
/* PackNormal (helper.glsl:51-63) */
static inline void or_pack_normal(vec3 n, float* a, float* b) {
  if (n.x == 0.0f) *a = (n.y > 0.0f ? 1.0f : (n.y < 0.0f ? -1.0f : 0.0f)) * OR_PI / 2.0f;
  else *a = or_atan2f(n.y, n.x);
  *b = n.z;
}
/* GetLuminance (helper.glsl:31-35) */
static inline float or_luminance(vec3 c) { return vdot(c, v3(0.212671f, 0.715160f, 0.072169f)); }
/* pathtracer.comp:39-46 */
static inline vec3 or_primary_dir(float j0, float j1, int px, int py, int w, int h, vec3 U, vec3 V, vec3 W) {
  float sx = (((j0 - 0.5f) + (float)px) / (float)w) * 2.0f - 1.0f;
  float sy = (((j1 - 0.5f) + (float)py) / (float)h) * 2.0f - 1.0f;
  return vnormalize(vadd(vadd(vscale(sx, U), vscale(sy, V)), W));
}

typedef struct {
  float refl[4];
  vec3 opacity, diffuse, emissivity, fresnel0;
  float rIndex;
  vec3 fresnel1;
} or_material;

or_material or_mat_of(const or_scene* s, int id);
vec3 or_sample_bsdf(vec3 inc, const or_material* m, uint32_t* seed, vec3 N, vec3* tp, int adjoint, float* pdf,
                    int* diffuseOut);
vec3 or_bsdf(vec3 inc, vec3 exc, const or_material* m, vec3 N, float* pdf, int adjoint, int specularOnly);
vec3 or_env_lookup(const float* tex, int size, vec3 d);
or_material or_mat_at(const or_scene* s, int id, float u, float v);
void or_hit_texcoord(const or_scene* s, const int32_t* tri, vec3 b, float* u, float* v);
vec3 or_shading_normal(const or_scene* s, const int32_t* tri, vec3 b);
int or_any_hit(const or_scene* s, vec3 o, vec3 d, float len);

int or_triangle(vec3 o, vec3 d, vec3 p0, vec3 p1, vec3 p2, float* t, vec3* b, vec3* n);
typedef struct {
  int kx, ky, kz;
  float sx, sy, sz;
} or_wtray;
or_wtray or_wt_setup(vec3 d);
int or_triangle_wt(vec3 o, const or_wtray* r, vec3 p0, vec3 p1, vec3 p2, float* t, vec3* b);
void or_count_hits(const or_scene* s, vec3 o, vec3 d, float rayLength, float w, float* importance);
int or_trace_one(const or_scene* s, int any, vec3 o, vec3 d, float* rayLength, int32_t* slot, vec3* bary,
                 vec3* triNormal, int32_t tri[4], uint32_t* nbox, uint32_t* ntri);

#endif
