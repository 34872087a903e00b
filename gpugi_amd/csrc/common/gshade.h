// Sampling and encoding helpers of the reference renderer, shared by host C++
// and HIP device code (compiled with -ffp-contract=off, see gmath.h).
#pragma once
#include "gvec.h"

namespace gg {

constexpr float kRayHitEpsilon = 0.0001f;          // stdheader.glsl:5 RAY_HIT_EPSILON
constexpr float kRayMax = 3.402823466e+38f;        // stdheader.glsl:6 RAY_MAX
constexpr float kIntersectEpsilon = 0.0001f;       // intersectiontests.glsl:1
constexpr float kDivisorEpsilon = 1e-5f;           // materials.glsl:1 DIVISOR_EPSILON

// CreateONB (helper.glsl:22-29): fall back to the x axis when n is (anti)parallel to y.
GG_HD void create_onb(V3 n, V3& U, V3& V) {
  U = cross(n, v3(0.0f, 1.0f, 0.0f));
  if (fabsf(U.x) < 0.0001f && fabsf(U.y) < 0.0001f && fabsf(U.z) < 0.0001f)
    U = cross(n, v3(1.0f, 0.0f, 0.0f));
  U = normalize(U);
  V = cross(n, U);
}

// GetLuminance (helper.glsl:31-35).
GG_HD float luminance(V3 c) { return dot(c, v3(0.212671f, 0.715160f, 0.072169f)); }

// UnpackNormal(vec2(phi, z)) (helper.glsl:38-42).
GG_HD V3 unpack_normal(float phi, float z) {
  float sinPhi = sqrtf(saturate((1.0f - z) * (1.0f + z)));
  float s, c;
  sincos_core(phi, &s, &c);
  return v3(c * sinPhi, s * sinPhi, z);
}
// UnpackNormal(float, float) (helper.glsl:43-47): no saturate.
GG_HD V3 unpack_normal_raw(float phi, float z) {
  float sinPhi = sqrtf(1.0f - z * z);
  float s, c;
  sincos_core(phi, &s, &c);
  return v3(c * sinPhi, s * sinPhi, z);
}
// PackNormal (helper.glsl:51-63): (atan2(y, x) or sign(y)*PI/2 when x == 0, z).
GG_HD void pack_normal(V3 n, float& a, float& b) {
  if (n.x == 0.0f) {
    float sgn = n.y > 0.0f ? 1.0f : (n.y < 0.0f ? -1.0f : 0.0f);
    a = sgn * kPI / 2.0f;
  } else {
    a = atan2f_(n.y, n.x);
  }
  b = n.z;
}

// SampleDirection (random.glsl:52-58): uniform sphere.
GG_HD V3 sample_direction(float r0, float r1) {
  float cosTheta = r0 * 2.0f - 1.0f;
  float sinTheta = sqrtf((1.0f - cosTheta) * (1.0f + cosTheta));
  float phi = r1 * kPI_2;
  float s, c;
  sincos_core(phi, &s, &c);
  return normalize(v3(sinTheta * s, sinTheta * c, cosTheta));
}

// SampleHemisphereCosine (random.glsl:62-71).
GG_HD V3 sample_hemisphere_cosine(float r0, float r1, V3 U, V3 V, V3 W) {
  float phi = kPI_2 * r0;
  float sinTheta = sqrtf(r1);
  float s, c;
  sincos_core(phi, &s, &c);
  float x = sinTheta * c;
  float y = sinTheta * s;
  float z = sqrtf(1.0f - r1);
  return x * U + y * V + z * W;
}

// SamplePhongLobe (random.glsl:75-87).
GG_HD V3 sample_phong_lobe(float r0, float r1, float exponent, V3 U, V3 V, V3 W) {
  float phi = kPI_2 * r0;
  float power = powf_(r1, 1.0f / (exponent + 1.0f));
  float scale = sqrtf(1.0f - power * power);
  float s, c;
  sincos_core(phi, &s, &c);
  float x = c * scale;
  float y = s * scale;
  float z = power;
  return x * U + y * V + z * W;
}

// Primary ray direction of pathtracer.comp:39-46 for pixel (px, py) and a jitter pair.
GG_HD V3 primary_direction(float j0, float j1, int px, int py, int width, int height, V3 U, V3 V, V3 W) {
  float sx = (((j0 - 0.5f) + (float)px) / (float)width) * 2.0f - 1.0f;
  float sy = (((j1 - 0.5f) + (float)py) / (float)height) * 2.0f - 1.0f;
  return normalize(sx * U + sy * V + W);
}

}  // namespace gg
