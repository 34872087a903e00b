// Deterministic FP32 math shared by the host C++ code and the HIP kernels.
//
// Every function here is written with plain IEEE-754 single-precision
// add/sub/mul/div/sqrt/floor and integer bit operations only, and the code that
// includes it is compiled with -ffp-contract=off.  The result is therefore the
// same bit pattern on the x86-64 host (SSE) and on gfx950, which is what lets
// a device render be compared per pixel against a CPU restatement.
//
// The transcendental kernels follow the published Cephes single-precision
// algorithms (S. L. Moshier, "Methods and Programs for Mathematical
// Functions", 1989): Cody-Waite reduction plus short minimax polynomials.
// Accuracy is a few ulp over the ranges the renderer uses (checked in tests
// against libm), which is well inside what GLSL allows for sin/cos/pow.
#pragma once

#include <stdint.h>
#include <string.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define GG_HD __host__ __device__ __forceinline__
#else
#include <math.h>
#define GG_HD static inline
#endif

namespace gg {

GG_HD uint32_t f2u(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }
GG_HD float u2f(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }
// min/max with IEEE minNum/maxNum semantics (a NaN operand yields the other
// operand): fminf/fmaxf on the host, v_min_f32/v_max_f32 on gfx950.
GG_HD float fminf_(float a, float b) { return fminf(a, b); }
GG_HD float fmaxf_(float a, float b) { return fmaxf(a, b); }
GG_HD float saturate(float x) { return fminf(fmaxf(x, 0.0f), 1.0f); }

constexpr float kPI = 3.14159265358979f;     // helper.glsl:1 PI
constexpr float kPI_2 = 6.28318530717958f;   // helper.glsl:2 PI_2 (= 2*pi)

// sin/cos: argument reduction by pi/2 with a three-part constant
// (12 + 12 + 24 significant bits; k*P1 and k*P2 are exact for |k| < 4096).
GG_HD void sincos_core(float x, float* s, float* c) {
  const float kTwoOverPi = 0.636619746685028076171875f;
  const float P1 = 1.5703125f;
  const float P2 = 4.837512969970703125e-4f;
  const float P3 = 7.54978995489188216e-8f;
  float k = floorf(x * kTwoOverPi + 0.5f);
  float r = ((x - k * P1) - k * P2) - k * P3;
  float z = r * r;
  float sr = ((-1.9515295891e-4f * z + 8.3321608736e-3f) * z - 1.6666654611e-1f) * z * r + r;
  float cr = ((2.443315711809948e-5f * z - 1.388731625493765e-3f) * z + 4.166664568298827e-2f) * z * z
             - 0.5f * z + 1.0f;
  int q = ((int)k) & 3;
  float so, co;
  if (q == 0) { so = sr; co = cr; }
  else if (q == 1) { so = cr; co = -sr; }
  else if (q == 2) { so = -sr; co = -cr; }
  else { so = -cr; co = sr; }
  *s = so; *c = co;
}
GG_HD float sinf_(float x) { float s, c; sincos_core(x, &s, &c); return s; }
GG_HD float cosf_(float x) { float s, c; sincos_core(x, &s, &c); return c; }

// log2 via x = m * 2^e, m in [sqrt(1/2), sqrt(2)), ln(1+f) polynomial.
GG_HD float log2f_(float x) {
  if (x != x) return x;
  if (x < 0.0f) return u2f(0x7fc00000u);
  if (x == 0.0f) return u2f(0xff800000u);
  if (x == u2f(0x7f800000u)) return x;
  int e = 0;
  if (x < 1.17549435e-38f) { x = x * 8388608.0f; e = -23; }
  uint32_t u = f2u(x);
  e += (int)((u >> 23) & 0xff) - 126;
  float m = u2f((u & 0x007fffffu) | 0x3f000000u);  // [0.5, 1)
  if (m < 0.70710678118654752440f) { e -= 1; m = m + m - 1.0f; }
  else { m = m - 1.0f; }
  float z = m * m;
  float y = ((((((((7.0376836292e-2f * m - 1.1514610310e-1f) * m + 1.1676998740e-1f) * m
              - 1.2420140846e-1f) * m + 1.4249322787e-1f) * m - 1.6668057665e-1f) * m
              + 2.0000714765e-1f) * m - 2.4999993993e-1f) * m + 3.3333331174e-1f) * m * z;
  y = y - 0.5f * z;
  const float kLog2e = 1.44269504088896341f;
  // log2(1+m) = (m + y) * log2(e), split so that the large part keeps precision
  float r = y * kLog2e + m * kLog2e;
  return r + (float)e;
}

// 2^x: x = i + f, |f| <= 1/2, polynomial for 2^f, exponent by bit construction.
GG_HD float exp2f_(float x) {
  if (x != x) return x;
  if (x > 128.0f) return u2f(0x7f800000u);
  if (x < -150.0f) return 0.0f;
  float i = floorf(x + 0.5f);
  float f = x - i;
  float p = ((((( 1.535336188319500e-4f * f + 1.339887440266574e-3f) * f + 9.618437357674640e-3f) * f
              + 5.550332471162809e-2f) * f + 2.402264791363012e-1f) * f + 6.931472028550421e-1f) * f + 1.0f;
  int n = (int)i;                                   // n in [-150, 128]
  if (n > 127) return (p * u2f(0x7f000000u)) * 2.0f;  // may overflow to inf
  if (n < -126)                                      // denormal result: one rounding at the end
    return (p * u2f((uint32_t)(n + 126 + 127) << 23)) * u2f(0x00800000u);
  return p * u2f((uint32_t)(n + 127) << 23);
}

// pow(x, y) = 2^(y * log2 x), the GLSL definition (GLSL 4.50 spec 8.2).
GG_HD float powf_(float x, float y) { return exp2f_(y * log2f_(x)); }

GG_HD float atanf_(float x) {
  float sign = 1.0f;
  if (x < 0.0f) { sign = -1.0f; x = -x; }
  float y0 = 0.0f;
  if (x > 2.414213562373095f) { y0 = 1.5707963267948966f; x = -1.0f / x; }
  else if (x > 0.4142135623730950f) { y0 = 0.7853981633974483f; x = (x - 1.0f) / (x + 1.0f); }
  float z = x * x;
  float y = ((((8.05374449538e-2f * z - 1.38776856032e-1f) * z + 1.99777106478e-1f) * z
             - 3.33329491539e-1f) * z * x + x) + y0;
  return sign * y;
}

GG_HD float atan2f_(float y, float x) {
  if (x == 0.0f) {
    if (y > 0.0f) return 1.5707963267948966f;
    if (y < 0.0f) return -1.5707963267948966f;
    return 0.0f;
  }
  float z = atanf_(y / x);
  if (x < 0.0f) z = (y < 0.0f) ? z - 3.14159265358979323846f : z + 3.14159265358979323846f;
  return z;
}

// ---------------------------------------------------------------------------
// RNG of the reference: random.glsl:1-49, host twin utilities/random.cpp:5-43.
GG_HD uint32_t wanghash(uint32_t x) {
  x = (x ^ 61u) ^ (x >> 16);
  x *= 9u;
  x = x ^ (x >> 4);
  x *= 0x27d4eb2du;
  x = x ^ (x >> 15);
  return x;
}
GG_HD uint32_t xorshift32(uint32_t& s) {
  s ^= (s << 13);
  s ^= (s >> 17);
  s ^= (s << 5);
  return s;
}
// Random(): float(u) / 4294967297.0 -- the FP32 literal is exactly 2^32, so
// the result lies in [0, 1] inclusive (random.glsl:42).
GG_HD float random01(uint32_t& s) { return (float)xorshift32(s) / 4294967296.0f; }
GG_HD uint32_t init_random_seed(uint32_t frameSeed, uint32_t invocation) {
  return wanghash(frameSeed + invocation);
}
// Host light-sample RNG: xorshift64* and XorshiftF (utilities/random.cpp:25-43).
GG_HD uint64_t xorshift64star(uint64_t s) {
  s ^= s >> 12;
  s ^= s << 25;
  s ^= s >> 27;
  return s * 2685821657736338717ull;
}
GG_HD float xorshiftf(uint64_t& s) {
  s = xorshift64star(s);
  return (float)(uint32_t)(s % 8388593ull) / 8388593.0f;
}

}  // namespace gg
