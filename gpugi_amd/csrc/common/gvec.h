// Small FP32 vector helpers with a fixed evaluation order, shared by host C++
// and HIP device code.  dot() sums left to right, cross() is the textbook
// component formula, normalize() divides every component by the length; with
// -ffp-contract=off these give identical bits on the host and on gfx950.
#pragma once
#include "gmath.h"

namespace gg {

struct V3 { float x, y, z; };

GG_HD V3 v3(float x, float y, float z) { V3 r; r.x = x; r.y = y; r.z = z; return r; }
GG_HD V3 v3s(float s) { return v3(s, s, s); }
GG_HD V3 operator+(V3 a, V3 b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
GG_HD V3 operator-(V3 a, V3 b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
GG_HD V3 operator-(V3 a) { return v3(-a.x, -a.y, -a.z); }
GG_HD V3 operator*(V3 a, V3 b) { return v3(a.x * b.x, a.y * b.y, a.z * b.z); }
GG_HD V3 operator*(V3 a, float s) { return v3(a.x * s, a.y * s, a.z * s); }
GG_HD V3 operator*(float s, V3 a) { return v3(s * a.x, s * a.y, s * a.z); }
GG_HD V3 operator/(V3 a, float s) { return v3(a.x / s, a.y / s, a.z / s); }
GG_HD V3 operator/(V3 a, V3 b) { return v3(a.x / b.x, a.y / b.y, a.z / b.z); }
GG_HD float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
GG_HD V3 cross(V3 a, V3 b) {
  return v3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
GG_HD float length(V3 a) { return sqrtf(dot(a, a)); }
GG_HD V3 normalize(V3 a) { float l = length(a); return a / l; }
GG_HD V3 vmin(V3 a, V3 b) { return v3(fminf(a.x, b.x), fminf(a.y, b.y), fminf(a.z, b.z)); }
GG_HD V3 vmax(V3 a, V3 b) { return v3(fmaxf(a.x, b.x), fmaxf(a.y, b.y), fmaxf(a.z, b.z)); }
GG_HD V3 saturate(V3 a) { return v3(saturate(a.x), saturate(a.y), saturate(a.z)); }
GG_HD V3 vabs(V3 a) { return v3(fabsf(a.x), fabsf(a.y), fabsf(a.z)); }

}  // namespace gg
