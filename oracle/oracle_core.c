/*
 * CPU oracle: RNG, FP32 math, traversal, brute force, FP64 classifier and the
 * microbenchmark ray generators.  TEST INFRASTRUCTURE ONLY (see oracle.h).
 */
#include <math.h>
#include <string.h>

#include "oracle.h"
#include "oracle_vec.h"

/* ---------------------------------------------------------------- RNG --- */
/* random.glsl:1-9 / utilities/random.cpp:5-13 */
uint32_t or_wanghash(uint32_t x) {
  x = (x ^ 61u) ^ (x >> 16);
  x = x * 9u;
  x = x ^ (x >> 4);
  x = x * 0x27d4eb2du;
  x = x ^ (x >> 15);
  return x;
}
/* random.glsl:11-19 (Marsaglia xorshift32, shifts 13/17/5) */
uint32_t or_xorshift32(uint32_t* s) {
  uint32_t x = *s;
  x ^= x << 13;
  x ^= x >> 17;
  x ^= x << 5;
  *s = x;
  return x;
}
/* random.glsl:38-43: float(u) / 4294967297.0, the FP32 literal being 2^32 */
float or_random01(uint32_t* s) {
  const float denom = 4294967297.0f;
  return (float)or_xorshift32(s) / denom;
}
/* utilities/random.cpp:25-31 */
uint64_t or_xorshift64star(uint64_t s) {
  s ^= s >> 12;
  s ^= s << 25;
  s ^= s >> 27;
  return s * UINT64_C(2685821657736338717);
}
/* utilities/random.cpp:39-43 */
float or_xorshiftf(uint64_t* s) {
  *s = or_xorshift64star(*s);
  return (float)(*s % 8388593u) / 8388593.0f;
}

/* --------------------------------------------------------------- math --- */
static uint32_t bits_of(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }
static float float_of(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }

/* Cody-Waite reduction by pi/2 (split 1.5703125 + 4.837512969970703125e-4 +
 * 7.54978995489188216e-8) and the Cephes sinf/cosf polynomials on [-pi/4, pi/4]. */
static void or_sincos(float x, float* s, float* c) {
  float k = floorf(x * 0.636619746685028076171875f + 0.5f);
  float r = x - k * 1.5703125f;
  r = r - k * 4.837512969970703125e-4f;
  r = r - k * 7.54978995489188216e-8f;
  float z = r * r;
  float ps = -1.9515295891e-4f * z + 8.3321608736e-3f;
  ps = ps * z - 1.6666654611e-1f;
  float sr = ps * z * r + r;
  float pc = 2.443315711809948e-5f * z - 1.388731625493765e-3f;
  pc = pc * z + 4.166664568298827e-2f;
  float cr = pc * z * z - 0.5f * z + 1.0f;
  switch (((int)k) & 3) {
    case 0: *s = sr; *c = cr; break;
    case 1: *s = cr; *c = -sr; break;
    case 2: *s = -sr; *c = -cr; break;
    default: *s = -cr; *c = sr; break;
  }
}
float or_sinf(float x) { float s, c; or_sincos(x, &s, &c); return s; }
float or_cosf(float x) { float s, c; or_sincos(x, &s, &c); return c; }

/* log2: mantissa in [sqrt(1/2), sqrt(2)) and the Cephes logf polynomial. */
float or_log2f(float x) {
  if (isnan(x)) return x;
  if (x < 0.0f) return NAN;
  if (x == 0.0f) return -INFINITY;
  if (isinf(x)) return x;
  int e = 0;
  if (x < 1.17549435e-38f) { x = x * 8388608.0f; e = -23; }
  uint32_t u = bits_of(x);
  e += (int)((u >> 23) & 0xff) - 126;
  float m = float_of((u & 0x007fffffu) | 0x3f000000u);
  if (m < 0.70710678118654752440f) { e -= 1; m = m + m - 1.0f; }
  else m = m - 1.0f;
  float z = m * m;
  float p = 7.0376836292e-2f * m - 1.1514610310e-1f;
  p = p * m + 1.1676998740e-1f;
  p = p * m - 1.2420140846e-1f;
  p = p * m + 1.4249322787e-1f;
  p = p * m - 1.6668057665e-1f;
  p = p * m + 2.0000714765e-1f;
  p = p * m - 2.4999993993e-1f;
  p = p * m + 3.3333331174e-1f;
  float y = p * m * z;
  y = y - 0.5f * z;
  float r = y * 1.44269504088896341f + m * 1.44269504088896341f;
  return r + (float)e;
}

/* 2^x: nearest integer plus the Cephes exp2f polynomial on [-1/2, 1/2]. */
float or_exp2f(float x) {
  if (isnan(x)) return x;
  if (x > 128.0f) return INFINITY;
  if (x < -150.0f) return 0.0f;
  float i = floorf(x + 0.5f);
  float f = x - i;
  float p = 1.535336188319500e-4f * f + 1.339887440266574e-3f;
  p = p * f + 9.618437357674640e-3f;
  p = p * f + 5.550332471162809e-2f;
  p = p * f + 2.402264791363012e-1f;
  p = p * f + 6.931472028550421e-1f;
  p = p * f + 1.0f;
  int n = (int)i;
  if (n > 127) return (p * float_of(0x7f000000u)) * 2.0f;
  if (n < -126) return (p * float_of((uint32_t)(n + 253) << 23)) * float_of(0x00800000u);
  return p * float_of((uint32_t)(n + 127) << 23);
}
/* GLSL pow(x, y) = exp2(y * log2(x)) (GLSL 4.50 sec. 8.2) */
float or_powf(float x, float y) { return or_exp2f(y * or_log2f(x)); }

/* Cephes atanf: reduction at tan(3pi/8) and tan(pi/8), odd polynomial. */
float or_atanf(float x) {
  float sign = 1.0f;
  if (x < 0.0f) { sign = -1.0f; x = -x; }
  float y0 = 0.0f;
  if (x > 2.414213562373095f) { y0 = 1.5707963267948966f; x = -1.0f / x; }
  else if (x > 0.4142135623730950f) { y0 = 0.7853981633974483f; x = (x - 1.0f) / (x + 1.0f); }
  float z = x * x;
  float p = 8.05374449538e-2f * z - 1.38776856032e-1f;
  p = p * z + 1.99777106478e-1f;
  p = p * z - 3.33329491539e-1f;
  float y = (p * z * x + x) + y0;
  return sign * y;
}
float or_atan2f(float y, float x) {
  if (x == 0.0f) return y > 0.0f ? 1.5707963267948966f : (y < 0.0f ? -1.5707963267948966f : 0.0f);
  float z = or_atanf(y / x);
  if (x < 0.0f) z = (y < 0.0f) ? z - 3.14159265358979323846f : z + 3.14159265358979323846f;
  return z;
}

/* ------------------------------------------------------------ traversal --- */
static vec3 pos_of(const or_scene* s, int v) {
  return v3(s->positions[3 * v], s->positions[3 * v + 1], s->positions[3 * v + 2]);
}

/* IntersectBox (intersectiontests.glsl:9-20) */
static int or_box(vec3 o, vec3 inv, vec3 mn, vec3 mx, float* first, float* last) {
  vec3 tbot = vmul(inv, vsub(mn, o));
  vec3 ttop = vmul(inv, vsub(mx, o));
  vec3 tmn = vmin3(ttop, tbot);
  vec3 tmx = vmax3(ttop, tbot);
  float a = fmaxf(tmn.x, tmn.y), b = fmaxf(tmn.x, tmn.z);
  *first = fmaxf(0.0f, fmaxf(a, b));
  a = fminf(tmx.x, tmx.y);
  b = fminf(tmx.x, tmx.z);
  *last = fminf(a, b);
  return *first <= *last;
}

/* rotate (intersectiontests.glsl:82-88) */
static vec3 or_rotate(vec3 x, const float* q) {
  vec3 t = vcross(v3(q[0], q[1], q[2]), x);
  vec3 r = v3(q[3] * t.x + q[1] * t.z - q[2] * t.y, q[3] * t.y + q[2] * t.x - q[0] * t.z,
              q[3] * t.z + q[0] * t.y - q[1] * t.x);
  return vadd(x, vscale(2.0f, r));
}

/* FetchIntersectOBoxNode (intersectiontests.glsl:114-128) for node record h (12 floats) */
static int or_obox(vec3 o, vec3 d, const float* h, float* first, float* last) {
  vec3 lo = or_rotate(vsub(o, v3(h[0], h[1], h[2])), h + 8);
  vec3 ld = or_rotate(d, h + 8);
  vec3 half = v3(h[4], h[5], h[6]);
  return or_box(lo, v3(1.0f / ld.x, 1.0f / ld.y, 1.0f / ld.z), vscale(-1.0f, half), half, first, last);
}

/* IntersectTriangle (intersectiontests.glsl:63-79) */
int or_triangle(vec3 o, vec3 d, vec3 p0, vec3 p1, vec3 p2, float* t, vec3* b, vec3* n) {
  vec3 e0 = vsub(p1, p0);
  vec3 e1 = vsub(p0, p2);
  *n = vcross(e1, e0);
  vec3 e2 = vscale(1.0f / vdot(*n, d), vsub(p0, o));
  vec3 i = vcross(d, e2);
  b->y = vdot(i, e1);
  b->z = vdot(i, e0);
  b->x = 1.0f - (b->z + b->y);
  *t = vdot(*n, e2);
  return (*t > 0.0001f) && b->x >= 0.0f && b->y >= 0.0f && b->z >= 0.0f;
}

/* Watertight ray/triangle test of Woop, Benthin and Wald (2013), "Watertight
 * Ray/Triangle Intersection", JCGT 2(1): vertices are sheared into a frame
 * whose z axis is the ray's dominant axis; the 2D edge functions U, V, W are
 * differences of products evaluated with Kahan's fma form, whose sign is
 * exact, so an edge shared by two triangles gets opposite signs in both and
 * no ray slips through.  Not a reference function (the reference test is
 * or_triangle); the HIP kernels' watertight mode restates this one.
 * Barycentrics: b.x, b.y, b.z weight p0, p1, p2; hits need t > 1e-4. */
static float comp_of(vec3 v, int k) { return k == 0 ? v.x : (k == 1 ? v.y : v.z); }

or_wtray or_wt_setup(vec3 d) {
  or_wtray r;
  const float ax = fabsf(d.x), ay = fabsf(d.y), az = fabsf(d.z);
  r.kz = (ax >= ay && ax >= az) ? 0 : (ay >= az ? 1 : 2);
  r.kx = r.kz == 2 ? 0 : r.kz + 1;
  r.ky = r.kx == 2 ? 0 : r.kx + 1;
  if (comp_of(d, r.kz) < 0.0f) { int t = r.kx; r.kx = r.ky; r.ky = t; }
  r.sx = comp_of(d, r.kx) / comp_of(d, r.kz);
  r.sy = comp_of(d, r.ky) / comp_of(d, r.kz);
  r.sz = 1.0f / comp_of(d, r.kz);
  return r;
}

/* a * b - c * d with an exact sign (Kahan): w = c*d, e = c*d - w exactly. */
static float diff_of_products(float a, float b, float c, float d) {
  const float w = c * d;
  const float e = fmaf(-c, d, w);
  const float f = fmaf(a, b, -w);
  return f + e;
}

int or_triangle_wt(vec3 o, const or_wtray* r, vec3 p0, vec3 p1, vec3 p2, float* t, vec3* b) {
  const vec3 A = vsub(p0, o), B = vsub(p1, o), C = vsub(p2, o);
  const float Akz = comp_of(A, r->kz), Bkz = comp_of(B, r->kz), Ckz = comp_of(C, r->kz);
  const float Ax = comp_of(A, r->kx) - r->sx * Akz, Ay = comp_of(A, r->ky) - r->sy * Akz;
  const float Bx = comp_of(B, r->kx) - r->sx * Bkz, By = comp_of(B, r->ky) - r->sy * Bkz;
  const float Cx = comp_of(C, r->kx) - r->sx * Ckz, Cy = comp_of(C, r->ky) - r->sy * Ckz;
  const float U = diff_of_products(Cx, By, Cy, Bx);
  const float V = diff_of_products(Ax, Cy, Ay, Cx);
  const float W = diff_of_products(Bx, Ay, By, Ax);
  if ((U < 0.0f || V < 0.0f || W < 0.0f) && (U > 0.0f || V > 0.0f || W > 0.0f)) return 0;
  const float det = U + V + W;
  if (det == 0.0f) return 0;
  const float Az = r->sz * Akz, Bz = r->sz * Bkz, Cz = r->sz * Ckz;
  const float T = U * Az + V * Bz + W * Cz;
  const float inv = 1.0f / det;
  *t = T * inv;
  b->x = U * inv;
  b->y = V * inv;
  b->z = W * inv;
  return *t > 0.0001f;
}

static int tri_test(int wt, vec3 o, vec3 d, const or_wtray* r, vec3 p0, vec3 p1, vec3 p2, float* t, vec3* b,
                    vec3* n) {
  if (!wt) return or_triangle(o, d, p0, p1, p2, t, b, n);
  *n = v3(0, 0, 0);
  return or_triangle_wt(o, r, p0, p1, p2, t, b);
}

/* TraceRay / TraceRayAnyHit, TRIANGLES_PER_LEAF = 2 (traceray.glsl:41-175),
 * with the reference triangle test (wt = 0) or the watertight one. */
/* TRACERAY_IMPORTANCE_BREAK parameters (traceray.glsl:48-112): one {x, y}
 * pair per node, the IMPORTANCE_THRESHOLD define, the _importanceThreshold
 * argument. */
typedef struct {
  const float* importance;
  float threshold, rayThreshold;
} or_imp;

/* nodeOut (closest hit, optional): {last hit node, its importance.x, its size^2} */
static int trace_one_ib(const or_scene* s, int any, int wt, vec3 o, vec3 d, float* rayLength, int32_t* slot,
                        vec3* bary, vec3* triNormal, int32_t tri[4], uint32_t* nbox, uint32_t* ntri,
                        const or_imp* ib, float* nodeOut) {
  const or_wtray wr = or_wt_setup(d);
  int lastNode = 0, hitNode = -1;
  float lastImp = 0.0f, lastSize = 0.0f, hitImp = 0.0f, hitSize = 0.0f;
  int node = 0, leaf = 0;
  int nextIsLeaf = 0;
  vec3 inv = v3(1.0f / d.x, 1.0f / d.y, 1.0f / d.z);
  *slot = -1;
  do {
    if (!nextIsLeaf) {
      const float* h = s->obox ? s->obox + (size_t)node * 12 : s->hierarchy + (size_t)node * 8;
      uint32_t childCode = bits_of(h[3]);
      int escape = (int)bits_of(h[7]);
      float newHit, exitDist;
      if (nbox) ++*nbox;
      int boxHit = s->obox ? or_obox(o, d, h, &newHit, &exitDist)
                           : or_box(o, inv, v3(h[0], h[1], h[2]), v3(h[4], h[5], h[6]), &newHit, &exitDist);
      if (boxHit && newHit <= *rayLength) {
        int broke = 0;
        lastNode = node;
        if (ib) {
          const float ix = ib->importance[2 * (size_t)node], iy = ib->importance[2 * (size_t)node + 1];
          float sizeSq;
          if (s->obox) {
            vec3 hh = v3(h[4], h[5], h[6]);
            sizeSq = vdot(hh, hh) * 4.0f;
          } else {
            vec3 e = vsub(v3(h[4], h[5], h[6]), v3(h[0], h[1], h[2]));
            sizeSq = vdot(e, e);
          }
          if (!any) {
            lastImp = ix;
            lastSize = sizeSq;
          }
          if ((ix < ib->threshold && newHit > 0.0f) || iy <= ib->rayThreshold * newHit) {
            broke = 1;
            if (any) {
              if (exitDist <= *rayLength) return 1;
            } else {
              hitImp = ix;
              hitSize = lastSize;
              hitNode = node;
              *rayLength = newHit + (exitDist - newHit) * ((float)(or_wanghash(bits_of(exitDist)) % 8388593u) / 8388593.0f);
              *slot = -1;
            }
            node = escape;
          }
        }
        if (!broke) {
          node = (int)(childCode & 0x7fffffffu);
          nextIsLeaf = (uint32_t)node != childCode;
          if (nextIsLeaf) {
            leaf = 2 * node;
            node = escape;
          }
        }
      } else {
        node = escape;
      }
    }
    if (nextIsLeaf) {
      const int32_t* t = s->triangles + (size_t)leaf * 4;
      if (t[0] == t[1]) {
        nextIsLeaf = 0;
      } else {
        float hit;
        vec3 b, n;
        if (ntri) ++*ntri;
        if (tri_test(wt, o, d, &wr, pos_of(s, t[0]), pos_of(s, t[1]), pos_of(s, t[2]), &hit, &b, &n) &&
            hit < *rayLength) {
          if (any) return 1;
          *rayLength = hit;
          *slot = leaf;
          hitNode = lastNode;
          hitImp = lastImp;
          hitSize = lastSize;
          *bary = b;
          if (triNormal) *triNormal = n;
          if (tri) memcpy(tri, t, 16);
        }
        ++leaf;
        nextIsLeaf = (leaf % 2) != 0;
      }
    }
  } while (node != 0 || nextIsLeaf);
  if (nodeOut) {
    memcpy(&nodeOut[0], &hitNode, 4);
    nodeOut[1] = hitImp;
    nodeOut[2] = hitSize;
  }
  return 0;
}

static int trace_one(const or_scene* s, int any, int wt, vec3 o, vec3 d, float* rayLength, int32_t* slot,
                     vec3* bary, vec3* triNormal, int32_t tri[4], uint32_t* nbox, uint32_t* ntri) {
  return trace_one_ib(s, any, wt, o, d, rayLength, slot, bary, triNormal, tri, nbox, ntri, NULL, NULL);
}

void or_trace_importance(const or_scene* s, int any, const float* rays, int32_t* hits, int32_t* nodes, size_t n,
                         const float* importance, float threshold, float rayThreshold) {
  or_imp ib = {importance, threshold, rayThreshold};
  for (size_t i = 0; i < n; ++i) {
    const float* r = rays + i * 8;
    vec3 o = v3(r[0], r[1], r[2]), d = v3(r[4], r[5], r[6]);
    float len = r[3];
    int32_t slot;
    vec3 b = v3(0, 0, 0);
    float info[3];
    int a = trace_one_ib(s, any, 0, o, d, &len, &slot, &b, NULL, NULL, NULL, NULL, &ib, info);
    int32_t* h = hits + i * 4;
    if (any) {
      h[0] = a; h[1] = h[2] = h[3] = 0;
    } else {
      h[0] = slot;
      h[1] = (int32_t)bits_of(len);
      h[2] = (int32_t)bits_of(slot >= 0 ? b.y : 0.0f);
      h[3] = (int32_t)bits_of(slot >= 0 ? b.z : 0.0f);
      if (nodes) {
        memcpy(&nodes[4 * i], &info[0], 4);
        nodes[4 * i + 1] = (int32_t)bits_of(info[1]);
        nodes[4 * i + 2] = (int32_t)bits_of(info[2]);
        nodes[4 * i + 3] = 0;
      }
    }
  }
}

/* TraceRayCountHits (hierarchy/counthits.glsl:1-71): the escape walk that adds
 * w to importance[2 * node] for every node whose box the ray enters at most
 * rayLength ahead, and to importance[2 * (num_nodes + slot)] for every valid
 * triangle slot it hits at most rayLength ahead. */
void or_count_hits(const or_scene* s, vec3 o, vec3 d, float rayLength, float w, float* importance) {
  int node = 0, leaf = 0, nextIsLeaf = 0;
  const vec3 inv = v3(1.0f / d.x, 1.0f / d.y, 1.0f / d.z);
  do {
    if (!nextIsLeaf) {
      const float* h = s->obox ? s->obox + (size_t)node * 12 : s->hierarchy + (size_t)node * 8;
      const uint32_t childCode = bits_of(h[3]);
      const int escape = (int)bits_of(h[7]);
      float newHit, exitDist;
      const int boxHit = s->obox ? or_obox(o, d, h, &newHit, &exitDist)
                                 : or_box(o, inv, v3(h[0], h[1], h[2]), v3(h[4], h[5], h[6]), &newHit, &exitDist);
      if (boxHit && newHit <= rayLength) {
        importance[2 * (size_t)node] += w;
        const int child = (int)(childCode & 0x7fffffffu);
        nextIsLeaf = (uint32_t)child != childCode;
        if (nextIsLeaf) {
          leaf = 2 * child;
          node = escape;
        } else {
          node = child;
        }
      } else {
        node = escape;
      }
    }
    if (nextIsLeaf) {
      const int32_t* t = s->triangles + (size_t)leaf * 4;
      if (t[0] == t[1]) {
        nextIsLeaf = 0;
      } else {
        float newHit;
        vec3 b, n;
        if (or_triangle(o, d, pos_of(s, t[0]), pos_of(s, t[1]), pos_of(s, t[2]), &newHit, &b, &n) &&
            newHit <= rayLength)
          importance[2 * ((size_t)s->num_nodes + (size_t)leaf)] += w;
        ++leaf;
        nextIsLeaf = (leaf % 2) != 0;
      }
    }
  } while (node != 0 || nextIsLeaf);
}

void or_count_hits_rays(const or_scene* s, const float* rays, const float* w, size_t n, float* importance) {
  for (size_t i = 0; i < n; ++i) {
    const float* r = rays + i * 8;
    or_count_hits(s, v3(r[0], r[1], r[2]), v3(r[4], r[5], r[6]), r[3], w[i], importance);
  }
}

/* Brute force over every slot (checker of the count walk): IntersectTriangle
 * accepts the slot at t <= the ray's length. */
void or_slot_hits(const or_scene* s, const float* rays, size_t n, uint8_t* out) {
  const size_t slots = 2 * (size_t)s->num_leaves;
  for (size_t i = 0; i < n; ++i) {
    const float* r = rays + i * 8;
    const vec3 o = v3(r[0], r[1], r[2]), d = v3(r[4], r[5], r[6]);
    for (size_t k = 0; k < slots; ++k) {
      const int32_t* t = s->triangles + k * 4;
      float th;
      vec3 b, nn;
      out[i * slots + k] = t[0] != t[1] && or_triangle(o, d, pos_of(s, t[0]), pos_of(s, t[1]), pos_of(s, t[2]), &th,
                                                        &b, &nn) && th <= r[3];
    }
  }
}

int or_trace_one(const or_scene* s, int any, vec3 o, vec3 d, float* rayLength, int32_t* slot, vec3* bary,
                 vec3* triNormal, int32_t tri[4], uint32_t* nbox, uint32_t* ntri) {
  return trace_one(s, any, 0, o, d, rayLength, slot, bary, triNormal, tri, nbox, ntri);
}

void or_trace(const or_scene* s, int any, const float* rays, int32_t* hits, uint32_t* nbox, uint32_t* ntri,
              size_t n) {
  or_trace_mode(s, any, 0, rays, hits, nbox, ntri, n);
}

void or_trace_mode(const or_scene* s, int any, int wt, const float* rays, int32_t* hits, uint32_t* nbox,
                   uint32_t* ntri, size_t n) {
  for (size_t i = 0; i < n; ++i) {
    const float* r = rays + i * 8;
    vec3 o = v3(r[0], r[1], r[2]), d = v3(r[4], r[5], r[6]);
    float len = r[3];
    int32_t slot;
    vec3 b = v3(0, 0, 0);
    uint32_t cb = 0, ct = 0;
    int a = trace_one(s, any, wt, o, d, &len, &slot, &b, NULL, NULL, nbox ? &cb : NULL, ntri ? &ct : NULL);
    int32_t* h = hits + i * 4;
    if (any) {
      h[0] = a; h[1] = h[2] = h[3] = 0;
    } else {
      h[0] = slot;
      h[1] = (int32_t)bits_of(len);
      h[2] = (int32_t)bits_of(slot >= 0 ? b.y : 0.0f);
      h[3] = (int32_t)bits_of(slot >= 0 ? b.z : 0.0f);
    }
    if (nbox) nbox[i] = cb;
    if (ntri) ntri[i] = ct;
  }
}

void or_bruteforce(const or_scene* s, int any, const float* rays, int32_t* hits, size_t n) {
  or_bruteforce_mode(s, any, 0, rays, hits, n);
}

void or_bruteforce_mode(const or_scene* s, int any, int wt, const float* rays, int32_t* hits, size_t n) {
  const size_t slots = (size_t)s->num_leaves * 2;
  for (size_t i = 0; i < n; ++i) {
    const float* r = rays + i * 8;
    vec3 o = v3(r[0], r[1], r[2]), d = v3(r[4], r[5], r[6]);
    const or_wtray wr = or_wt_setup(d);
    float best = r[3];
    int32_t bestSlot = -1;
    vec3 bb = v3(0, 0, 0);
    int found = 0;
    for (size_t k = 0; k < slots && !found; ++k) {
      const int32_t* t = s->triangles + k * 4;
      if (t[0] == t[1]) continue;
      float hit;
      vec3 b, nn;
      if (!tri_test(wt, o, d, &wr, pos_of(s, t[0]), pos_of(s, t[1]), pos_of(s, t[2]), &hit, &b, &nn)) continue;
      if (any) {
        if (hit < best) found = 1;
      } else if (hit < best) {  /* slots ascend, so the first of equal t is kept */
        best = hit;
        bestSlot = (int32_t)k;
        bb = b;
      }
    }
    int32_t* h = hits + i * 4;
    if (any) {
      h[0] = found; h[1] = h[2] = h[3] = 0;
    } else {
      h[0] = bestSlot;
      h[1] = (int32_t)bits_of(best);
      h[2] = (int32_t)bits_of(bestSlot >= 0 ? bb.y : 0.0f);
      h[3] = (int32_t)bits_of(bestSlot >= 0 ? bb.z : 0.0f);
    }
  }
}

/* FP64 ray/triangle geometry: t and the three barycentrics. */
static int fp64_hit(const or_scene* s, const float* ray, int32_t slot, double* t, double bary[3], double* cosang) {
  const int32_t* tri = s->triangles + (size_t)slot * 4;
  double p[3][3];
  for (int k = 0; k < 3; ++k)
    for (int a = 0; a < 3; ++a) p[k][a] = s->positions[3 * tri[k] + a];
  double o[3] = {ray[0], ray[1], ray[2]}, d[3] = {ray[4], ray[5], ray[6]};
  double e1[3], e2[3], pv[3], tv[3], qv[3];
  for (int a = 0; a < 3; ++a) { e1[a] = p[1][a] - p[0][a]; e2[a] = p[2][a] - p[0][a]; }
  pv[0] = d[1] * e2[2] - d[2] * e2[1];
  pv[1] = d[2] * e2[0] - d[0] * e2[2];
  pv[2] = d[0] * e2[1] - d[1] * e2[0];
  double det = e1[0] * pv[0] + e1[1] * pv[1] + e1[2] * pv[2];
  double nrm[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2], e1[0] * e2[1] - e1[1] * e2[0]};
  double nl = sqrt(nrm[0] * nrm[0] + nrm[1] * nrm[1] + nrm[2] * nrm[2]);
  double dl = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
  *cosang = (nl > 0 && dl > 0) ? fabs(nrm[0] * d[0] + nrm[1] * d[1] + nrm[2] * d[2]) / (nl * dl) : 0.0;
  if (det == 0.0) return 0;
  for (int a = 0; a < 3; ++a) tv[a] = o[a] - p[0][a];
  double u = (tv[0] * pv[0] + tv[1] * pv[1] + tv[2] * pv[2]) / det;
  qv[0] = tv[1] * e1[2] - tv[2] * e1[1];
  qv[1] = tv[2] * e1[0] - tv[0] * e1[2];
  qv[2] = tv[0] * e1[1] - tv[1] * e1[0];
  double v = (d[0] * qv[0] + d[1] * qv[1] + d[2] * qv[2]) / det;
  *t = (e2[0] * qv[0] + e2[1] * qv[1] + e2[2] * qv[2]) / det;
  bary[0] = 1.0 - u - v;
  bary[1] = u;
  bary[2] = v;
  return 1;
}

int or_classify(const or_scene* s, const float* ray, int32_t slot_a, int32_t slot_b, double rel_eps) {
  int flags = 0;
  double ta = 0, tb = 0;
  int ha = 0, hb = 0;
  const double tmax = ray[3];
  int32_t sl[2] = {slot_a, slot_b};
  for (int k = 0; k < 2; ++k) {
    if (sl[k] < 0) continue;
    double t, b[3], ca;
    if (!fp64_hit(s, ray, sl[k], &t, b, &ca)) { flags |= 16; continue; }
    double mb = b[0] < b[1] ? (b[0] < b[2] ? b[0] : b[2]) : (b[1] < b[2] ? b[1] : b[2]);
    if (fabs(mb) <= rel_eps) flags |= (k == 0 ? 1 : 2);
    if (fabs(t - 1e-4) <= rel_eps * 1e-4 + 1e-12 || fabs(t - tmax) <= rel_eps * fabs(tmax)) flags |= 8;
    if (ca < 1e-6) flags |= 16;
    if (k == 0) { ta = t; ha = 1; } else { tb = t; hb = 1; }
  }
  if (ha && hb) {
    double m = fabs(ta) > fabs(tb) ? fabs(ta) : fabs(tb);
    if (fabs(ta - tb) <= rel_eps * m) flags |= 4;
  }
  return flags;
}

/* ------------------------------------------------------------ raygen --- */
/* Primary ray as pathtracer.comp:39-46 with InitRandomSeed(seed, i). */
void or_raygen_primary(const or_camera* c, int width, int height, uint32_t seed, float* rays, size_t first,
                       size_t n) {
  vec3 U = v3(c->u[0], c->u[1], c->u[2]), V = v3(c->v[0], c->v[1], c->v[2]), W = v3(c->w[0], c->w[1], c->w[2]);
  for (size_t k = 0; k < n; ++k) {
    uint32_t i = (uint32_t)(first + k);
    uint32_t pix = i % (uint32_t)(width * height);
    int px = (int)(pix % (uint32_t)width), py = (int)(pix / (uint32_t)width);
    uint32_t rs = or_wanghash(seed + i);
    float j0 = or_random01(&rs), j1 = or_random01(&rs);
    vec3 d = or_primary_dir(j0, j1, px, py, width, height, U, V, W);
    float* r = rays + k * 8;
    r[0] = c->position[0]; r[1] = c->position[1]; r[2] = c->position[2]; r[3] = 3.402823466e+38f;
    r[4] = d.x; r[5] = d.y; r[6] = d.z; r[7] = 0.0f;
  }
}

static void hit_list(const int32_t* hits, size_t n, uint32_t** list, size_t* count) {
  static uint32_t* buf = NULL;
  static size_t cap = 0;
  if (cap < n) { free(buf); buf = (uint32_t*)malloc(n * 4); cap = n; }
  size_t c = 0;
  for (size_t i = 0; i < n; ++i)
    if (hits[i * 4] >= 0) buf[c++] = (uint32_t)i;
  *list = buf;
  *count = c;
}

void or_raygen_diffuse(const or_scene* s, const float* prim, const int32_t* prim_hits, size_t n_prim, uint32_t seed,
                       float* rays, size_t first, size_t n) {
  uint32_t* list;
  size_t nh;
  hit_list(prim_hits, n_prim, &list, &nh);
  for (size_t k = 0; k < n; ++k) {
    uint32_t i = (uint32_t)(first + k);
    float* r = rays + k * 8;
    if (nh == 0) { memset(r, 0, 32); r[5] = 1.0f; continue; }
    uint32_t j = list[i % nh];
    const float* pr = prim + (size_t)j * 8;
    const int32_t* h = prim_hits + (size_t)j * 4;
    vec3 o = v3(pr[0], pr[1], pr[2]), din = v3(pr[4], pr[5], pr[6]);
    vec3 p = vadd(o, vscale(float_of((uint32_t)h[1]), din));
    const int32_t* t = s->triangles + (size_t)h[0] * 4;
    vec3 p0 = pos_of(s, t[0]), p1 = pos_of(s, t[1]), p2 = pos_of(s, t[2]);
    vec3 N = vnormalize(vcross(vsub(p0, p2), vsub(p1, p0)));
    if (vdot(N, din) > 0.0f) N = vneg(N);
    vec3 U, V;
    or_create_onb(N, &U, &V);
    uint32_t rs = or_wanghash(seed + i);
    float r0 = or_random01(&rs), r1 = or_random01(&rs);
    vec3 d = or_hemisphere_cosine(r0, r1, U, V, N);
    vec3 oo = vadd(p, vscalev(d, 0.0001f));
    r[0] = oo.x; r[1] = oo.y; r[2] = oo.z; r[3] = 3.402823466e+38f;
    r[4] = d.x; r[5] = d.y; r[6] = d.z; r[7] = 0.0f;
  }
}

void or_raygen_shadow(const float* prim, const int32_t* prim_hits, size_t n_prim, const float* lights,
                      uint32_t n_lights, float* rays, size_t first, size_t n) {
  uint32_t* list;
  size_t nh;
  hit_list(prim_hits, n_prim, &list, &nh);
  for (size_t k = 0; k < n; ++k) {
    uint32_t i = (uint32_t)(first + k);
    float* r = rays + k * 8;
    if (nh == 0) { memset(r, 0, 32); r[5] = 1.0f; continue; }
    uint32_t j = list[i % nh];
    const float* pr = prim + (size_t)j * 8;
    vec3 p = vadd(v3(pr[0], pr[1], pr[2]), vscale(float_of((uint32_t)prim_hits[(size_t)j * 4 + 1]), v3(pr[4], pr[5], pr[6])));
    uint32_t l = i % n_lights;
    vec3 L = v3(lights[3 * l], lights[3 * l + 1], lights[3 * l + 2]);
    /* lightsource.glsl:18-21, 39-40 */
    vec3 dir = vsub(L, p);
    float distSq = vdot(dir, dir) + 1e-5f;
    float dist = sqrtf(distSq);
    dir = vdiv(dir, dist);
    vec3 o = vadd(vscale(0.0001f, dir), p);
    r[0] = o.x; r[1] = o.y; r[2] = o.z; r[3] = dist - 0.0001f * 2.0f;
    r[4] = dir.x; r[5] = dir.y; r[6] = dir.z; r[7] = 0.0f;
  }
}
