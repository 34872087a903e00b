/*
 * CPU oracle: the importance and material passes of the HierarchyImportance
 * renderer (renderer/hierarchyimportance.cpp:84-178; shader/hierarchy/
 * acquisition2.comp, counthits.glsl, hierarchypropagation_init.comp,
 * hierarchypropagation_nodes.comp, sampleleafmaterials.comp, pullmaterials.comp).
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * The importance buffer holds {x, y} per node, then {x, y} per triangle slot.
 * The reference propagates and pulls with repeated passes over all nodes until
 * nothing changes; every node's value depends only on its children's final
 * values, summed in sibling order, so one sweep in reverse preorder (children
 * after their parent) gives the converged result.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"
#include "oracle_vec.h"

static uint32_t fbits(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }

static const float* node_rec(const or_scene* s, int n) {
  return s->obox ? s->obox + (size_t)n * 12 : s->hierarchy + (size_t)n * 8;
}

/* acquisition2.comp:23-112 (NEXT_EVENT_IMPORTANCE, MAX_PATHLENGTH 8, no
 * roulette) for iteration `iteration` (NumImportanceIterations, 1-based) over
 * the W/8*8 x H/8*8 dispatch grid. */
void or_hier_acquire(const or_scene* s, const float* lights, uint32_t numLights, const or_camera* c, int W, int H,
                     int iteration, float* importance) {
  const int gw = (W / 8) * 8, gh = (H / 8) * 8;
  const vec3 U = v3(c->u[0], c->u[1], c->u[2]), V = v3(c->v[0], c->v[1], c->v[2]), Wd = v3(c->w[0], c->w[1], c->w[2]);
  const uint32_t frameSeed = or_wanghash((uint32_t)iteration);
  for (int y = 0; y < gh; ++y) {
    for (int x = 0; x < gw; ++x) {
      uint32_t seed = or_wanghash(frameSeed + (uint32_t)(x + y * W));
      const float j0 = or_random01(&seed), j1 = or_random01(&seed);
      /* (Random2 + gridPosition) / BackbufferSize * 2 - 1: no half-pixel shift here */
      const float sx = ((j0 + (float)x) / (float)W) * 2.0f - 1.0f;
      const float sy = ((j1 + (float)y) / (float)H) * 2.0f - 1.0f;
      vec3 o = v3(c->position[0], c->position[1], c->position[2]);
      vec3 d = vnormalize(vadd(vadd(vscale(sx, U), vscale(sy, V)), Wd));
      float pathImportance = 1.0f;
      for (int i = 0; i < 8; ++i) {
        float t = 3.402823466e+38f;
        int32_t slot, tri[4];
        vec3 b;
        or_trace_one(s, 0, o, d, &t, &slot, &b, NULL, tri, NULL, NULL);
        or_count_hits(s, o, d, t * 1.000001f, pathImportance, importance);
        if (t == 3.402823466e+38f) break;
        const vec3 n = or_shading_normal(s, tri, b);
        float tu, tv;
        or_hit_texcoord(s, tri, b, &tu, &tv);
        const or_material m = or_mat_at(s, tri[3], tu, tv);
        o = vadd(o, vscale(t, d));
        pathImportance *= 0.5f;
        const int idx = (int)(or_xorshift32(&seed) % numLights);
        const float* ls = lights + (size_t)idx * 8;
        vec3 ld = vsub(v3(ls[0], ls[1], ls[2]), o);
        const float dist = sqrtf(vdot(ld, ld));
        ld = vdiv(ld, dist);
        const vec3 lsd = or_unpack_normal(ls[3], ls[7]);
        const float f = fsat(vdot(lsd, vneg(ld)));
        if (f > 0.0f) {
          float lt = dist;
          int32_t ls2, lt4[4];
          vec3 lb;
          or_trace_one(s, 0, o, ld, &lt, &ls2, &lb, NULL, lt4, NULL, NULL);
          if (lt < dist) or_count_hits(s, o, ld, lt * 1.000001f, pathImportance, importance);
        }
        vec3 tp = v3(1, 1, 1);
        float pdf;
        int dif;
        d = or_sample_bsdf(d, &m, &seed, n, &tp, 0, &pdf, &dif);
        o = vadd(o, vscalev(d, 0.0001f));
      }
    }
  }
}

/* Parent of every node (-1 for the root): an inner node's children are its
 * first child and the siblings its escape chain reaches inside its own
 * preorder range (the reference's parent-pointer buffer, scene.cpp:125). */
void or_hier_parents(const or_scene* s, int32_t* parent) {
  const int n = (int)s->num_nodes;
  for (int i = 0; i < n; ++i) parent[i] = -1;
  for (int p = 0; p < n; ++p) {
    const float* h = node_rec(s, p);
    const uint32_t cc = fbits(h[3]);
    if (cc & 0x80000000u) continue;
    const int end = (int)fbits(h[7]) == 0 ? n : (int)fbits(h[7]);
    for (int k = (int)cc; k > p && k < end; k = (int)fbits(node_rec(s, k)[7])) {
      parent[k] = p;
      if ((int)fbits(node_rec(s, k)[7]) == 0) break;
    }
  }
}

/* avgArea of hierarchypropagation_*.comp: a third of a box's surface */
static float avg_area(const or_scene* s, int n) {
  const float* h = node_rec(s, n);
  if (s->obox) /* half side lengths: x*y*4 per face, two faces each */
    return (h[4] * h[5] + h[4] * h[6] + h[5] * h[6]) * 8.0f / 6.0f;
  const float sx = h[4] - h[0], sy = h[5] - h[1], sz = h[6] - h[2];
  return (sx * sy + sx * sz + sy * sz) * 2.0f / 6.0f;
}

/* HierarchyImportance::UpdateHierarchyNodeImportance (hierarchyimportance.cpp:120-152):
 * hierarchypropagation_init.comp:22-69 then hierarchypropagation_nodes.comp:22-81
 * until every node has a subtree importance.  importance: in/out; subtree: out. */
void or_hier_propagate(const or_scene* s, int iterations, float* importance, float* subtree) {
  const int n = (int)s->num_nodes;
  const float N = (float)iterations;
  for (int i = 0; i < n; ++i) {
    subtree[i] = 0.0f;
    importance[2 * (size_t)i] /= N;
    importance[2 * (size_t)i + 1] = 0.0f;
    const uint32_t cc = fbits(node_rec(s, i)[3]);
    if (!(cc & 0x80000000u)) continue;
    const size_t first = 2 * (size_t)(cc & 0x7fffffffu);
    float children = 0.0f;
    for (int k = 0; k < 2; ++k) children += importance[2 * ((size_t)n + first + (size_t)k)];
    children /= N;
    subtree[i] = fmaxf(children, 1.19e-07f); /* NotificationBias */
    const float approximation = importance[2 * (size_t)i] / children;
    importance[2 * (size_t)i + 1] = approximation * approximation * avg_area(s, i);
  }
  int32_t* parent = (int32_t*)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
  or_hier_parents(s, parent);
  for (int i = n - 1; i >= 0; --i) {
    if (subtree[i] != 0.0f) continue;
    int child = (int)(fbits(node_rec(s, i)[3]) & 0x7fffffffu);
    float children = 0.0f;
    for (;;) {
      children += subtree[child];
      child = (int)fbits(node_rec(s, child)[7]);
      if (child == 0 || parent[child] != i) break;
    }
    subtree[i] = children;
    const float approximation = importance[2 * (size_t)i] / children;
    importance[2 * (size_t)i + 1] = approximation * approximation * avg_area(s, i);
  }
  free(parent);
}

/* packHalf2x16's float -> half conversion, rounded to nearest even (IEEE 754
 * binary16: overflow to infinity, subnormals, NaN kept quiet). */
uint16_t or_half_bits(float f) {
  const uint32_t u = fbits(f);
  const uint32_t sign = (u >> 16) & 0x8000u;
  const uint32_t ex = (u >> 23) & 0xffu;
  uint32_t man = u & 0x7fffffu;
  if (ex == 0xffu) return (uint16_t)(sign | 0x7c00u | (man ? 0x200u | (man >> 13) : 0u));
  const int e = (int)ex - 127 + 15;
  if (e >= 31) return (uint16_t)(sign | 0x7c00u);
  if (e <= 0) {
    if (e < -10) return (uint16_t)sign;
    man |= 0x800000u;
    const int shift = 14 - e;
    uint32_t h = man >> shift;
    const uint32_t rem = man & ((1u << shift) - 1u), half = 1u << (shift - 1);
    if (rem > half || (rem == half && (h & 1u))) ++h;
    return (uint16_t)(sign | h);
  }
  uint32_t h = ((uint32_t)e << 10) | (man >> 13);
  const uint32_t rem = man & 0x1fffu;
  if (rem > 0x1000u || (rem == 0x1000u && (h & 1u))) ++h; /* may carry into the exponent: still right */
  return (uint16_t)(sign | h);
}

float or_half_value(uint16_t h) {
  const uint32_t sign = (uint32_t)(h & 0x8000u) << 16;
  const uint32_t ex = (h >> 10) & 0x1fu, man = h & 0x3ffu;
  uint32_t u;
  if (ex == 0) {
    if (man == 0) {
      u = sign;
    } else { /* subnormal: normalize */
      int e = -1;
      uint32_t m = man;
      do { ++e; m <<= 1; } while (!(m & 0x400u));
      u = sign | ((uint32_t)(127 - 15 - e) << 23) | ((m & 0x3ffu) << 13);
    }
  } else if (ex == 31) {
    u = sign | 0x7f800000u | (man << 13);
  } else {
    u = sign | ((ex - 15 + 127) << 23) | (man << 13);
  }
  float f;
  memcpy(&f, &u, 4);
  return f;
}

static uint32_t pack2(float a, float b) { return (uint32_t)or_half_bits(a) | ((uint32_t)or_half_bits(b) << 16); }
static float lo_of(uint32_t w) { return or_half_value((uint16_t)(w & 0xffffu)); }
static float hi_of(uint32_t w) { return or_half_value((uint16_t)(w >> 16)); }

/* HierarchyMaterial (hierarchymaterialbuffer.glsl:1-15): 12 words per node */
static void store_material(uint32_t* w, const float refl[4], const float op[3], const float dif[3], const float em[3],
                           const float f0[3], float rIdx, const float f1[3], float area) {
  w[0] = pack2(refl[0], refl[1]);
  w[1] = pack2(refl[2], refl[3]);
  w[2] = pack2(op[0], op[1]);
  w[3] = pack2(op[2], area);
  w[4] = pack2(dif[0], dif[1]);
  w[5] = pack2(dif[2], 0.0f);
  w[6] = pack2(em[0], em[1]);
  w[7] = pack2(em[2], 0.0f);
  w[8] = pack2(f0[0], f0[1]);
  w[9] = pack2(f0[2], rIdx);
  w[10] = pack2(f1[0], f1[1]);
  w[11] = pack2(f1[2], 0.0f);
}

/* HierarchyImportance::ComputeHierarchyMaterials (hierarchyimportance.cpp:154-178):
 * sampleleafmaterials.comp:7-98 then pullmaterials.comp:9-94 over the tree. */
void or_hier_materials(const or_scene* s, uint32_t* out) {
  const int n = (int)s->num_nodes;
  for (int i = 0; i < n; ++i) {
    const uint32_t cc = fbits(node_rec(s, i)[3]);
    if (!(cc & 0x80000000u)) continue;
    const int base = 2 * (int)(cc & 0x7fffffffu);
    float refl[4] = {0, 0, 0, 0}, op[3] = {0, 0, 0}, dif[3] = {0, 0, 0}, em[3] = {0, 0, 0}, f0[3] = {0, 0, 0},
          f1[3] = {0, 0, 0}, rIdx = 0.0f, area = 0.0f;
    for (int k = 0; k < 2; ++k) {
      const int32_t* t = s->triangles + (size_t)(base + k) * 4;
      if (t[0] == t[1]) break;
      const float* p0 = s->positions + 3 * (size_t)t[0];
      const float* p1 = s->positions + 3 * (size_t)t[1];
      const float* p2 = s->positions + 3 * (size_t)t[2];
      const vec3 a = v3(p0[0], p0[1], p0[2]);
      const vec3 cr = vcross(vsub(v3(p1[0], p1[1], p1[2]), a), vsub(v3(p2[0], p2[1], p2[2]), a));
      const float triangleArea = sqrtf(vdot(cr, cr)) / 2.0f;
      area += triangleArea;
      const float third = 1.0f / 3.0f;
      const float* va = s->vertex_info + 4 * (size_t)t[0];
      const float* vb = s->vertex_info + 4 * (size_t)t[1];
      const float* vc = s->vertex_info + 4 * (size_t)t[2];
      const float u = (va[2] * third + vb[2] * third) + vc[2] * third;
      const float v = (va[3] * third + vb[3] * third) + vc[3] * third;
      const or_material m = or_mat_at(s, t[3], u, v);
      for (int q = 0; q < 4; ++q) refl[q] += m.refl[q] * triangleArea;
      const float mo[3] = {m.opacity.x, m.opacity.y, m.opacity.z}, md[3] = {m.diffuse.x, m.diffuse.y, m.diffuse.z};
      const float me[3] = {m.emissivity.x, m.emissivity.y, m.emissivity.z};
      const float m0[3] = {m.fresnel0.x, m.fresnel0.y, m.fresnel0.z}, m1[3] = {m.fresnel1.x, m.fresnel1.y, m.fresnel1.z};
      for (int q = 0; q < 3; ++q) {
        op[q] += mo[q] * triangleArea;
        dif[q] += md[q] * triangleArea;
        em[q] += me[q] * triangleArea;
        f0[q] += m0[q] * triangleArea;
        f1[q] += m1[q] * triangleArea;
      }
      rIdx += m.rIndex * triangleArea;
    }
    for (int q = 0; q < 4; ++q) refl[q] /= area;
    for (int q = 0; q < 3; ++q) {
      op[q] /= area;
      dif[q] /= area;
      em[q] /= area;
      f0[q] /= area;
      f1[q] /= area;
    }
    rIdx /= area;
    store_material(out + 12 * (size_t)i, refl, op, dif, em, f0, rIdx, f1, area);
  }
  int32_t* parent = (int32_t*)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
  or_hier_parents(s, parent);
  for (int i = n - 1; i >= 0; --i) {
    const uint32_t cc = fbits(node_rec(s, i)[3]);
    if (cc & 0x80000000u) continue;
    float refl[4] = {0, 0, 0, 0}, op[3] = {0, 0, 0}, dif[3] = {0, 0, 0}, em[3] = {0, 0, 0}, f0[3] = {0, 0, 0},
          f1[3] = {0, 0, 0}, rIdx = 0.0f, area = 0.0f;
    int child = (int)cc;
    for (;;) {
      const uint32_t* w = out + 12 * (size_t)child;
      const float ca = hi_of(w[3]);
      area += ca;
      refl[0] += lo_of(w[0]) * ca;
      refl[1] += hi_of(w[0]) * ca;
      refl[2] += lo_of(w[1]) * ca;
      refl[3] += hi_of(w[1]) * ca;
      op[0] += lo_of(w[2]) * ca;
      op[1] += hi_of(w[2]) * ca;
      op[2] += lo_of(w[3]) * ca;
      dif[0] += lo_of(w[4]) * ca;
      dif[1] += hi_of(w[4]) * ca;
      dif[2] += lo_of(w[5]) * ca;
      em[0] += lo_of(w[6]) * ca;
      em[1] += hi_of(w[6]) * ca;
      em[2] += lo_of(w[7]) * ca;
      f0[0] += lo_of(w[8]) * ca;
      f0[1] += hi_of(w[8]) * ca;
      f0[2] += lo_of(w[9]) * ca;
      rIdx += hi_of(w[9]) * ca;
      f1[0] += lo_of(w[10]) * ca;
      f1[1] += hi_of(w[10]) * ca;
      f1[2] += lo_of(w[11]) * ca;
      child = (int)fbits(node_rec(s, child)[7]);
      if (child == 0 || parent[child] != i) break;
    }
    for (int q = 0; q < 4; ++q) refl[q] /= area;
    for (int q = 0; q < 3; ++q) {
      op[q] /= area;
      dif[q] /= area;
      em[q] /= area;
      f0[q] /= area;
      f1[q] /= area;
    }
    rIdx /= area;
    store_material(out + 12 * (size_t)i, refl, op, dif, em, f0, rIdx, f1, area);
  }
  free(parent);
}
