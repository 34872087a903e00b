// Escape-pointer walk over the reference BVH2 layout (device code).
//
// This is the bit-exact traversal: same node order, same FP32 operations in
// the same order as traceray.glsl:41-175 with intersectiontests.glsl:9-20,
// 63-79, 100-111, so every ray gets the reference's hit slot, t and
// barycentrics including exact ties (first visited wins) and box grazes.
#pragma once
#include <hip/hip_runtime.h>

#include "../common/gshade.h"

namespace gpugi {

struct RefScene {
  const float4* __restrict__ nodes;     // 2 texels per node (HierachyBuffer, TBO 4)
  const int4* __restrict__ slots;       // TriangleBuffer, TBO 1
  const float* __restrict__ positions;  // VertexPositionBuffer RGB32F, TBO 2
  const float4* __restrict__ vinfo;     // VertexInfoBuffer, TBO 3
  int obox = 0;                         // nodes are TreeNode<OBox>: 3 texels each (OBOX_BVH)
};

// TRACERAY_IMPORTANCE_BREAK (traceray.glsl:48-112): per-node importance {x, y}
// (HierarchyImportanceBuffer, RG32F), the IMPORTANCE_THRESHOLD define and the
// _importanceThreshold argument.  A hit node with (x < threshold and newHit > 0)
// or y <= rayThreshold * newHit stops the descent: its subtree is skipped, and
// the closest-hit walk records a hit at a pseudo-random distance inside the
// node's box interval, the any-hit walk returns true when that interval ends
// before the ray does.  importance == nullptr: the plain walk.
struct ImportanceBreak {
  const float2* __restrict__ importance = nullptr;
  float threshold = 0.0f;
  float rayThreshold = 0.0f;
};

struct RefHit {
  float t;           // rayLength
  int slot;          // _hitIndex.y
  int node;          // _hitIndex.x
  gg::V3 bary;       // barycentricCoord
  gg::V3 triNormal;  // TRINORMAL_OUTPUT (not normalized)
  int4 tri;          // outTriangle
  float nodeImportance = 0.0f;  // _nodeImportance (importance break)
  float nodeSizeSq = 0.0f;      // _nodeSizeSq
};

__device__ __forceinline__ gg::V3 load_pos(const float* __restrict__ p, int v) {
  return gg::v3(p[3 * v + 0], p[3 * v + 1], p[3 * v + 2]);
}

// IntersectBox (intersectiontests.glsl:9-20).
__device__ __forceinline__ bool intersect_box(gg::V3 o, gg::V3 inv, gg::V3 bmin, gg::V3 bmax,
                                              float& firstHit, float& lastHit) {
  gg::V3 tbot = inv * (bmin - o);
  gg::V3 ttop = inv * (bmax - o);
  gg::V3 tmin = gg::vmin(ttop, tbot);
  gg::V3 tmax = gg::vmax(ttop, tbot);
  float tx = fmaxf(tmin.x, tmin.y), ty = fmaxf(tmin.x, tmin.z);
  firstHit = fmaxf(0.0f, fmaxf(tx, ty));
  tx = fminf(tmax.x, tmax.y);
  ty = fminf(tmax.x, tmax.z);
  lastHit = fminf(tx, ty);
  return firstHit <= lastHit;
}

// rotate (intersectiontests.glsl:82-88): x rotated by the quaternion q.
__device__ __forceinline__ gg::V3 rotate_q(gg::V3 x, float4 q) {
  const gg::V3 t = gg::cross(gg::v3(q.x, q.y, q.z), x);
  return x + 2.0f * gg::v3(q.w * t.x + q.y * t.z - q.z * t.y, q.w * t.y + q.z * t.x - q.x * t.z,
                           q.w * t.z + q.x * t.y - q.y * t.x);
}

// FetchIntersectOBoxNode (intersectiontests.glsl:114-128) after the fetch: the
// ray in the box's frame against the centred box.
__device__ __forceinline__ bool intersect_obox(gg::V3 o, gg::V3 d, float4 a, float4 b, float4 q, float& firstHit,
                                               float& lastHit) {
  const gg::V3 lo = rotate_q(o - gg::v3(a.x, a.y, a.z), q);
  const gg::V3 ld = rotate_q(d, q);
  const gg::V3 half = gg::v3(b.x, b.y, b.z);
  return intersect_box(lo, gg::v3(1.0f / ld.x, 1.0f / ld.y, 1.0f / ld.z), -half, half, firstHit, lastHit);
}

// IntersectTriangle (intersectiontests.glsl:63-79), the OptiX-style branchless
// Moeller-Trumbore variant: not watertight, t > 1e-4, all barycentrics >= 0.
__device__ __forceinline__ bool intersect_triangle(gg::V3 o, gg::V3 d, gg::V3 p0, gg::V3 p1, gg::V3 p2,
                                                   float& hit, gg::V3& bary, gg::V3& n) {
  gg::V3 e0 = p1 - p0;
  gg::V3 e1 = p0 - p2;
  n = gg::cross(e1, e0);
  gg::V3 e2 = (1.0f / gg::dot(n, d)) * (p0 - o);
  gg::V3 i = gg::cross(d, e2);
  bary.y = gg::dot(i, e1);
  bary.z = gg::dot(i, e0);
  bary.x = 1.0f - (bary.z + bary.y);
  hit = gg::dot(n, e2);
  return (hit > gg::kIntersectEpsilon) && bary.x >= 0.0f && bary.y >= 0.0f && bary.z >= 0.0f;
}

// Watertight ray/triangle test (Woop, Benthin and Wald 2013, "Watertight
// Ray/Triangle Intersection", JCGT 2(1)); no reference counterpart.  The
// vertices are sheared into a frame whose z axis is the ray's dominant axis
// and the 2D edge functions are differences of products in Kahan's fma form,
// whose sign is exact: an edge shared by two triangles gets opposite signs in
// both, so no ray passes between them.  b.x, b.y, b.z weight p0, p1, p2.
struct WtRay {
  int kx, ky, kz;
  float sx, sy, sz;
};

__device__ __forceinline__ float comp(gg::V3 v, int k) { return k == 0 ? v.x : (k == 1 ? v.y : v.z); }

__device__ __forceinline__ WtRay wt_setup(gg::V3 d) {
  WtRay r;
  const float ax = fabsf(d.x), ay = fabsf(d.y), az = fabsf(d.z);
  r.kz = (ax >= ay && ax >= az) ? 0 : (ay >= az ? 1 : 2);
  r.kx = r.kz == 2 ? 0 : r.kz + 1;
  r.ky = r.kx == 2 ? 0 : r.kx + 1;
  if (comp(d, r.kz) < 0.0f) {
    const int t = r.kx;
    r.kx = r.ky;
    r.ky = t;
  }
  r.sx = comp(d, r.kx) / comp(d, r.kz);
  r.sy = comp(d, r.ky) / comp(d, r.kz);
  r.sz = 1.0f / comp(d, r.kz);
  return r;
}

// a * b - c * d with an exact sign: w = c*d and e = c*d - w exactly.
__device__ __forceinline__ float diff_of_products(float a, float b, float c, float d) {
  const float w = c * d;
  const float e = __builtin_fmaf(-c, d, w);
  const float f = __builtin_fmaf(a, b, -w);
  return f + e;
}

__device__ __forceinline__ bool intersect_triangle_wt(gg::V3 o, const WtRay& r, gg::V3 p0, gg::V3 p1, gg::V3 p2,
                                                      float& t, gg::V3& bary) {
  const gg::V3 A = p0 - o, B = p1 - o, C = p2 - o;
  const float Akz = comp(A, r.kz), Bkz = comp(B, r.kz), Ckz = comp(C, r.kz);
  const float Ax = comp(A, r.kx) - r.sx * Akz, Ay = comp(A, r.ky) - r.sy * Akz;
  const float Bx = comp(B, r.kx) - r.sx * Bkz, By = comp(B, r.ky) - r.sy * Bkz;
  const float Cx = comp(C, r.kx) - r.sx * Ckz, Cy = comp(C, r.ky) - r.sy * Ckz;
  const float U = diff_of_products(Cx, By, Cy, Bx);
  const float V = diff_of_products(Ax, Cy, Ay, Cx);
  const float W = diff_of_products(Bx, Ay, By, Ax);
  if ((U < 0.0f || V < 0.0f || W < 0.0f) && (U > 0.0f || V > 0.0f || W > 0.0f)) return false;
  const float det = U + V + W;
  if (det == 0.0f) return false;
  const float Az = r.sz * Akz, Bz = r.sz * Bkz, Cz = r.sz * Ckz;
  const float T = U * Az + V * Bz + W * Cz;
  const float inv = 1.0f / det;
  t = T * inv;
  bary.x = U * inv;
  bary.y = V * inv;
  bary.z = W * inv;
  return t > gg::kIntersectEpsilon;
}

// TraceRay / TraceRayAnyHit (traceray.glsl:18-180) for TRIANGLES_PER_LEAF == 2.
// ANY: returns at the first accepted triangle.  COUNT: box/triangle visit
// counters exactly as TRACERAY_DEBUG_VARS counts them.  WT: the watertight
// triangle test in place of IntersectTriangle (triNormal is then not set).
template <bool ANY, bool COUNT, bool WT = false>
__device__ __forceinline__ bool trace_ref(const RefScene& s, gg::V3 o, gg::V3 d, float& rayLength,
                                          RefHit& out, uint32_t& nbox, uint32_t& ntri,
                                          const ImportanceBreak& ib = ImportanceBreak{}) {
  float lastNodeImportance = 0.0f, lastNodeSizeSq = 0.0f;
  const WtRay wr = WT ? wt_setup(d) : WtRay{};
  int currentNodeIndex = 0;
  int currentLeafIndex = 0;
  int lastNodeIndex = 0;
  if (!ANY) { out.slot = -1; out.node = -1; }
  gg::V3 inv = gg::v3(1.0f / d.x, 1.0f / d.y, 1.0f / d.z);
  bool nextIsLeafNode = false;
  do {
    if (!nextIsLeafNode) {
      if (COUNT) ++nbox;
      float newHit, exitDist;
      bool boxHit;
      float4 a, b;
      if (s.obox) {  // OBOX_BVH (traceray.glsl:68-69)
        a = s.nodes[3 * currentNodeIndex];
        b = s.nodes[3 * currentNodeIndex + 1];
        boxHit = intersect_obox(o, d, a, b, s.nodes[3 * currentNodeIndex + 2], newHit, exitDist);
      } else {
        a = s.nodes[2 * currentNodeIndex];
        b = s.nodes[2 * currentNodeIndex + 1];
        boxHit = intersect_box(o, inv, gg::v3(a.x, a.y, a.z), gg::v3(b.x, b.y, b.z), newHit, exitDist);
      }
      uint32_t childCode = __float_as_uint(a.w);
      int escape = __float_as_int(b.w);
      bool broke = false;
      if (boxHit && newHit <= rayLength) {
        lastNodeIndex = currentNodeIndex;
        if (ib.importance) {
          const float2 importance = ib.importance[currentNodeIndex];
          // nodeSizeSq of FetchIntersect(O)BoxNode (intersectiontests.glsl:110,126)
          float sizeSq;
          if (s.obox) {
            const gg::V3 h = gg::v3(b.x, b.y, b.z);
            sizeSq = gg::dot(h, h) * 4.0f;
          } else {
            const gg::V3 e = gg::v3(b.x, b.y, b.z) - gg::v3(a.x, a.y, a.z);
            sizeSq = gg::dot(e, e);
          }
          if (!ANY) {
            lastNodeImportance = importance.x;
            lastNodeSizeSq = sizeSq;
          }
          if ((importance.x < ib.threshold && newHit > 0.0f) || importance.y <= ib.rayThreshold * newHit) {
            broke = true;
            if (ANY) {
              if (exitDist <= rayLength) return true;
            } else {
              out.nodeImportance = importance.x;
              out.nodeSizeSq = lastNodeSizeSq;
              out.node = currentNodeIndex;
              rayLength = newHit + (exitDist - newHit) *
                                       ((float)(gg::wanghash(__float_as_uint(exitDist)) % 8388593u) / 8388593.0f);
              out.slot = -1;
            }
            currentNodeIndex = escape;
          }
        }
        if (!broke) {
          currentNodeIndex = (int)(childCode & 0x7fffffffu);
          nextIsLeafNode = (uint32_t)currentNodeIndex != childCode;
          if (nextIsLeafNode) {
            currentLeafIndex = 2 * currentNodeIndex;
            currentNodeIndex = escape;
          }
        }
      } else {
        currentNodeIndex = escape;
      }
    }
    if (nextIsLeafNode) {
      int4 tri = s.slots[currentLeafIndex];
      if (tri.x == tri.y) {
        nextIsLeafNode = false;
      } else {
        if (COUNT) ++ntri;
        gg::V3 p0 = load_pos(s.positions, tri.x);
        gg::V3 p1 = load_pos(s.positions, tri.y);
        gg::V3 p2 = load_pos(s.positions, tri.z);
        float newHit;
        gg::V3 bary, n = gg::v3s(0.0f);
        const bool ok = WT ? intersect_triangle_wt(o, wr, p0, p1, p2, newHit, bary)
                           : intersect_triangle(o, d, p0, p1, p2, newHit, bary, n);
        if (ok && newHit < rayLength) {
          if (ANY) return true;
          rayLength = newHit;
          out.tri = tri;
          out.bary = bary;
          out.triNormal = n;
          out.node = lastNodeIndex;
          out.slot = currentLeafIndex;
          out.nodeImportance = lastNodeImportance;
          out.nodeSizeSq = lastNodeSizeSq;
        }
        ++currentLeafIndex;
        nextIsLeafNode = (currentLeafIndex % 2) != 0;
      }
    }
  } while (currentNodeIndex != 0 || nextIsLeafNode);
  return false;
}

}  // namespace gpugi
