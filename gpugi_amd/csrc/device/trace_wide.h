// Ordered traversal of the 4-wide quantized BVH (wide_bvh.h), one ray per lane.
//
// The walk is a state machine: wide_step() visits one node (four child boxes,
// the triangles of hit leaf children, ordered push of hit inner children) and
// reports whether the ray is finished.  trace_wide_ray() runs it to the end for
// one ray; the persistent trace kernel interleaves steps with wave-level ray
// refills instead.
//
// Per lane, a short stack of (node, entry distance) pairs lives in LDS laid
// out [depth][thread], so lanes at different depths touch different banks;
// deeper entries spill to a per-thread slice of global memory whose size the
// host derives from the exact worst case of the tree.  Children are visited
// nearest first, and a popped entry whose entry distance lies beyond the
// current hit is dropped without fetching its node.
//
// Box tests use the quantized boxes the host made conservative, slab distances
// computed with one fma per plane and a slack that covers their rounding (see
// wide_step), and axis-parallel direction components replaced by +-1e-20 so
// that 0 * inf never occurs: a box that contains the closest hit is never
// rejected.  The
// triangle test is the reference formula (intersectiontests.glsl:63-79), and
// equal distances resolve to the smaller reference slot index, which is the
// triangle the reference's preorder walk meets first.
#pragma once
#include <hip/hip_runtime.h>

#include "../common/gvec.h"
#include "trace_ref.h"
#include "wide_bvh.h"

namespace gpugi {

struct WideResult {
  float t;     // final ray length (tmax on a miss)
  int slot;    // reference slot index, -1 on a miss, -2 on stack overflow
  float by, bz;
  bool any;    // any-hit result
};

struct WideState {
  gg::V3 o, d;
  float ix, iy, iz;
  float tbest;
  int bestSlot;
  float bu, bv;
  uint32_t node;
  int sp;
  bool found;
  WtRay wr;  // watertight mode only
};

__device__ __forceinline__ float safe_inv(float d) {
  const float tiny = 1e-20f;
  return 1.0f / (fabsf(d) > tiny ? d : copysignf(tiny, d));
}

__device__ __forceinline__ float byte_f(uint32_t w, int k) { return (float)((w >> (8 * k)) & 0xffu); }

template <bool WT = false>
__device__ __forceinline__ void wide_init(WideState& st, gg::V3 o, gg::V3 d, float tmax) {
  st.o = o;
  st.d = d;
  st.ix = safe_inv(d.x);
  st.iy = safe_inv(d.y);
  st.iz = safe_inv(d.z);
  st.tbest = tmax;
  st.bestSlot = -1;
  st.bu = st.bv = 0.0f;
  st.node = 0;
  st.sp = 0;
  st.found = false;
  if (WT) st.wr = wt_setup(d);
}

// One node visit.  Returns true when the ray is finished.  WT: the watertight
// triangle test on the raw-vertex triangle layout {p0, slot}{p1, mat}{p2, 0}.
template <bool ANY, int B, int S, bool WT = false>
__device__ __forceinline__ bool wide_step(const WideScene& s, WideState& st, uint32_t* stkNode, float* stkT, int tid,
                                          uint2* spill) {
  const gg::V3 o = st.o, d = st.d;
  const uint32_t node = st.node;
  const uint4 q0 = s.nodes[4 * (size_t)node + 0];
  const uint4 q1 = s.nodes[4 * (size_t)node + 1];
  const uint4 q2 = s.nodes[4 * (size_t)node + 2];
  const uint4 q3 = s.nodes[4 * (size_t)node + 3];
  const float ox = __uint_as_float(q0.x), oy = __uint_as_float(q0.y), oz = __uint_as_float(q0.z);
  const float sx = __uint_as_float((q0.w & 0xffu) << 23);
  const float sy = __uint_as_float(((q0.w >> 8) & 0xffu) << 23);
  const float sz = __uint_as_float(((q0.w >> 16) & 0xffu) << 23);
  const int nchild = (int)(q0.w >> 24);
  const uint32_t childs[4] = {q1.x, q1.y, q1.z, q1.w};
  // Slab distances straight from the quantized bytes: along axis a,
  //   t(q) = (origin_a + q * s_a - o_a) / d_a ~ q * A_a + B_a,
  //   A_a = s_a * inv_a (exact: s_a is a power of two),  B_a = (origin_a - o_a) * inv_a.
  // B_a carries at most ~3 roundings of |B_a|, so near distances use B_a - e_a and
  // far ones B_a + e_a with e_a = |B_a| * 2^-21; the byte word giving the near
  // plane (lo or hi) is picked per axis from the sign of inv_a.  The host checks
  // in extended precision that origin + q * s encloses every child box, so with
  // the far distance widened by 1 + 2^-20 a box that holds the closest hit is
  // never rejected (the same argument as Ize 2013, "Robust BVH ray traversal").
  const float Ax = sx * st.ix, Ay = sy * st.iy, Az = sz * st.iz;
  const float Bx = (ox - o.x) * st.ix, By = (oy - o.y) * st.iy, Bz = (oz - o.z) * st.iz;
  const float ex = fabsf(Bx) * 4.76837158e-7f, ey = fabsf(By) * 4.76837158e-7f, ez = fabsf(Bz) * 4.76837158e-7f;
  const float bnx = Bx - ex, bny = By - ey, bnz = Bz - ez;
  const float bfx = Bx + ex, bfy = By + ey, bfz = Bz + ez;
  const uint32_t nearX = st.ix >= 0.0f ? q2.x : q2.w, farX = st.ix >= 0.0f ? q2.w : q2.x;
  const uint32_t nearY = st.iy >= 0.0f ? q2.y : q3.x, farY = st.iy >= 0.0f ? q3.x : q2.y;
  const uint32_t nearZ = st.iz >= 0.0f ? q2.z : q3.y, farZ = st.iz >= 0.0f ? q3.y : q2.z;
  const float tbScaled = st.tbest * 1.00000095f;
  float tn[4];
  bool hit[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const float nx = __builtin_fmaf(byte_f(nearX, k), Ax, bnx);
    const float ny = __builtin_fmaf(byte_f(nearY, k), Ay, bny);
    const float nz = __builtin_fmaf(byte_f(nearZ, k), Az, bnz);
    const float fx = __builtin_fmaf(byte_f(farX, k), Ax, bfx);
    const float fy = __builtin_fmaf(byte_f(farY, k), Ay, bfy);
    const float fz = __builtin_fmaf(byte_f(farZ, k), Az, bfz);
    const float tnear = fmaxf(fmaxf(nx, ny), fmaxf(nz, 0.0f));
    const float tfar = fminf(fminf(fx, fy), fz) * 1.00000095f;
    tn[k] = tnear;
    hit[k] = (k < nchild) && tnear <= tfar && tnear <= tbScaled;
  }
  // Leaf children first: they can only shorten tbest.  The triangles of all
  // hit leaf children form one list per lane and each lane walks its own, so a
  // wave spends max over lanes of their triangle counts, not the sum over the
  // four child slots of per-slot maxima.
  uint32_t leafMask = 0;
#pragma unroll
  for (int k = 0; k < 4; ++k) leafMask |= (hit[k] && (childs[k] & kWideLeafFlag)) ? (1u << k) : 0u;
  uint32_t m = 0;
  while (leafMask) {
    const int k = __builtin_ctz(leafMask);
    const uint32_t ck = k == 0 ? childs[0] : (k == 1 ? childs[1] : (k == 2 ? childs[2] : childs[3]));
    const uint32_t first = ck & kWideLeafFirstMask;
    const uint32_t cnt = ((ck >> 27) & 0xfu) + 1u;
    const size_t ti = 3 * (size_t)(first + m);
    const float4 a = s.tris[ti + 0];
    const float4 b = s.tris[ti + 1];
    const float4 c = s.tris[ti + 2];
    const gg::V3 p0 = gg::v3(a.x, a.y, a.z);
    const int slot = __float_as_int(a.w);
    float t, by, bz;
    bool ok;
    if (WT) {
      gg::V3 bary;
      ok = intersect_triangle_wt(o, st.wr, p0, gg::v3(b.x, b.y, b.z), gg::v3(c.x, c.y, c.z), t, bary);
      by = bary.y;
      bz = bary.z;
    } else {
      const gg::V3 e0 = gg::v3(b.x, b.y, b.z);
      const gg::V3 e1 = gg::v3(c.x, c.y, c.z);
      const gg::V3 nn = gg::cross(e1, e0);
      const gg::V3 e2 = (1.0f / gg::dot(nn, d)) * (p0 - o);
      const gg::V3 ii = gg::cross(d, e2);
      by = gg::dot(ii, e1);
      bz = gg::dot(ii, e0);
      const float bx = 1.0f - (bz + by);
      t = gg::dot(nn, e2);
      ok = t > 0.0001f && bx >= 0.0f && by >= 0.0f && bz >= 0.0f;
    }
    if (ANY) {
      if (ok && t < st.tbest) {
        st.found = true;
        break;
      }
    } else if (ok && (t < st.tbest || (t == st.tbest && slot < st.bestSlot))) {
      st.tbest = t;
      st.bestSlot = slot;
      st.bu = by;
      st.bv = bz;
    }
    if (++m == cnt) {
      m = 0;
      leafMask &= leafMask - 1u;
    }
  }
  if (ANY && st.found) return true;
  // Inner children, nearest first.
  const float tb = st.tbest * 1.00000095f;  // entry distances are lower bounds with the same slack
  float t0 = (hit[0] && !(childs[0] & kWideLeafFlag) && tn[0] <= tb) ? tn[0] : INFINITY;
  float t1 = (hit[1] && !(childs[1] & kWideLeafFlag) && tn[1] <= tb) ? tn[1] : INFINITY;
  float t2 = (hit[2] && !(childs[2] & kWideLeafFlag) && tn[2] <= tb) ? tn[2] : INFINITY;
  float t3 = (hit[3] && !(childs[3] & kWideLeafFlag) && tn[3] <= tb) ? tn[3] : INFINITY;
  uint32_t c0 = childs[0], c1 = childs[1], c2 = childs[2], c3 = childs[3];
  if (!ANY) {
#define GPUGI_CSWAP(ta, ca, tb_, cb)    \
  if (tb_ < ta) {                       \
    float tt = ta; ta = tb_; tb_ = tt;  \
    uint32_t cc = ca; ca = cb; cb = cc; \
  }
    GPUGI_CSWAP(t0, c0, t1, c1);
    GPUGI_CSWAP(t2, c2, t3, c3);
    GPUGI_CSWAP(t0, c0, t2, c2);
    GPUGI_CSWAP(t1, c1, t3, c3);
    GPUGI_CSWAP(t1, c1, t2, c2);
#undef GPUGI_CSWAP
  }
  bool overflow = false;
  int sp = st.sp;
#define GPUGI_PUSH(nd, tv)                                   \
  do {                                                       \
    if (sp < S) {                                            \
      stkNode[sp * B + tid] = (nd);                          \
      stkT[sp * B + tid] = (tv);                             \
      ++sp;                                                  \
    } else if ((uint32_t)(sp - S) < s.spillDepth) {          \
      spill[sp - S] = make_uint2((nd), __float_as_uint(tv)); \
      ++sp;                                                  \
    } else {                                                 \
      overflow = true;                                       \
    }                                                        \
  } while (0)
  // Sorted (closest hit): finite entries form a prefix, c0 is the nearest.
  // Unsorted (any hit): the first finite entry is taken next.
  uint32_t next = 0xffffffffu;
  if (t3 != INFINITY) { if (ANY && next == 0xffffffffu) next = c3; else GPUGI_PUSH(c3, t3); }
  if (t2 != INFINITY) { if (ANY && next == 0xffffffffu) next = c2; else GPUGI_PUSH(c2, t2); }
  if (t1 != INFINITY) { if (ANY && next == 0xffffffffu) next = c1; else GPUGI_PUSH(c1, t1); }
  if (t0 != INFINITY) { if (next == 0xffffffffu) next = c0; else GPUGI_PUSH(c0, t0); }
#undef GPUGI_PUSH
  if (overflow) {  // impossible when the host sized the spill from the tree; fail loudly
    atomicOr(s.errorFlag, 1u);
    st.bestSlot = -2;
    st.sp = sp;
    return true;
  }
  if (next != 0xffffffffu) {
    st.node = next;
    st.sp = sp;
    return false;
  }
  // Pop the nearest pending node still in front of the closest hit.
  while (sp > 0) {
    --sp;
    uint32_t nd;
    float t;
    if (sp < S) {
      nd = stkNode[sp * B + tid];
      t = stkT[sp * B + tid];
    } else {
      const uint2 e = spill[sp - S];
      nd = e.x;
      t = __uint_as_float(e.y);
    }
    if (t <= st.tbest * 1.00000095f) {
      st.node = nd;
      st.sp = sp;
      return false;
    }
  }
  st.sp = 0;
  return true;
}

// stkNode/stkT: LDS arrays of S * B entries; tid in [0, B).
template <bool ANY, int B, int S, bool WT = false>
__device__ __forceinline__ WideResult trace_wide_ray(const WideScene& s, gg::V3 o, gg::V3 d, float tmax,
                                                     uint32_t* stkNode, float* stkT, int tid, uint2* spill) {
  WideState st;
  wide_init<WT>(st, o, d, tmax);
  while (!wide_step<ANY, B, S, WT>(s, st, stkNode, stkT, tid, spill)) {
  }
  WideResult r;
  r.t = st.tbest;
  r.slot = st.bestSlot;
  r.by = st.bu;
  r.bz = st.bv;
  r.any = st.found;
  return r;
}

}  // namespace gpugi
