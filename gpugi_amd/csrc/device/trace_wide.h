// Ordered traversal of the 4-wide quantized BVH (wide_bvh.h), one ray per lane.
//
// A ray's state holds the next inner node to visit, up to two pending leaves
// and a stack of (child word, entry distance) pairs.  Work comes in two kinds
// of wave-wide passes:
//   box pass   every lane with a node visits it: four child box tests, the
//              nearest hit inner child becomes the next node, the nearest hit
//              leaf takes a free pending slot, the other hit children are
//              pushed farthest first;
//   leaf pass  every lane with pending leaves tests their triangles.
// A lane with a pending leaf keeps traversing (speculative traversal,
// Aila and Laine 2009, "Understanding the efficiency of ray traversal on
// GPUs"), so a wave runs a leaf pass only once enough lanes wait for one and
// does not pay a divergent triangle loop on every node visit.  The caller
// chooses the pass from two ballots (wide_choose_leaf_pass).  Lanes left
// without a node pop: entries beyond the current hit are dropped, a leaf
// entry fills a free pending slot, an inner entry becomes the next node.
// The result does not depend on the order of the work: the closest hit is the
// minimum distance, ties resolved to the smaller reference slot index (the
// triangle the reference's preorder walk meets first).
//
// Per lane, the first S stack entries live in LDS laid out [depth][thread],
// so lanes at different depths touch different banks; deeper entries spill to
// global memory (WideStack), sized by the host from the exact worst case of
// the tree.
//
// Box tests use the quantized boxes the host made conservative, slab distances
// computed with one fma per plane and a slack that covers their rounding (see
// wide_visit), and axis-parallel direction components replaced by +-1e-20 so
// that 0 * inf never occurs: a box that contains the closest hit is never
// rejected.  The triangle test is the reference formula
// (intersectiontests.glsl:63-79) or, in watertight mode, the watertight test.
#pragma once
#include <hip/hip_runtime.h>

#include "../common/gvec.h"
#include "trace_ref.h"
#include "wide_bvh.h"

namespace gpugi {

constexpr uint32_t kWideNoNode = 0xffffffffu;

struct WideResult {
  float t;     // final ray length (tmax on a miss)
  int slot;    // reference slot index, -1 on a miss, -2 on stack overflow
  float by, bz;
  bool any;    // any-hit result
};

// A lane's traversal stack of {child word, entry distance bits}: entry i < S at
// LDS lds[i * B + tid], deeper entries at spill[(i - S) * spillStride]
// (spillStride = number of threads sharing the spill buffer when it is laid out
// [depth][thread], so a wave's spills at one depth are contiguous; 1 for a
// per-thread slice).
// The LDS part is typed as LDS memory, so its accesses stay ds_* instructions
// and are never merged with the spill's global accesses into flat ones.
// (The host pass of the compiler sees the device code too and has no LDS.)
#if defined(__HIP_DEVICE_COMPILE__)
typedef __attribute__((address_space(3))) uint2 LdsEntry;
#else
typedef uint2 LdsEntry;
#endif
__device__ __forceinline__ LdsEntry* as_lds(uint2* p) { return (LdsEntry*)p; }

struct WideStack {
  LdsEntry* lds;
  int tid;
  uint2* spill;
  uint32_t spillStride;
};

// Pending-leaf slots per lane, filled in order (slot j is taken only while
// slots 0..j-1 are).
constexpr int kWidePending = 2;

struct WideState {
  gg::V3 o, d;
  float ix, iy, iz;
  float tbest;
  int bestSlot;
  float bu, bv;
  uint32_t node;  // next inner node to visit, kWideNoNode if none
  uint32_t leaf[kWidePending];  // pending leaf child words (leaf flag set), 0 if free
  int sp;
  bool found;
  WtRay wr;  // watertight mode only
};

// 1/d by v_rcp_f32 (within 1 ulp): a reciprocal with relative error
// below 2^-23 scales every slab distance of a ray by the same factor (1 + δ),
// which the far-distance and hit-distance widening of 1 + 2^-20 covers.
__device__ __forceinline__ float safe_inv(float d) {
  const float tiny = 1e-20f;
  return __builtin_amdgcn_rcpf(fabsf(d) > tiny ? d : copysignf(tiny, d));
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
#pragma unroll
  for (int j = 0; j < kWidePending; ++j) st.leaf[j] = 0;
  st.sp = 0;
  st.found = false;
  if (WT) st.wr = wt_setup(d);
}

// Box pass for a lane with st.node != kWideNoNode.  Returns false if the stack
// would overflow (impossible when the host sized the spill from the tree).
template <bool ANY, int B, int S>
__device__ __forceinline__ bool wide_visit(const WideScene& s, WideState& st, const WideStack& stk) {
  const gg::V3 o = st.o;
  const uint32_t nodeIndex = st.node;
  const uint4 q0 = s.nodes[4 * (size_t)nodeIndex + 0];
  const uint4 q1 = s.nodes[4 * (size_t)nodeIndex + 1];
  const uint4 q2 = s.nodes[4 * (size_t)nodeIndex + 2];
  const uint4 q3 = s.nodes[4 * (size_t)nodeIndex + 3];
  const float ox = __uint_as_float(q0.x), oy = __uint_as_float(q0.y), oz = __uint_as_float(q0.z);
  const float sx = __uint_as_float((q0.w & 0xffu) << 23);
  const float sy = __uint_as_float(((q0.w >> 8) & 0xffu) << 23);
  const float sz = __uint_as_float(((q0.w >> 16) & 0xffu) << 23);
  const int nchild = (int)(q0.w >> 24);
  // Slab distances straight from the quantized bytes: along axis a,
  //   t(q) = (origin_a + q * s_a - o_a) / d_a ~ q * A_a + B_a,
  //   A_a = s_a * inv_a (exact: s_a is a power of two),  B_a = (origin_a - o_a) * inv_a.
  // With inv_a = (1/d_a)(1 + δ), |δ| <= 2^-23 (safe_inv), every exact distance of
  // the ray is scaled by 1 + δ; beyond that B_a carries two roundings, so near
  // distances use B_a - e_a and far ones B_a + e_a with e_a = |B_a| * 2^-21; the
  // byte word giving the near plane (lo or hi) is picked per axis from the sign of
  // inv_a.  The host checks in extended precision that origin + q * s encloses
  // every child box, so with the far distance and the hit distance widened by
  // 1 + 2^-20 (> 1 + 2^-23 + the fma's rounding) a box that holds the closest hit
  // is never rejected (the same argument as Ize 2013, "Robust BVH ray traversal").
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
    tn[k] = ((k < nchild) && tnear <= tfar && tnear <= tbScaled) ? tnear : INFINITY;
  }
  float t0 = tn[0], t1 = tn[1], t2 = tn[2], t3 = tn[3];
  uint32_t c0 = q1.x, c1 = q1.y, c2 = q1.z, c3 = q1.w;
  if (!ANY) {  // nearest first; missed children (INFINITY) sink to the end
#define GPUGI_CSWAP(ta, ca, tb_, cb)                       \
  {                                                        \
    const bool sw = tb_ < ta;                              \
    const float lo = sw ? tb_ : ta, hi = sw ? ta : tb_;    \
    const uint32_t cl = sw ? cb : ca, ch = sw ? ca : cb;   \
    ta = lo;                                               \
    tb_ = hi;                                              \
    ca = cl;                                               \
    cb = ch;                                               \
  }
    GPUGI_CSWAP(t0, c0, t1, c1);
    GPUGI_CSWAP(t2, c2, t3, c3);
    GPUGI_CSWAP(t0, c0, t2, c2);
    GPUGI_CSWAP(t1, c1, t3, c3);
    GPUGI_CSWAP(t1, c1, t2, c2);
#undef GPUGI_CSWAP
  }
  // The nearest hit inner child is visited next, the nearest hit leaves fill the
  // free pending slots, every other hit child is pushed (p_k).  Ranks count the
  // hit inner and hit leaf children nearer than child k; the claims are selects.
  const float tk[4] = {t0, t1, t2, t3};
  const uint32_t ck[4] = {c0, c1, c2, c3};
  int used = 0;
#pragma unroll
  for (int j = 0; j < kWidePending; ++j) used += st.leaf[j] != 0u ? 1 : 0;
  const int freeSlots = kWidePending - used;
  uint32_t node = kWideNoNode;
  uint32_t claimed[kWidePending];
#pragma unroll
  for (int j = 0; j < kWidePending; ++j) claimed[j] = 0u;
  int nInner = 0, nLeaf = 0;
  bool pk[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const bool hit = tk[k] != INFINITY;
    const bool isLeaf = (int32_t)ck[k] < 0;
    const bool takeNode = hit && !isLeaf && nInner == 0;
    const bool takeLeaf = hit && isLeaf && nLeaf < freeSlots;
    node = takeNode ? ck[k] : node;
#pragma unroll
    for (int j = 0; j < kWidePending; ++j) claimed[j] = (takeLeaf && nLeaf == j) ? ck[k] : claimed[j];
    nInner += (hit && !isLeaf) ? 1 : 0;
    nLeaf += (hit && isLeaf) ? 1 : 0;
    pk[k] = hit && !takeNode && !takeLeaf;
  }
  // free slot j (j >= used) takes claimed[j - used]
  uint32_t leaf[kWidePending];
#pragma unroll
  for (int j = 0; j < kWidePending; ++j) {
    uint32_t w = st.leaf[j];
#pragma unroll
    for (int u = 0; u <= j; ++u) w = (used == u) ? claimed[j - u] : w;
    leaf[j] = w;
  }
  const bool p0 = pk[0], p1 = pk[1], p2 = pk[2], p3 = pk[3];
  bool overflow = false;
  int sp = st.sp;
  if (sp + 4 <= S) {  // every push stays in LDS: one store and one increment each
    LdsEntry* e = stk.lds + sp * B + stk.tid;
    int pushed = 0;
    if (p3) { e[pushed * B] = make_uint2(c3, __float_as_uint(t3)); ++pushed; }
    if (p2) { e[pushed * B] = make_uint2(c2, __float_as_uint(t2)); ++pushed; }
    if (p1) { e[pushed * B] = make_uint2(c1, __float_as_uint(t1)); ++pushed; }
    if (p0) { e[pushed * B] = make_uint2(c0, __float_as_uint(t0)); ++pushed; }
    sp += pushed;
  } else {
#define GPUGI_PUSH(nd, tv)                                                                  \
  do {                                                                                      \
    if (sp < S) {                                                                           \
      stk.lds[sp * B + stk.tid] = make_uint2((nd), __float_as_uint(tv));                    \
      ++sp;                                                                                 \
    } else if ((uint32_t)(sp - S) < s.spillDepth) {                                         \
      stk.spill[(size_t)(sp - S) * stk.spillStride] = make_uint2((nd), __float_as_uint(tv)); \
      ++sp;                                                                                 \
    } else {                                                                                \
      overflow = true;                                                                      \
    }                                                                                       \
  } while (0)
    if (p3) GPUGI_PUSH(c3, t3);
    if (p2) GPUGI_PUSH(c2, t2);
    if (p1) GPUGI_PUSH(c1, t1);
    if (p0) GPUGI_PUSH(c0, t0);
#undef GPUGI_PUSH
  }
  st.node = node;
#pragma unroll
  for (int j = 0; j < kWidePending; ++j) st.leaf[j] = leaf[j];
  st.sp = sp;
  if (overflow) {
    atomicOr(s.errorFlag, 1u);
    st.bestSlot = -2;
    return false;
  }
  return true;
}

// Leaf pass for a lane with pending leaves: tests their triangles (one loop
// over the slots in order) and frees every slot.
template <bool ANY, bool WT>
__device__ __forceinline__ void wide_leaf(const WideScene& s, WideState& st) {
  const gg::V3 o = st.o, d = st.d;
  uint32_t first[kWidePending], end[kWidePending];  // end[j]: triangles of slots 0..j
  uint32_t cnt = 0;
#pragma unroll
  for (int j = 0; j < kWidePending; ++j) {
    const uint32_t w = st.leaf[j];
    first[j] = (w & kWideLeafFirstMask) - cnt;  // triangle of loop index m is first[j] + m
    cnt += w != 0u ? ((w >> 27) & 0xfu) + 1u : 0u;
    end[j] = cnt;
    st.leaf[j] = 0;
  }
  for (uint32_t m = 0; m < cnt; ++m) {
    uint32_t tri = first[kWidePending - 1] + m;
#pragma unroll
    for (int j = kWidePending - 2; j >= 0; --j) tri = m < end[j] ? first[j] + m : tri;
    const size_t ti = 3 * (size_t)tri;
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
        return;
      }
    } else if (ok && (t < st.tbest || (t == st.tbest && slot < st.bestSlot))) {
      st.tbest = t;
      st.bestSlot = slot;
      st.bu = by;
      st.bv = bz;
    }
  }
}

// For a lane without a next node: pops entries still in front of the closest
// hit until one yields a node, or a leaf finds both pending slots taken (the
// entry stays), or the stack is empty.
template <int B, int S>
__device__ __forceinline__ void wide_pop(WideState& st, const WideStack& stk) {
  const float tb = st.tbest * 1.00000095f;  // entry distances are lower bounds with this slack
  int sp = st.sp;
  uint32_t leaf[kWidePending];
#pragma unroll
  for (int j = 0; j < kWidePending; ++j) leaf[j] = st.leaf[j];
  uint32_t node = kWideNoNode;
  while (sp > 0) {
    const int top = sp - 1;
    uint32_t nd;
    float t;
    if (top < S) {
      const uint2 e = stk.lds[top * B + stk.tid];
      nd = e.x;
      t = __uint_as_float(e.y);
    } else {
      const uint2 e = stk.spill[(size_t)(top - S) * stk.spillStride];
      nd = e.x;
      t = __uint_as_float(e.y);
    }
    if (t > tb) {
      sp = top;
      continue;
    }
    if (nd & kWideLeafFlag) {
      if (leaf[kWidePending - 1] != 0u) break;
      bool placed = false;
#pragma unroll
      for (int j = 0; j < kWidePending; ++j) {
        const bool take = !placed && leaf[j] == 0u;
        leaf[j] = take ? nd : leaf[j];
        placed = placed || take;
      }
      sp = top;
      continue;
    }
    node = nd;
    sp = top;
    break;
  }
  st.sp = sp;
#pragma unroll
  for (int j = 0; j < kWidePending; ++j) st.leaf[j] = leaf[j];
  st.node = node;
}

// Wave-uniform pass choice for the working lanes (mask `working`, a ballot
// taken by the caller): a leaf pass when no working lane has a node to visit
// or at least `waitLanes` of them have only a pending leaf.
__device__ __forceinline__ bool wide_choose_leaf_pass(uint64_t working, const WideState& st, int waitLanes) {
  const uint64_t hasNode = __ballot(st.node != kWideNoNode);
  return (working & hasNode) == 0 || __popcll(working & ~hasNode) >= waitLanes;
}

// One pass for a working lane.  Returns true when the ray is finished.
template <bool ANY, int B, int S, bool WT>
__device__ __forceinline__ bool wide_pass(const WideScene& s, WideState& st, bool leafPass, const WideStack& stk) {
  if (leafPass) {
    if (st.leaf[0] == 0u) return false;  // still traversing; unaffected by this pass
    wide_leaf<ANY, WT>(s, st);
    if (ANY && st.found) return true;
  } else {
    if (st.node == kWideNoNode) return false;  // waiting for a leaf pass
    if (!wide_visit<ANY, B, S>(s, st, stk)) return true;  // overflow, slot -2
  }
  if (st.node == kWideNoNode) {
    wide_pop<B, S>(st, stk);
    return st.node == kWideNoNode && st.leaf[0] == 0u;
  }
  return false;
}

// Traces one ray per calling lane to the end; the calling lanes form the wave
// for pass choices.  lds: LDS array of S * B stack entries; tid in [0, B);
// spill: this thread's own slice of s.spillDepth entries.
template <bool ANY, int B, int S, bool WT = false>
__device__ __forceinline__ WideResult trace_wide_ray(const WideScene& s, gg::V3 o, gg::V3 d, float tmax,
                                                     uint2* lds, int tid, uint2* spill) {
  const WideStack stk{as_lds(lds), tid, spill, 1u};
  WideState st;
  wide_init<WT>(st, o, d, tmax);
  bool live = true;
  for (;;) {
    const uint64_t act = __ballot(live);
    if (act == 0) break;
    const bool leafPass = wide_choose_leaf_pass(act, st, (__popcll(act) + 1) / 2);
    if (live && wide_pass<ANY, B, S, WT>(s, st, leafPass, stk)) live = false;
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
