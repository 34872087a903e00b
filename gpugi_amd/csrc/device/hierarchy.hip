// Importance and material passes of the HierarchyImportance renderer
// (renderer/hierarchyimportance.cpp:84-178):
//   k_hier_acquire        acquisition2.comp: per pixel one path of up to 8
//                         vertices; every ray (camera/bounce ray up to its hit,
//                         blocked light ray up to its blocker) adds the path's
//                         importance to the .x of each node box it enters and
//                         each triangle it hits (counthits.glsl), by float
//                         atomics; the importance halves at every vertex;
//   k_hier_init           hierarchypropagation_init.comp: x /= iterations,
//                         leaf nodes' subtree importance and approximation y;
//   k_hier_nodes          hierarchypropagation_nodes.comp for one tree level:
//                         subtree importance = sum over the children, y;
//   k_hier_leaf_materials sampleleafmaterials.comp: area-weighted material of
//                         each leaf node's triangles at their centroids, packed
//                         to halves;
//   k_hier_pull           pullmaterials.comp for one tree level.
// The reference repeats the propagation and pull passes over all nodes until
// nothing changes; here the host launches one pass per tree level, deepest
// first, over that level's nodes, which gives the converged values directly
// (every node's value depends only on its children's final values, summed in
// sibling order).
#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>

#include <algorithm>
#include <cstring>

#include "hierarchy.h"
#include "integrator_common.h"

namespace gpugi {

namespace {

constexpr int kHierBlock = 256;

struct NodeView {
  const float4* nodes;
  int obox;
  __device__ __forceinline__ float4 t0(int n) const { return nodes[(obox ? 3 : 2) * n]; }
  __device__ __forceinline__ float4 t1(int n) const { return nodes[(obox ? 3 : 2) * n + 1]; }
};

// The acquisition sums each entry's importance in 64-bit fixed point, in
// units of 2^-kImpFracBits (the smallest addend: the path importance halves at
// each of the kHierPathLength vertices and once more for the light sample), so
// every sum is exact and independent of the order of the additions.
constexpr int kImpFracBits = 8;
static_assert(kHierPathLength <= kImpFracBits, "every addend is a whole number of units");

// acc[key] += w for every calling lane.  Lanes with the same (key, w) are
// combined into one atomic of w times their number, up to kCombineGroups
// groups per call (each group led by its lowest lane); the lanes left after
// that add on their own.  The rays of a wave enter the top nodes together and
// split into a few groups below them, which would otherwise take one
// contended atomic per ray.
constexpr int kCombineGroups = 8;
// children an inner node can have in the count walk's child loop (the
// reference trees are binary)
constexpr int kMaxChildren = 16;

// w may be negative (a node's term w * (1 - c) below); the sums wrap around in
// 64 bits and end non-negative.
__device__ __forceinline__ void add_importance(unsigned long long* __restrict__ acc, uint32_t key, float w) {
  const unsigned long long mag = (unsigned long long)(fabsf(w) * (float)(1 << kImpFracBits));  // exact: |w| = m 2^-k
  const unsigned long long units = w < 0.0f ? 0ull - mag : mag;
  const int lane = (int)(threadIdx.x & 63);
  uint64_t remaining = __ballot(true);  // wave-uniform: the lanes whose addition is still to be made
  for (int g = 0; g < kCombineGroups && remaining != 0; ++g) {
    const int leader = __builtin_ctzll(remaining);  // a lane still in the loop
    const uint32_t lk = (uint32_t)__shfl((int)key, leader);
    const float lw = __shfl(w, leader);
    const uint64_t group = __ballot(key == lk && w == lw) & remaining;
    if (lane == leader) atomicAdd(&acc[lk], units * (unsigned long long)__popcll(group));
    remaining &= ~group;
    if ((group >> lane) & 1ull) return;  // this lane's addition was made
  }
  atomicAdd(&acc[key], units);
}

// Whether the count walk enters node n (its box is hit within rayLength); a
// and b receive the node's first two texels (child code in a.w, escape in b.w).
__device__ __forceinline__ bool count_enters(const RefScene& s, gg::V3 o, gg::V3 d, gg::V3 inv, float rayLength,
                                             int n, float4& a, float4& b) {
  float newHit, exitDist;
  bool boxHit;
  if (s.obox) {
    a = s.nodes[3 * n];
    b = s.nodes[3 * n + 1];
    boxHit = intersect_obox(o, d, a, b, s.nodes[3 * n + 2], newHit, exitDist);
  } else {
    a = s.nodes[2 * n];
    b = s.nodes[2 * n + 1];
    boxHit = intersect_box(o, inv, gg::v3(a.x, a.y, a.z), gg::v3(b.x, b.y, b.z), newHit, exitDist);
  }
  return boxHit && newHit <= rayLength;
}

// TraceRayCountHits (counthits.glsl:1-71) over the reference layout.  The
// nodes a walk enters are closed under parents, so instead of w at every
// entered node it adds w * (1 - c), c being how many of the node's children
// the walk enters (tested here with the walk's own test), and
// k_hier_subtree_sum later gives each node the sum over its subtree, which is
// w for each walk that entered it.  Nodes with one entered child, most of a
// walk, then need no atomic.  Triangles keep their own w.
__device__ void count_hits(const RefScene& s, gg::V3 o, gg::V3 d, float rayLength, float w,
                           unsigned long long* __restrict__ imp, uint32_t numNodes) {
  int node = 0, leaf = 0;
  bool nextIsLeaf = false;
  const gg::V3 inv = gg::v3(1.0f / d.x, 1.0f / d.y, 1.0f / d.z);
  do {
    if (!nextIsLeaf) {
      float4 a, b;
      const bool entered = count_enters(s, o, d, inv, rayLength, node, a, b);
      const uint32_t childCode = __float_as_uint(a.w);
      const int escape = __float_as_int(b.w);
      if (entered) {
        const int child = (int)(childCode & 0x7fffffffu);
        nextIsLeaf = (uint32_t)child != childCode;
        int c = 0;
        if (!nextIsLeaf) {
          // the children: the first child and the siblings its escape chain
          // reaches before this node's escape (the parent rule of BuildHierTree)
          const int end = escape == 0 ? (int)numNodes : escape;
          int k = child;
          for (int it = 0; it < kMaxChildren && k > node && k < end; ++it) {
            float4 ca, cb;
            if (count_enters(s, o, d, inv, rayLength, k, ca, cb)) ++c;
            const int next = __float_as_int(cb.w);
            if (next == 0) break;
            k = next;
          }
        }
        if (c != 1) add_importance(imp, (uint32_t)node, w * (float)(1 - c));
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
      const int4 tri = s.slots[leaf];
      if (tri.x == tri.y) {
        nextIsLeaf = false;
      } else {
        float newHit;
        gg::V3 bary, n;
        if (intersect_triangle(o, d, load_pos(s.positions, tri.x), load_pos(s.positions, tri.y),
                               load_pos(s.positions, tri.z), newHit, bary, n) &&
            newHit <= rayLength)
          add_importance(imp, numNodes + (uint32_t)leaf, w);
        ++leaf;
        nextIsLeaf = (leaf % 2) != 0;
      }
    }
  } while (node != 0 || nextIsLeaf);
}

// Eight waves per SIMD (64 VGPRs, a 48-byte spill) instead of five (88 VGPRs,
// no spill): the two walks per ray wait on memory, and the extra waves cover
// it (93 against 123 ms per 10-pass update on a 2M-triangle heightfield).
__global__ __launch_bounds__(kHierBlock) __attribute__((amdgpu_waves_per_eu(8))) void k_hier_acquire(
    DrawParams p, uint32_t frameSeed, unsigned long long* __restrict__ imp, uint32_t numNodes, uint32_t tilesW,
    uint32_t numPixels) {
  const uint32_t i = blockIdx.x * kHierBlock + threadIdx.x;
  if (i >= numPixels) return;
  // pixel of thread i: 8x8 tiles in row order (the 8x8 work groups of the dispatch)
  const uint32_t tile = i / kTileThreads, k = i % kTileThreads;
  const int gx = (int)((tile % tilesW) * kTile + k % kTile), gy = (int)((tile / tilesW) * kTile + k / kTile);
  uint32_t seed = gg::init_random_seed(frameSeed, (uint32_t)(gx + gy * p.width));
  const float j0 = gg::random01(seed);
  const float j1 = gg::random01(seed);
  // (Random2 + gridPosition) / BackbufferSize * 2 - 1 (acquisition2.comp:28): no half-pixel shift
  const float sx = ((j0 + (float)gx) / (float)p.width) * 2.0f - 1.0f;
  const float sy = ((j1 + (float)gy) / (float)p.height) * 2.0f - 1.0f;
  const gg::V3 U = gg::v3(p.camera.u[0], p.camera.u[1], p.camera.u[2]);
  const gg::V3 V = gg::v3(p.camera.v[0], p.camera.v[1], p.camera.v[2]);
  const gg::V3 W = gg::v3(p.camera.w[0], p.camera.w[1], p.camera.w[2]);
  gg::V3 o = gg::v3(p.camera.position[0], p.camera.position[1], p.camera.position[2]);
  gg::V3 d = gg::normalize(sx * U + sy * V + W);
  float pathImportance = 1.0f;
  for (int v = 0; v < kHierPathLength; ++v) {
    float t = gg::kRayMax;
    RefHit h;
    uint32_t nb = 0, nt = 0;
    trace_ref<false, false>(p.ref, o, d, t, h, nb, nt);
    count_hits(p.ref, o, d, t * 1.000001f, pathImportance, imp, numNodes);
    if (t == gg::kRayMax) break;
    const gg::V3 n = shading_normal(p.ref.vinfo, h.tri, h.bary);
    const MaterialData m = load_material_hit(p.materials, p.tex, p.ref.vinfo, h.tri, h.bary);
    o = o + t * d;
    // NEXT_EVENT_IMPORTANCE (acquisition2.comp:70-99)
    pathImportance *= 0.5f;
    const int sampleIndex = (int)(gg::xorshift32(seed) % p.numLights);
    const gpugi_light_sample ls = p.lights[sampleIndex];
    gg::V3 ld = gg::v3(ls.position[0], ls.position[1], ls.position[2]) - o;
    const float lightDist = sqrtf(gg::dot(ld, ld));
    ld = ld / lightDist;
    const gg::V3 lsd = gg::unpack_normal(ls.normal_phi, ls.normal_theta_cos);
    if (gg::saturate(gg::dot(lsd, -ld)) > 0.0f) {
      float lt = lightDist;
      RefHit lh;
      trace_ref<false, false>(p.ref, o, ld, lt, lh, nb, nt);
      if (lt < lightDist) count_hits(p.ref, o, ld, lt * 1.000001f, pathImportance, imp, numNodes);
    }
    gg::V3 tp = gg::v3s(1.0f);
    float pdf;
    bool diffuse;
    d = sample_bsdf(d, m, seed, n, tp, false, pdf, diffuse);
    o = o + d * gg::kRayHitEpsilon;
  }
}

// The fixed-point sums to the FP32 {sum, 0} pairs the propagation reads (in
// place: each entry is 8 bytes either way).  The conversion is exact while a
// sum stays below 2^24 units.
// One tree level of the subtree sums of count_hits' node terms: an inner node
// adds its children's totals (final, as their level ran before).  The
// children follow the rule of BuildHierTree, which also checked the tree.
__global__ __launch_bounds__(kHierBlock) void k_hier_subtree_sum(NodeView nv, unsigned long long* __restrict__ acc,
                                                                 const uint32_t* __restrict__ order, uint32_t count,
                                                                 uint32_t numNodes) {
  const uint32_t i = blockIdx.x * kHierBlock + threadIdx.x;
  if (i >= count) return;
  const int n = (int)order[i];
  const uint32_t cc = __float_as_uint(nv.t0(n).w);
  if (cc & 0x80000000u) return;  // a leaf node: no child nodes
  const int escape = __float_as_int(nv.t1(n).w);
  const int end = escape == 0 ? (int)numNodes : escape;
  unsigned long long sum = 0;
  int k = (int)cc;
  for (int it = 0; it < kMaxChildren && k > n && k < end; ++it) {
    sum += acc[k];
    const int next = __float_as_int(nv.t1(k).w);
    if (next == 0) break;
    k = next;
  }
  acc[n] += sum;
}

__global__ __launch_bounds__(kHierBlock) void k_hier_to_float(unsigned long long* __restrict__ acc, size_t n) {
  const size_t i = (size_t)blockIdx.x * kHierBlock + threadIdx.x;
  if (i >= n) return;
  const float x = (float)acc[i] * (1.0f / (float)(1 << kImpFracBits));
  reinterpret_cast<float2*>(acc)[i] = make_float2(x, 0.0f);
}

// avgArea of hierarchypropagation_*.comp: a third of the box's surface
__device__ __forceinline__ float avg_area(const NodeView& nv, int n) {
  const float4 a = nv.t0(n), b = nv.t1(n);
  if (nv.obox) return (b.x * b.y + b.x * b.z + b.y * b.z) * 8.0f / 6.0f;  // half side lengths
  const float sx = b.x - a.x, sy = b.y - a.y, sz = b.z - a.z;
  return (sx * sy + sx * sz + sy * sz) * 2.0f / 6.0f;
}

__global__ __launch_bounds__(kHierBlock) void k_hier_init(NodeView nv, float2* __restrict__ imp,
                                                          float* __restrict__ subtree, uint32_t numNodes,
                                                          float iterations) {
  const uint32_t i = blockIdx.x * kHierBlock + threadIdx.x;
  if (i >= numNodes) return;
  const float x = imp[i].x / iterations;
  float y = 0.0f, sub = 0.0f;
  const uint32_t cc = __float_as_uint(nv.t0((int)i).w);
  if (cc & 0x80000000u) {
    const uint32_t first = 2u * (cc & 0x7fffffffu);
    float children = 0.0f;
    for (uint32_t k = 0; k < 2; ++k) children += imp[numNodes + first + k].x;
    children /= iterations;
    sub = fmaxf(children, 1.19e-07f);  // NotificationBias
    const float approximation = x / children;
    y = approximation * approximation * avg_area(nv, (int)i);
  }
  imp[i] = make_float2(x, y);
  subtree[i] = sub;
}

__global__ __launch_bounds__(kHierBlock) void k_hier_nodes(NodeView nv, float2* __restrict__ imp,
                                                           float* __restrict__ subtree,
                                                           const int32_t* __restrict__ parent,
                                                           const uint32_t* __restrict__ level, uint32_t count) {
  const uint32_t j = blockIdx.x * kHierBlock + threadIdx.x;
  if (j >= count) return;
  const int i = (int)level[j];
  if (subtree[i] != 0.0f) return;  // a leaf node, done by k_hier_init
  int child = (int)(__float_as_uint(nv.t0(i).w) & 0x7fffffffu);
  float children = 0.0f;
  for (;;) {
    children += subtree[child];
    child = __float_as_int(nv.t1(child).w);  // escape: the next sibling
    if (child == 0 || parent[child] != i) break;
  }
  subtree[i] = children;
  const float approximation = imp[i].x / children;
  imp[i].y = approximation * approximation * avg_area(nv, i);
}

// packHalf2x16: two halves rounded to nearest even
__device__ __forceinline__ uint32_t pack2(float a, float b) {
  return (uint32_t)__half_as_ushort(__float2half_rn(a)) | ((uint32_t)__half_as_ushort(__float2half_rn(b)) << 16);
}
__device__ __forceinline__ float lo_of(uint32_t w) { return __half2float(__ushort_as_half((unsigned short)(w & 0xffffu))); }
__device__ __forceinline__ float hi_of(uint32_t w) { return __half2float(__ushort_as_half((unsigned short)(w >> 16))); }

struct MatAcc {
  float refl[4], op[3], dif[3], em[3], f0[3], f1[3], rIdx, area;
};

__device__ __forceinline__ void store_material(uint32_t* __restrict__ out, int i, MatAcc a) {
  for (int q = 0; q < 4; ++q) a.refl[q] /= a.area;
  for (int q = 0; q < 3; ++q) {
    a.op[q] /= a.area;
    a.dif[q] /= a.area;
    a.em[q] /= a.area;
    a.f0[q] /= a.area;
    a.f1[q] /= a.area;
  }
  a.rIdx /= a.area;
  uint4* w = reinterpret_cast<uint4*>(out + 12 * (size_t)i);
  w[0] = make_uint4(pack2(a.refl[0], a.refl[1]), pack2(a.refl[2], a.refl[3]), pack2(a.op[0], a.op[1]),
                    pack2(a.op[2], a.area));
  w[1] = make_uint4(pack2(a.dif[0], a.dif[1]), pack2(a.dif[2], 0.0f), pack2(a.em[0], a.em[1]), pack2(a.em[2], 0.0f));
  w[2] = make_uint4(pack2(a.f0[0], a.f0[1]), pack2(a.f0[2], a.rIdx), pack2(a.f1[0], a.f1[1]), pack2(a.f1[2], 0.0f));
}

__global__ __launch_bounds__(kHierBlock) void k_hier_leaf_materials(DrawParams p, NodeView nv,
                                                                    uint32_t* __restrict__ out, uint32_t numNodes) {
  const uint32_t i = blockIdx.x * kHierBlock + threadIdx.x;
  if (i >= numNodes) return;
  const uint32_t cc = __float_as_uint(nv.t0((int)i).w);
  if (!(cc & 0x80000000u)) return;
  const uint32_t base = 2u * (cc & 0x7fffffffu);
  MatAcc a = {};
  for (uint32_t k = 0; k < 2; ++k) {
    const int4 tri = p.ref.slots[base + k];
    if (tri.x == tri.y) break;  // padding slot
    const gg::V3 p0 = load_pos(p.ref.positions, tri.x);
    const gg::V3 cr = gg::cross(load_pos(p.ref.positions, tri.y) - p0, load_pos(p.ref.positions, tri.z) - p0);
    const float triangleArea = sqrtf(gg::dot(cr, cr)) / 2.0f;
    a.area += triangleArea;
    const float third = 1.0f / 3.0f;
    const float4 va = p.ref.vinfo[tri.x], vb = p.ref.vinfo[tri.y], vc = p.ref.vinfo[tri.z];
    const float2 uv = make_float2((va.z * third + vb.z * third) + vc.z * third, (va.w * third + vb.w * third) + vc.w * third);
    const MaterialData m = load_material_uv(p.materials, p.tex, tri.w, uv);
    for (int q = 0; q < 4; ++q) a.refl[q] += m.refl[q] * triangleArea;
    const float mo[3] = {m.opacity.x, m.opacity.y, m.opacity.z}, md[3] = {m.diffuse.x, m.diffuse.y, m.diffuse.z};
    const float me[3] = {m.emissivity.x, m.emissivity.y, m.emissivity.z};
    const float m0[3] = {m.fresnel0.x, m.fresnel0.y, m.fresnel0.z}, m1[3] = {m.fresnel1.x, m.fresnel1.y, m.fresnel1.z};
    for (int q = 0; q < 3; ++q) {
      a.op[q] += mo[q] * triangleArea;
      a.dif[q] += md[q] * triangleArea;
      a.em[q] += me[q] * triangleArea;
      a.f0[q] += m0[q] * triangleArea;
      a.f1[q] += m1[q] * triangleArea;
    }
    a.rIdx += m.refractionIndexAvg * triangleArea;
  }
  store_material(out, (int)i, a);
}

__global__ __launch_bounds__(kHierBlock) void k_hier_pull(NodeView nv, uint32_t* __restrict__ out,
                                                          const int32_t* __restrict__ parent,
                                                          const uint32_t* __restrict__ level, uint32_t count) {
  const uint32_t j = blockIdx.x * kHierBlock + threadIdx.x;
  if (j >= count) return;
  const int i = (int)level[j];
  const uint32_t cc = __float_as_uint(nv.t0(i).w);
  if (cc & 0x80000000u) return;  // a leaf node
  MatAcc a = {};
  int child = (int)cc;
  for (;;) {
    const uint32_t* w = out + 12 * (size_t)child;
    const float ca = hi_of(w[3]);
    a.area += ca;
    a.refl[0] += lo_of(w[0]) * ca;
    a.refl[1] += hi_of(w[0]) * ca;
    a.refl[2] += lo_of(w[1]) * ca;
    a.refl[3] += hi_of(w[1]) * ca;
    a.op[0] += lo_of(w[2]) * ca;
    a.op[1] += hi_of(w[2]) * ca;
    a.op[2] += lo_of(w[3]) * ca;
    a.dif[0] += lo_of(w[4]) * ca;
    a.dif[1] += hi_of(w[4]) * ca;
    a.dif[2] += lo_of(w[5]) * ca;
    a.em[0] += lo_of(w[6]) * ca;
    a.em[1] += hi_of(w[6]) * ca;
    a.em[2] += lo_of(w[7]) * ca;
    a.f0[0] += lo_of(w[8]) * ca;
    a.f0[1] += hi_of(w[8]) * ca;
    a.f0[2] += lo_of(w[9]) * ca;
    a.rIdx += hi_of(w[9]) * ca;
    a.f1[0] += lo_of(w[10]) * ca;
    a.f1[1] += hi_of(w[10]) * ca;
    a.f1[2] += lo_of(w[11]) * ca;
    child = __float_as_int(nv.t1(child).w);
    if (child == 0 || parent[child] != i) break;
  }
  store_material(out, i, a);
}

dim3 Blocks(size_t n) { return dim3((unsigned)((n + kHierBlock - 1) / kHierBlock)); }

}  // namespace

bool BuildHierTree(const float* hierarchy, uint32_t numNodes, std::vector<int32_t>* parent,
                   std::vector<uint32_t>* levelOrder, std::vector<uint32_t>* levelStart) {
  const int n = (int)numNodes;
  auto bits = [&](int node, int k) {
    uint32_t u;
    std::memcpy(&u, &hierarchy[(size_t)node * 8 + (size_t)k], 4);
    return u;
  };
  parent->assign((size_t)n, -1);
  for (int p = 0; p < n; ++p) {
    const uint32_t cc = bits(p, 3);
    if (cc & 0x80000000u) continue;
    const int end = bits(p, 7) == 0u ? n : (int)bits(p, 7);
    int k = (int)cc;
    while (k > p && k < end) {
      (*parent)[(size_t)k] = p;
      const int next = (int)bits(k, 7);
      if (next == 0) break;
      k = next;
    }
  }
  // preorder: a parent precedes its children
  std::vector<uint32_t> depth((size_t)n, 0u);
  uint32_t maxDepth = 0;
  for (int i = 1; i < n; ++i) {
    const int pi = (*parent)[(size_t)i];
    if (pi < 0 || pi >= i) return false;
    depth[(size_t)i] = depth[(size_t)pi] + 1u;
    maxDepth = std::max(maxDepth, depth[(size_t)i]);
  }
  levelStart->assign((size_t)maxDepth + 2, 0u);
  for (int i = 0; i < n; ++i) ++(*levelStart)[depth[(size_t)i] + 1];
  for (size_t d = 1; d < levelStart->size(); ++d) (*levelStart)[d] += (*levelStart)[d - 1];
  levelOrder->assign((size_t)n, 0u);
  std::vector<uint32_t> fill(levelStart->begin(), levelStart->end() - 1);
  for (int i = 0; i < n; ++i) (*levelOrder)[fill[depth[(size_t)i]]++] = (uint32_t)i;
  return true;
}

hipError_t LaunchHierAcquire(const DrawParams& p, int iteration, unsigned long long* importance, uint32_t numNodes,
                             hipStream_t stream) {
  const uint32_t tilesW = (uint32_t)(p.width / kTile), tilesH = (uint32_t)(p.height / kTile);
  const uint32_t numPixels = tilesW * tilesH * kTileThreads;
  if (numPixels == 0) return hipSuccess;
  if (p.numLights == 0 || !p.lights) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_hier_acquire, Blocks(numPixels), dim3(kHierBlock), 0, stream, p,
                     gg::wanghash((uint32_t)iteration), importance, numNodes, tilesW, numPixels);
  return hipGetLastError();
}

hipError_t LaunchHierSubtreeSum(const RefScene& s, const HierTree& tree, unsigned long long* importance,
                                hipStream_t stream) {
  const NodeView nv{s.nodes, s.obox};
  for (int lv = (int)tree.levelStart.size() - 2; lv >= 0; --lv) {  // deepest level first
    const uint32_t first = tree.levelStart[lv], count = tree.levelStart[lv + 1] - first;
    if (count == 0) continue;
    hipLaunchKernelGGL(k_hier_subtree_sum, Blocks(count), dim3(kHierBlock), 0, stream, nv, importance,
                       tree.levelOrder + first, count, tree.numNodes);
  }
  return hipGetLastError();
}

hipError_t LaunchHierToFloat(unsigned long long* importance, size_t entries, hipStream_t stream) {
  if (entries == 0) return hipSuccess;
  hipLaunchKernelGGL(k_hier_to_float, Blocks(entries), dim3(kHierBlock), 0, stream, importance, entries);
  return hipGetLastError();
}

hipError_t LaunchHierPropagate(const RefScene& s, const HierTree& tree, int iterations, float2* importance,
                               float* subtree, hipStream_t stream) {
  const NodeView nv{s.nodes, s.obox};
  hipLaunchKernelGGL(k_hier_init, Blocks(tree.numNodes), dim3(kHierBlock), 0, stream, nv, importance, subtree,
                     tree.numNodes, (float)iterations);
  for (int lv = (int)tree.levelStart.size() - 2; lv >= 0; --lv) {
    const uint32_t first = tree.levelStart[lv], count = tree.levelStart[lv + 1] - first;
    if (count == 0) continue;
    hipLaunchKernelGGL(k_hier_nodes, Blocks(count), dim3(kHierBlock), 0, stream, nv, importance, subtree,
                       tree.parent, tree.levelOrder + first, count);
  }
  return hipGetLastError();
}

hipError_t LaunchHierMaterials(const DrawParams& p, const HierTree& tree, uint32_t* materials, hipStream_t stream) {
  const NodeView nv{p.ref.nodes, p.ref.obox};
  hipLaunchKernelGGL(k_hier_leaf_materials, Blocks(tree.numNodes), dim3(kHierBlock), 0, stream, p, nv, materials,
                     tree.numNodes);
  for (int lv = (int)tree.levelStart.size() - 2; lv >= 0; --lv) {
    const uint32_t first = tree.levelStart[lv], count = tree.levelStart[lv + 1] - first;
    if (count == 0) continue;
    hipLaunchKernelGGL(k_hier_pull, Blocks(count), dim3(kHierBlock), 0, stream, nv, materials, tree.parent,
                       tree.levelOrder + first, count);
  }
  return hipGetLastError();
}

}  // namespace gpugi
