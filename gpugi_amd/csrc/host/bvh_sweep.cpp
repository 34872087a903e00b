// SAH sweep BVH builder.
//
// Restates the algorithm of bvhmake's BuildSweep (bvhmake/buildmethods/sweep.cpp):
//  1. triangles are ordered once along a Hilbert curve of their centroids
//     (sweep.cpp:143-183: centroid - min, x4096, 16-bit chunks interleaved and
//     Gray-decoded, high chunk first);
//  2. a node over a contiguous range of that order becomes a leaf when it holds
//     at most `trianglesPerLeaf` triangles (sweep.cpp:195-211); otherwise every
//     split position is scored with SA(side)/SA(node) * count for both sides
//     (sweep.cpp:14-23, 217-239) and the cheapest split with at least two
//     triangles on the left wins, first minimum on ties (sweep.cpp:241-251);
//  3. nodes and leaves are numbered in preorder, the left child directly
//     follows its parent, the left child escapes to its right sibling and the
//     right child to the parent's escape (bvhbuilder.cpp:710-722).
// Differences, all without effect on the tree: ties of the Hilbert key are
// broken by triangle index (std::sort in the reference leaves them open), the
// recursion uses an explicit stack, and prefix/suffix boxes live in their own
// scratch arrays instead of the tail of the node box buffer (sweep.cpp:214-216).
#include <algorithm>
#include <cmath>
#include <cstring>

#include "scene.hpp"

namespace gpugi {
namespace {

struct Box3 {
  float mn[3], mx[3];
};

inline Box3 Merge(const Box3& a, const Box3& b) {
  Box3 r;
  for (int i = 0; i < 3; ++i) {
    r.mn[i] = std::min(a.mn[i], b.mn[i]);
    r.mx[i] = std::max(a.mx[i], b.mx[i]);
  }
  return r;
}

inline float Surface(const Box3& b) {
  float x = b.mx[0] - b.mn[0], y = b.mx[1] - b.mn[1], z = b.mx[2] - b.mn[2];
  return 2.0f * (x * y + y * z + z * x);
}

// Spread the low 16 bits of x so that bit i lands at bit 3i (3D Morton code).
inline uint64_t Spread3(uint64_t x) {
  x &= 0xffffull;
  x = (x | (x << 16)) & 0x0000ff0000ffull;
  x = (x | (x << 8)) & 0x00f00f00f00full;
  x = (x | (x << 4)) & 0x0c30c30c30c3ull;
  x = (x | (x << 2)) & 0x249249249249ull;
  return x;
}
inline uint64_t Morton3(uint32_t a, uint32_t b, uint32_t c) {
  return Spread3(a) | (Spread3(b) << 1) | (Spread3(c) << 2);
}
// Inverse of the binary-reflected Gray code: prefix XOR from the top bit down.
inline uint64_t GrayDecode(uint64_t g) {
  g ^= g >> 32;
  g ^= g >> 16;
  g ^= g >> 8;
  g ^= g >> 4;
  g ^= g >> 2;
  g ^= g >> 1;
  return g;
}

struct Task {
  uint32_t lo, hi;
  int64_t parent;  // -1 for the root
  bool isRight;
};

}  // namespace

std::string BuildSweepBVH(const std::vector<RawVertex>& vertices,
                          const std::vector<RawTriangle>& triangles, uint32_t trianglesPerLeaf,
                          HostScene* out) {
  if (trianglesPerLeaf == 0 || trianglesPerLeaf > 64) return "trianglesPerLeaf must be in [1, 64]";
  // Drop triangles that cannot be stored (repeated vertex index, bvhmake/filedef.hpp:42-48).
  std::vector<uint32_t> valid;
  valid.reserve(triangles.size());
  for (uint32_t i = 0; i < triangles.size(); ++i) {
    const RawTriangle& t = triangles[i];
    if (t.vertices[0] == t.vertices[1] || t.vertices[1] == t.vertices[2] || t.vertices[0] == t.vertices[2])
      continue;
    for (int j = 0; j < 3; ++j)
      if (t.vertices[j] >= vertices.size()) return "triangle references a vertex out of range";
    valid.push_back(i);
  }
  const uint32_t n = (uint32_t)valid.size();
  if (n == 0) return "no valid triangle to build a hierarchy from";

  // Per-triangle boxes and centroids.
  std::vector<Box3> triBox(n);
  std::vector<float> centers((size_t)n * 3);
  float minC[3];
  {
    const float* p0 = vertices[triangles[valid[0]].vertices[0]].position;
    for (int j = 0; j < 3; ++j) minC[j] = p0[j];  // sweep.cpp:166 seeds the minimum with v0 of triangle 0
  }
  for (uint32_t i = 0; i < n; ++i) {
    const RawTriangle& t = triangles[valid[i]];
    const float* a = vertices[t.vertices[0]].position;
    const float* b = vertices[t.vertices[1]].position;
    const float* c = vertices[t.vertices[2]].position;
    for (int j = 0; j < 3; ++j) {
      triBox[i].mn[j] = std::min(std::min(a[j], b[j]), c[j]);
      triBox[i].mx[j] = std::max(std::max(a[j], b[j]), c[j]);
      float cj = (a[j] + b[j] + c[j]) / 3.0f;
      centers[(size_t)i * 3 + j] = cj;
      minC[j] = std::min(minC[j], cj);
    }
  }
  // Hilbert keys: (high 16-bit chunk, low 16-bit chunk).
  std::vector<uint64_t> keyHi(n), keyLo(n);
  for (uint32_t i = 0; i < n; ++i) {
    uint32_t q[3];
    for (int j = 0; j < 3; ++j) q[j] = (uint32_t)((centers[(size_t)i * 3 + j] - minC[j]) * 4096.0f);
    keyHi[i] = GrayDecode(Morton3(q[0] >> 16, q[1] >> 16, q[2] >> 16));
    keyLo[i] = GrayDecode(Morton3(q[0] & 0xffff, q[1] & 0xffff, q[2] & 0xffff));
  }
  std::vector<uint32_t> order(n);
  for (uint32_t i = 0; i < n; ++i) order[i] = i;
  std::sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) {
    if (keyHi[a] != keyHi[b]) return keyHi[a] < keyHi[b];
    if (keyLo[a] != keyLo[b]) return keyLo[a] < keyLo[b];
    return a < b;
  });

  std::vector<Box3> sorted(n);
  for (uint32_t i = 0; i < n; ++i) sorted[i] = triBox[order[i]];
  std::vector<Box3> suffix(n);
  std::vector<float> costL(n);

  struct BuildNode {
    uint32_t left, right;
    Box3 box;
  };
  std::vector<BuildNode> nodes;
  nodes.reserve(2 * (size_t)n / std::max<uint32_t>(1, trianglesPerLeaf) + 2);
  std::vector<RawTriangle> slots;
  slots.reserve(((size_t)n / trianglesPerLeaf + 1) * trianglesPerLeaf * 2);
  const RawTriangle kInvalid = {{0, 0, 0}, 0};

  std::vector<Task> stack;
  stack.push_back({0, n - 1, -1, false});
  while (!stack.empty()) {
    Task task = stack.back();
    stack.pop_back();
    const uint32_t idx = (uint32_t)nodes.size();
    nodes.push_back(BuildNode());
    if (task.parent >= 0) {
      if (task.isRight) nodes[(size_t)task.parent].right = idx;
      else nodes[(size_t)task.parent].left = idx;
    }
    const uint32_t lo = task.lo, hi = task.hi;
    Box3 box = sorted[lo];
    for (uint32_t i = lo + 1; i <= hi; ++i) box = Merge(box, sorted[i]);
    nodes[idx].box = box;
    if (hi - lo < trianglesPerLeaf) {
      const uint32_t leaf = (uint32_t)(slots.size() / trianglesPerLeaf);
      for (uint32_t i = lo; i <= hi; ++i) {
        const RawTriangle& t = triangles[valid[order[i]]];
        slots.push_back(t);
      }
      for (uint32_t i = 0; i < trianglesPerLeaf - (hi - lo + 1); ++i) slots.push_back(kInvalid);
      nodes[idx].left = 0x80000000u | leaf;
      nodes[idx].right = 0;
      continue;
    }
    const uint32_t cnt = hi - lo + 1;
    const float parentSA = Surface(box);
    // Left costs for split k (left = [lo, lo+k]).
    Box3 acc = sorted[lo];
    for (uint32_t k = 0; k + 1 < cnt; ++k) {
      if (k > 0) acc = Merge(acc, sorted[lo + k]);
      costL[k] = Surface(acc) / parentSA * (float)(k + 1);
    }
    // Right boxes: suffix[j] = union of [j, hi].
    suffix[hi] = sorted[hi];
    for (uint32_t j = hi; j > lo + 1; --j) suffix[j - 1] = Merge(suffix[j], sorted[j - 1]);
    float minCost = INFINITY;
    uint32_t split = 0;
    for (uint32_t k = 1; k + 1 < cnt; ++k) {
      float costR = Surface(suffix[lo + k + 1]) / parentSA * (float)(cnt - 1 - k);
      float c = costL[k] + costR;
      if (c < minCost) { minCost = c; split = k; }
    }
    // Push right first so the left subtree is numbered next (preorder).
    stack.push_back({lo + split + 1, hi, (int64_t)idx, true});
    stack.push_back({lo, lo + split, (int64_t)idx, false});
  }

  // Escape pointers and parents (RecursiveWriteHierarchy, bvhbuilder.cpp:710-722).
  const uint32_t nn = (uint32_t)nodes.size();
  out->nodes.assign(nn, RawNode{0, 0, 0});
  out->boxes.resize(nn);
  for (uint32_t i = 0; i < nn; ++i) {
    out->nodes[i].firstChild = nodes[i].left;
    for (int j = 0; j < 3; ++j) {
      out->boxes[i].min[j] = nodes[i].box.mn[j];
      out->boxes[i].max[j] = nodes[i].box.mx[j];
    }
  }
  for (uint32_t i = 0; i < nn; ++i) {
    if (nodes[i].left & 0x80000000u) continue;
    const uint32_t l = nodes[i].left, r = nodes[i].right;
    out->nodes[l].parent = i;
    out->nodes[r].parent = i;
    out->nodes[l].escape = r;
    out->nodes[r].escape = out->nodes[i].escape;  // parent precedes children, so it is final
  }
  out->leafSlots.swap(slots);
  out->trianglesPerLeaf = trianglesPerLeaf;
  out->vertices = vertices;
  return std::string();
}

}  // namespace gpugi
