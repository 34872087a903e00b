// Re-pack of the reference BVH2 into the 4-wide quantized layout of wide_bvh.h.
#include "../device/wide_bvh.h"

#include <algorithm>
#include <array>
#include <limits>
#include <cmath>
#include <cstring>

namespace gpugi {
namespace {

struct Bvh2View {
  const gpugi_scene_desc* d;
  uint32_t n;
  const float* box(uint32_t i) const { return d->hierarchy + (size_t)i * 8; }
  uint32_t firstChild(uint32_t i) const { uint32_t u; std::memcpy(&u, box(i) + 3, 4); return u; }
  uint32_t escape(uint32_t i) const { uint32_t u; std::memcpy(&u, box(i) + 7, 4); return u; }
  bool isLeafNode(uint32_t i) const { return (firstChild(i) & 0x80000000u) != 0; }
  float surface(uint32_t i) const {
    const float* b = box(i);
    float x = b[4] - b[0], y = b[5] - b[1], z = b[6] - b[2];
    return 2.0f * (x * y + y * z + z * x);
  }
};

// Children of an inner node: firstChild, then escapes until the chain leaves
// the parent (the last child escapes to the parent's escape).
bool Children(const Bvh2View& v, uint32_t i, std::vector<uint32_t>* out, std::string* err) {
  out->clear();
  uint32_t c = v.firstChild(i);
  const uint32_t stop = v.escape(i);
  for (int guard = 0; guard < 64; ++guard) {
    if (c <= i || c >= v.n) { *err = "malformed hierarchy below node " + std::to_string(i); return false; }
    out->push_back(c);
    uint32_t e = v.escape(c);
    if (e == stop) return true;
    if (e <= c) { *err = "malformed escape chain below node " + std::to_string(i); return false; }
    c = e;
  }
  *err = "node " + std::to_string(i) + " has more than 64 children";
  return false;
}

// The exact real origin + q * scale (extended precision: the kernel never rounds a
// decoded bound, it folds q * scale into the slab distance).
long double Decode(float origin, uint32_t q, float scale) {
  return (long double)origin + (long double)q * (long double)scale;
}
float ScaleOf(uint32_t biased) { uint32_t u = biased << 23; float f; std::memcpy(&f, &u, 4); return f; }

}  // namespace

std::string BuildWideBVH(const gpugi_scene_desc& desc, uint32_t maxLeafTris, WideBVHHost* out, bool rawVertices) {
  if (maxLeafTris < 1 || maxLeafTris > kWideMaxLeafTris) return "maxLeafTris must be in [1, 16]";
  Bvh2View v{&desc, desc.num_nodes};
  if (v.n == 0) return "empty hierarchy";
  const uint32_t tpl = desc.triangles_per_leaf;
  std::string err;
  // Subtree valid-triangle counts, bottom-up (children follow parents in preorder).
  std::vector<uint32_t> count(v.n, 0);
  std::vector<std::vector<uint32_t>> kids(v.n);
  for (uint32_t i = 0; i < v.n; ++i) {
    if (v.isLeafNode(i)) continue;
    if (!Children(v, i, &kids[i], &err)) return err;
  }
  for (uint32_t ii = v.n; ii-- > 0;) {
    if (v.isLeafNode(ii)) {
      uint32_t leaf = v.firstChild(ii) & 0x7fffffffu;
      if (leaf >= desc.num_leaves) return "leaf index out of range";
      uint32_t c = 0;
      for (uint32_t k = 0; k < tpl; ++k) {
        const int32_t* t = desc.triangles + ((size_t)leaf * tpl + k) * 4;
        if (t[0] == t[1]) break;
        ++c;
      }
      count[ii] = c;
    } else {
      uint32_t c = 0;
      for (uint32_t k : kids[ii]) c += count[k];
      count[ii] = c;
    }
  }
  out->nodes.clear();
  out->tris.clear();
  out->tris.reserve(count[0]);

  auto appendTris = [&](uint32_t sub) -> uint32_t {
    // All valid slots of a BVH2 subtree in slot order of a preorder walk.
    uint32_t first = (uint32_t)out->tris.size();
    std::vector<uint32_t> st{sub};
    while (!st.empty()) {
      uint32_t i = st.back();
      st.pop_back();
      if (v.isLeafNode(i)) {
        uint32_t leaf = v.firstChild(i) & 0x7fffffffu;
        for (uint32_t k = 0; k < tpl; ++k) {
          size_t slot = (size_t)leaf * tpl + k;
          const int32_t* t = desc.triangles + slot * 4;
          if (t[0] == t[1]) break;
          const float* p0 = desc.positions + (size_t)t[0] * 3;
          const float* p1 = desc.positions + (size_t)t[1] * 3;
          const float* p2 = desc.positions + (size_t)t[2] * 3;
          WideTriHost w;
          for (int a = 0; a < 3; ++a) {
            w.p0[a] = p0[a];
            // reference test operands e0 = p1 - p0, e1 = p0 - p2, or the raw vertices
            w.e0[a] = rawVertices ? p1[a] : p1[a] - p0[a];
            w.e1[a] = rawVertices ? p2[a] : p0[a] - p2[a];
          }
          w.slot = (int32_t)slot;
          w.material = t[3];
          w.pad = 0;
          out->tris.push_back(w);
        }
      } else {
        for (size_t k = kids[i].size(); k-- > 0;) st.push_back(kids[i][k]);
      }
    }
    return first;
  };

  // SAH-optimal collapse of the BVH2 into 4-wide nodes (the dynamic program of
  // Ylitie, Karras and Laine 2017, "Efficient incoherent ray traversal on
  // GPUs through compressed wide BVHs").  dist[n][j] is the least expected
  // cost of covering n's subtree with at most j children of one wide node;
  // a child is a leaf (a reference leaf, or a subtree of at most maxLeafTris
  // valid triangles) or a wide node of its own.  A wide node costs
  // kNodeCost * area per visit, a leaf kTriCost * area * triangles; the ratio
  // is the measured cost of a box pass (four slab tests, ordering, stack) to
  // that of one triangle test.  The topology stays the reference's: only which
  // BVH2 nodes become wide nodes or leaves is chosen.
  constexpr float kNodeCost = 1.0f, kTriCost = 0.4f;
  constexpr float kInf = std::numeric_limits<float>::infinity();
  struct Dp {
    float dist[5];
    uint8_t pick[5];     // 0: no child (empty subtree), 1: n as one child, 2: n's children share j
    bool leaf;           // n as one child: a leaf (else a wide node)
  };
  std::vector<Dp> dp(v.n);
  // split[splitAt[i] + k][j]: children given to kid k when kids 0..k of node i
  // share j (knapsack back-pointers)
  std::vector<size_t> splitAt(v.n + 1, 0);
  for (uint32_t i = 0; i < v.n; ++i) splitAt[i + 1] = splitAt[i] + kids[i].size();
  std::vector<std::array<uint8_t, 5>> split(splitAt[v.n], std::array<uint8_t, 5>{0, 0, 0, 0, 0});
  for (uint32_t n = v.n; n-- > 0;) {
    Dp& d = dp[n];
    if (count[n] == 0) {
      for (int j = 0; j < 5; ++j) { d.dist[j] = 0.0f; d.pick[j] = 0; }
      d.leaf = true;
      continue;
    }
    const float area = v.surface(n);
    const bool leafOk = v.isLeafNode(n) || count[n] <= maxLeafTris;
    const float leafCost = leafOk ? kTriCost * area * (float)count[n] : kInf;
    std::array<float, 5> comb = {0.0f, kInf, kInf, kInf, kInf};
    if (!v.isLeafNode(n)) {
      for (size_t k = 0; k < kids[n].size(); ++k) {
        const Dp& c = dp[kids[n][k]];
        std::array<float, 5> next = {kInf, kInf, kInf, kInf, kInf};
        for (int j = 0; j < 5; ++j)
          for (int a = 0; a <= j; ++a) {
            const float cost = comb[j - a] + c.dist[a];
            if (cost < next[j]) { next[j] = cost; split[splitAt[n] + k][j] = (uint8_t)a; }
          }
        comb = next;
      }
    } else {
      comb = {kInf, kInf, kInf, kInf, kInf};
    }
    const float nodeCost = v.isLeafNode(n) ? kInf : kNodeCost * area + comb[4];
    const float single = std::min(leafCost, nodeCost);
    d.leaf = leafCost <= nodeCost;
    d.dist[0] = kInf;
    d.pick[0] = 0;
    for (int j = 1; j < 5; ++j) {
      if (single <= comb[j]) { d.dist[j] = single; d.pick[j] = 1; }
      else { d.dist[j] = comb[j]; d.pick[j] = 2; }
    }
  }
  // The children of the wide node for BVH2 node n: n's own children sharing 4.
  struct Item {
    uint32_t bvh2;
    bool leaf;
  };
  std::vector<Item> items;
  std::vector<std::pair<uint32_t, int>> work;  // (node, children it may use)
  auto wideChildren = [&](uint32_t n) {
    items.clear();
    if (v.isLeafNode(n)) {  // a one-leaf scene
      if (count[n] > 0) items.push_back({n, true});
      return;
    }
    // knapsack back-tracking over n's kids, then each kid's own choice
    work.clear();
    int j = 4;
    std::vector<std::pair<uint32_t, int>> alloc;
    for (size_t k = kids[n].size(); k-- > 0;) {
      const int a = split[splitAt[n] + k][j];
      alloc.push_back({kids[n][k], a});
      j -= a;
    }
    for (size_t k = alloc.size(); k-- > 0;) work.push_back(alloc[k]);
    // expand in child order
    for (size_t w = 0; w < work.size(); ++w) {
      const uint32_t c = work[w].first;
      const int jc = work[w].second;
      const Dp& d = dp[c];
      if (jc == 0 || d.pick[jc] == 0) continue;
      if (d.pick[jc] == 1) {
        items.push_back({c, d.leaf});
        continue;
      }
      // c opened: its kids share jc, inserted in place to keep the reference's child order
      std::vector<std::pair<uint32_t, int>> sub;
      int jj = jc;
      for (size_t k = kids[c].size(); k-- > 0;) {
        const int a = split[splitAt[c] + k][jj];
        sub.push_back({kids[c][k], a});
        jj -= a;
      }
      std::reverse(sub.begin(), sub.end());
      work.insert(work.begin() + (std::ptrdiff_t)w + 1, sub.begin(), sub.end());
    }
  };

  struct Pending {
    uint32_t bvh2;       // BVH2 node whose subtree this wide node covers
    uint32_t wideIndex;  // index of the wide node to fill
  };
  std::vector<Pending> todo;
  out->nodes.push_back(WideNodeHost());
  todo.push_back({0, 0});
  std::vector<uint32_t> needStack;   // per wide node: worst-case stack entries below it
  std::vector<std::vector<uint32_t>> wideKids;
  needStack.push_back(0);
  wideKids.emplace_back();
  while (!todo.empty()) {
    Pending p = todo.back();
    todo.pop_back();
    // Subtrees without a valid triangle never produce a hit and are left out.
    wideChildren(p.bvh2);
    if (items.size() > 4) return "BVH node with more than 4 children cannot be packed";
    if (items.empty()) {  // a scene without any valid triangle
      WideNodeHost empty;
      std::memset(&empty, 0, sizeof(empty));
      empty.exps = 1u | (1u << 8) | (1u << 16);
      out->nodes[p.wideIndex] = empty;
      continue;
    }
    WideNodeHost node;
    std::memset(&node, 0, sizeof(node));
    // Quantization frame.
    float origin[3], top[3];
    for (int a = 0; a < 3; ++a) {
      origin[a] = v.box(items[0].bvh2)[a];
      top[a] = v.box(items[0].bvh2)[4 + a];
      for (const Item& it : items) {
        origin[a] = std::min(origin[a], v.box(it.bvh2)[a]);
        top[a] = std::max(top[a], v.box(it.bvh2)[4 + a]);
      }
      node.origin[a] = origin[a];
    }
    uint32_t biased[3];
    for (int a = 0; a < 3; ++a) {
      double ext = (double)top[a] - (double)origin[a];
      int b = 1;
      if (ext > 0.0) b = std::max(1, std::min(254, (int)std::ceil(std::log2(ext / 255.0)) + 127));
      for (;; ++b) {
        if (b > 254) return "cannot quantize a node box";
        const float scale = ScaleOf((uint32_t)b);
        bool ok = true;
        for (size_t k = 0; k < items.size() && ok; ++k) {
          const float lo = v.box(items[k].bvh2)[a], hi = v.box(items[k].bvh2)[4 + a];
          double ql = std::floor(((double)lo - (double)origin[a]) / (double)scale);
          double qh = std::ceil(((double)hi - (double)origin[a]) / (double)scale);
          if (ql < 0) ql = 0;
          if (qh > 255.0) { ok = false; break; }
          uint32_t qlo = (uint32_t)ql, qhi = (uint32_t)std::max(0.0, qh);
          while (qlo > 0 && Decode(origin[a], qlo, scale) > lo) --qlo;
          while (qhi <= 255 && Decode(origin[a], qhi, scale) < hi) ++qhi;
          if (qhi > 255 || Decode(origin[a], qlo, scale) > lo) { ok = false; break; }
          node.lo[a][k] = (uint8_t)qlo;
          node.hi[a][k] = (uint8_t)qhi;
        }
        if (ok) break;
      }
      biased[a] = (uint32_t)b;
    }
    node.exps = biased[0] | (biased[1] << 8) | (biased[2] << 16) | ((uint32_t)items.size() << 24);
    uint32_t innerKids = 0;
    for (size_t k = 0; k < items.size(); ++k) {
      const uint32_t it = items[k].bvh2;
      if (items[k].leaf) {
        uint32_t c = count[it];
        uint32_t first = appendTris(it);
        if (first > kWideLeafFirstMask) return "too many triangles for the wide layout";
        node.child[k] = kWideLeafFlag | ((c - 1) << 27) | first;
      } else {
        uint32_t wi = (uint32_t)out->nodes.size();
        out->nodes.push_back(WideNodeHost());
        needStack.push_back(0);
        wideKids.emplace_back();
        wideKids[p.wideIndex].push_back(wi);
        node.child[k] = wi;
        todo.push_back({it, wi});
        ++innerKids;
      }
    }
    node.pad[0] = innerKids;  // used for the stack bound only
    out->nodes[p.wideIndex] = node;
  }
  // Worst-case stack.  A visit pushes every hit child except the nearest inner
  // one, which is visited next (leaf children are pushed when the lane's
  // pending-leaf slot is taken), so need(n) = (children - [n has an inner
  // child]) + max need(inner child); popped leaf entries push nothing.  Wide
  // nodes are numbered parent-before-child, so a reverse sweep is bottom-up.
  uint32_t depthMax = 1;
  std::vector<uint32_t> depth(out->nodes.size(), 1);
  for (size_t i = 0; i < out->nodes.size(); ++i)
    for (uint32_t c : wideKids[i]) { depth[c] = depth[i] + 1; depthMax = std::max(depthMax, depth[c]); }
  for (size_t i = out->nodes.size(); i-- > 0;) {
    uint32_t m = 0;
    for (uint32_t c : wideKids[i]) m = std::max(m, needStack[c]);
    const uint32_t k = out->nodes[i].pad[0];
    const uint32_t children = out->nodes[i].exps >> 24;
    needStack[i] = (k > 0 ? children - 1 : children) + m;
    out->nodes[i].pad[0] = 0;
  }
  out->maxStack = needStack[0];
  out->depth = depthMax;
  return std::string();
}

}  // namespace gpugi
