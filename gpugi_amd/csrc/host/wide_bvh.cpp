// Re-pack of the reference BVH2 into the 4-wide quantized layout of wide_bvh.h.
#include "../device/wide_bvh.h"

#include <algorithm>
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
    std::vector<uint32_t> items;
    if (v.isLeafNode(p.bvh2) || count[p.bvh2] <= maxLeafTris) items.push_back(p.bvh2);
    else items = kids[p.bvh2];
    // Open the largest-surface inner item while the node has room.
    for (;;) {
      int best = -1;
      float bestSA = -1.0f;
      for (size_t k = 0; k < items.size(); ++k) {
        uint32_t it = items[k];
        if (v.isLeafNode(it) || count[it] <= maxLeafTris) continue;
        if (items.size() - 1 + kids[it].size() > 4) continue;
        float sa = v.surface(it);
        if (sa > bestSA) { bestSA = sa; best = (int)k; }
      }
      if (best < 0) break;
      uint32_t it = items[(size_t)best];
      items.erase(items.begin() + best);
      items.insert(items.begin() + best, kids[it].begin(), kids[it].end());
    }
    // Subtrees without a valid triangle can never produce a hit: drop them.
    items.erase(std::remove_if(items.begin(), items.end(), [&](uint32_t it) { return count[it] == 0; }),
                items.end());
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
      origin[a] = v.box(items[0])[a];
      top[a] = v.box(items[0])[4 + a];
      for (uint32_t it : items) {
        origin[a] = std::min(origin[a], v.box(it)[a]);
        top[a] = std::max(top[a], v.box(it)[4 + a]);
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
          const float lo = v.box(items[k])[a], hi = v.box(items[k])[4 + a];
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
      uint32_t it = items[k];
      if (v.isLeafNode(it) || count[it] <= maxLeafTris) {
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
