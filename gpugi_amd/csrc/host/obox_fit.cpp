// Oriented-box hierarchy for the reference walk's OBOX_BVH mode
// (TreeNode<OBox>, scene.hpp:39-49; FetchIntersectOBoxNode,
// intersectiontests.glsl:114-128).  The reference reads the boxes from its
// scene library, which is not in the tree; here every node of the existing
// hierarchy gets its own box: the principal axes of its subtree's vertices
// (eigenvectors of their covariance) or the world axes, whichever box has the
// smaller surface area (the cost a ray pays to test it), made to enclose the
// vertices exactly under the float quaternion the walk rotates with, then
// widened by a small margin for the walk's float arithmetic.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <new>
#include <thread>
#include <vector>

#include "scene.hpp"

namespace gpugi {

namespace {

struct Q {
  double x, y, z, w;
};

// rotate(x, q) of intersectiontests.glsl:82-88, in double.
void Rotate(const Q& q, const double* v, double* out) {
  const double t[3] = {q.y * v[2] - q.z * v[1], q.z * v[0] - q.x * v[2], q.x * v[1] - q.y * v[0]};
  out[0] = v[0] + 2.0 * (q.w * t[0] + q.y * t[2] - q.z * t[1]);
  out[1] = v[1] + 2.0 * (q.w * t[1] + q.z * t[0] - q.x * t[2]);
  out[2] = v[2] + 2.0 * (q.w * t[2] + q.x * t[1] - q.y * t[0]);
}

// Unit quaternion q with rotate(x, q) = R x for a rotation matrix R (rows r[0..2]).
Q FromMatrix(const double r[3][3]) {
  Q q;
  const double tr = r[0][0] + r[1][1] + r[2][2];
  if (tr > 0.0) {
    const double s = std::sqrt(tr + 1.0) * 2.0;
    q.w = 0.25 * s;
    q.x = (r[2][1] - r[1][2]) / s;
    q.y = (r[0][2] - r[2][0]) / s;
    q.z = (r[1][0] - r[0][1]) / s;
  } else if (r[0][0] > r[1][1] && r[0][0] > r[2][2]) {
    const double s = std::sqrt(1.0 + r[0][0] - r[1][1] - r[2][2]) * 2.0;
    q.w = (r[2][1] - r[1][2]) / s;
    q.x = 0.25 * s;
    q.y = (r[0][1] + r[1][0]) / s;
    q.z = (r[0][2] + r[2][0]) / s;
  } else if (r[1][1] > r[2][2]) {
    const double s = std::sqrt(1.0 + r[1][1] - r[0][0] - r[2][2]) * 2.0;
    q.w = (r[0][2] - r[2][0]) / s;
    q.x = (r[0][1] + r[1][0]) / s;
    q.y = 0.25 * s;
    q.z = (r[1][2] + r[2][1]) / s;
  } else {
    const double s = std::sqrt(1.0 + r[2][2] - r[0][0] - r[1][1]) * 2.0;
    q.w = (r[1][0] - r[0][1]) / s;
    q.x = (r[0][2] + r[2][0]) / s;
    q.y = (r[1][2] + r[2][1]) / s;
    q.z = 0.25 * s;
  }
  const double n = std::sqrt(q.x * q.x + q.y * q.y + q.z * q.z + q.w * q.w);
  return Q{q.x / n, q.y / n, q.z / n, q.w / n};
}

// Eigenvectors of a symmetric 3x3 matrix (cyclic Jacobi), as the rows of v.
void Eigenvectors(double a[3][3], double v[3][3]) {
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) v[i][j] = i == j ? 1.0 : 0.0;
  for (int sweep = 0; sweep < 32; ++sweep) {
    const double off = a[0][1] * a[0][1] + a[0][2] * a[0][2] + a[1][2] * a[1][2];
    if (off < 1e-30) break;
    for (int p = 0; p < 2; ++p)
      for (int q = p + 1; q < 3; ++q) {
        if (std::fabs(a[p][q]) < 1e-300) continue;
        const double theta = (a[q][q] - a[p][p]) / (2.0 * a[p][q]);
        const double t = (theta >= 0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
        const double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
        for (int k = 0; k < 3; ++k) {  // A <- J^T A J
          const double akp = a[k][p], akq = a[k][q];
          a[k][p] = c * akp - s * akq;
          a[k][q] = s * akp + c * akq;
        }
        for (int k = 0; k < 3; ++k) {
          const double apk = a[p][k], aqk = a[q][k];
          a[p][k] = c * apk - s * aqk;
          a[q][k] = s * apk + c * aqk;
        }
        for (int k = 0; k < 3; ++k) {  // rows of v are the eigenvectors
          const double vpk = v[p][k], vqk = v[q][k];
          v[p][k] = c * vpk - s * vqk;
          v[q][k] = s * vpk + c * vqk;
        }
      }
  }
}

struct Box {
  float q[4];       // x, y, z, w (float, as stored)
  float center[3];  // float, as stored
  double half[3];   // exact half extents of the vertices about the float centre
  double area;
};

// The box of the points for the float quaternion of orientation qd.
Box BoxFor(const Q& qd, const std::vector<double>& pts) {
  Box b;
  b.q[0] = (float)qd.x;
  b.q[1] = (float)qd.y;
  b.q[2] = (float)qd.z;
  b.q[3] = (float)qd.w;
  const Q q{b.q[0], b.q[1], b.q[2], b.q[3]};
  const Q qi{-q.x, -q.y, -q.z, q.w};
  double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  double y[3];
  for (size_t i = 0; i < pts.size(); i += 3) {
    Rotate(q, &pts[i], y);
    for (int k = 0; k < 3; ++k) {
      lo[k] = std::min(lo[k], y[k]);
      hi[k] = std::max(hi[k], y[k]);
    }
  }
  // the float quaternion need not be exactly unit: the centre is taken back
  // through the conjugate and the extents are then measured about it
  const double m[3] = {0.5 * (lo[0] + hi[0]), 0.5 * (lo[1] + hi[1]), 0.5 * (lo[2] + hi[2])};
  const double n2 = q.x * q.x + q.y * q.y + q.z * q.z + q.w * q.w;
  double c[3];
  Rotate(qi, m, c);
  for (int k = 0; k < 3; ++k) b.center[k] = (float)(c[k] / n2);
  for (int k = 0; k < 3; ++k) b.half[k] = 0.0;
  for (size_t i = 0; i < pts.size(); i += 3) {
    const double rel[3] = {pts[i] - b.center[0], pts[i + 1] - b.center[1], pts[i + 2] - b.center[2]};
    Rotate(q, rel, y);
    for (int k = 0; k < 3; ++k) b.half[k] = std::max(b.half[k], std::fabs(y[k]));
  }
  b.area = b.half[0] * b.half[1] + b.half[1] * b.half[2] + b.half[2] * b.half[0];
  return b;
}

}  // namespace

std::vector<float> FitOBoxHierarchy(const float* hier8, uint32_t numNodes, const int32_t* slots, uint32_t numLeaves,
                                    const float* positions, uint32_t numVertices) {
  std::vector<float> out((size_t)numNodes * 12, 0.0f);
  // scene scale for the float margin
  double smin[3] = {INFINITY, INFINITY, INFINITY}, smax[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (uint32_t v = 0; v < numVertices; ++v)
    for (int k = 0; k < 3; ++k) {
      smin[k] = std::min(smin[k], (double)positions[3 * v + k]);
      smax[k] = std::max(smax[k], (double)positions[3 * v + k]);
    }
  double scale = 0.0;
  for (int k = 0; k < 3; ++k) scale = std::max(scale, std::max(std::fabs(smin[k]), std::fabs(smax[k])));
  const double margin = 4e-6 * std::max(scale, 1e-30);

  auto child = [&](uint32_t n) { uint32_t c; std::memcpy(&c, hier8 + (size_t)n * 8 + 3, 4); return c; };
  auto escape = [&](uint32_t n) { uint32_t e; std::memcpy(&e, hier8 + (size_t)n * 8 + 7, 4); return e; };
  // nodes are independent: interleaved over up to 16 threads
  const unsigned hw = std::thread::hardware_concurrency();
  const unsigned numThreads = numNodes < 4096 ? 1u : std::max(1u, std::min(16u, hw));
  std::atomic<bool> failed{false};
  auto work = [&](unsigned tIdx) {
    try {
      std::vector<double> pts;
      std::vector<uint32_t> stack;
      for (uint32_t n = tIdx; n < numNodes; n += numThreads) {
        // the subtree's valid triangle vertices: a node's children are its
        // firstChild and that child's escape (the left child escapes to its sibling)
        pts.clear();
        stack.assign(1, n);
        while (!stack.empty()) {
          const uint32_t m = stack.back();
          stack.pop_back();
          const uint32_t fc = child(m);
          if (fc & 0x80000000u) {
            const uint32_t leaf = fc & 0x7fffffffu;
            if (leaf >= numLeaves) continue;
            for (int s = 0; s < 2; ++s) {
              const int32_t* t = slots + ((size_t)leaf * 2 + s) * 4;
              if (t[0] == t[1]) break;
              for (int k = 0; k < 3; ++k)
                for (int a = 0; a < 3; ++a) pts.push_back(positions[(size_t)t[k] * 3 + a]);
            }
          } else if (fc < numNodes) {
            stack.push_back(fc);
            const uint32_t second = escape(fc);
            if (second != 0 && second < numNodes && second != escape(m)) stack.push_back(second);
          }
        }
        float* o = out.data() + (size_t)n * 12;
        std::memcpy(o + 3, hier8 + (size_t)n * 8 + 3, 4);  // firstChild
        std::memcpy(o + 7, hier8 + (size_t)n * 8 + 7, 4);  // escape
        if (pts.empty()) {  // an empty subtree: a point box far from nothing (never entered: the walk
          o[11] = 1.0f;     // still tests it, and a zero-size box at the origin holds no triangle)
          continue;
        }
        // candidate orientations: world axes and principal axes
        Box best = BoxFor(Q{0, 0, 0, 1}, pts);
        const size_t np = pts.size() / 3;
        if (np >= 4) {
          double mean[3] = {0, 0, 0};
          for (size_t i = 0; i < pts.size(); i += 3)
            for (int k = 0; k < 3; ++k) mean[k] += pts[i + k];
          for (int k = 0; k < 3; ++k) mean[k] /= (double)np;
          double cov[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
          for (size_t i = 0; i < pts.size(); i += 3)
            for (int a = 0; a < 3; ++a)
              for (int b2 = 0; b2 < 3; ++b2) cov[a][b2] += (pts[i + a] - mean[a]) * (pts[i + b2] - mean[b2]);
          double ev[3][3];
          Eigenvectors(cov, ev);
          // a right-handed frame: rows are the local axes (R maps world to local)
          const double cr[3] = {ev[0][1] * ev[1][2] - ev[0][2] * ev[1][1], ev[0][2] * ev[1][0] - ev[0][0] * ev[1][2],
                                ev[0][0] * ev[1][1] - ev[0][1] * ev[1][0]};
          if (cr[0] * ev[2][0] + cr[1] * ev[2][1] + cr[2] * ev[2][2] < 0.0)
            for (int k = 0; k < 3; ++k) ev[2][k] = -ev[2][k];
          const Box pca = BoxFor(FromMatrix(ev), pts);
          if (pca.area < best.area) best = pca;
        }
        for (int k = 0; k < 3; ++k) o[k] = best.center[k];
        for (int k = 0; k < 3; ++k) {
          // widened by the margin and rounded up to float
          const double h = best.half[k] * (1.0 + 1e-6) + margin;
          float hf = (float)h;
          if ((double)hf < h) hf = std::nextafter(hf, INFINITY);
          o[4 + k] = hf;
        }
        for (int k = 0; k < 4; ++k) o[8 + k] = best.q[k];
      }
    } catch (const std::bad_alloc&) {  // reported on the calling thread
      failed = true;
    }
  };
  if (numThreads == 1) {
    work(0);
  } else {
    std::vector<std::thread> threads;
    for (unsigned t = 0; t < numThreads; ++t) threads.emplace_back(work, t);
    for (auto& t : threads) t.join();
  }
  if (failed) throw std::bad_alloc();
  return out;
}

}  // namespace gpugi
