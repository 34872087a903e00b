// Re-packed 4-wide BVH with 8-bit quantized child boxes (host layout + device view).
//
// Node: 64 bytes, read as four 16-byte loads (one half of a 128-byte L2 line):
//   q0 = {origin.x, origin.y, origin.z, exps}   exps = ex | ey<<8 | ez<<16 | nchild<<24,
//                                              scale_a = 2^(e_a - 127)
//   q1 = {child[0..3]}                          inner: node index; leaf: MSB | (count-1)<<27 | first
//   q2 = {lo.x[4], lo.y[4], lo.z[4], hi.x[4]}   one byte per child
//   q3 = {hi.y[4], hi.z[4], 0, 0}
// child box along axis a = [origin_a + lo*scale_a, origin_a + hi*scale_a]; the
// host rounds lo down and hi up and checks in extended precision that the
// exact origin + q * scale encloses the child box, so the quantized tree is conservative.
// Triangle: 48 bytes {p0, slot}, {e0 = p1-p0, material}, {e1 = p0-p2, 0}: the
// operands of the reference triangle test, computed once in FP32 (in
// watertight mode the raw vertices {p0, slot}, {p1, material}, {p2, 0}).
#pragma once
#include <stdint.h>

#include <string>
#include <vector>

#include "gpugi_rt.h"

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#endif

namespace gpugi {

constexpr uint32_t kWideLeafFlag = 0x80000000u;
constexpr uint32_t kWideLeafFirstMask = (1u << 27) - 1;
constexpr uint32_t kWideMaxLeafTris = 16;

struct WideNodeHost {
  float origin[3];
  uint32_t exps;
  uint32_t child[4];
  uint8_t lo[3][4];
  uint8_t hi[3][4];
  uint32_t pad[2];
};
static_assert(sizeof(WideNodeHost) == 64, "wide node is 64 bytes");

struct WideTriHost {
  float p0[3];
  int32_t slot;
  float e0[3];
  int32_t material;
  float e1[3];
  int32_t pad;
};
static_assert(sizeof(WideTriHost) == 48, "wide triangle is 48 bytes");

struct WideBVHHost {
  std::vector<WideNodeHost> nodes;
  std::vector<WideTriHost> tris;
  uint32_t maxStack = 0;   // worst-case traversal stack entries (siblings summed along a path)
  uint32_t depth = 0;
};

// Re-packs the reference BVH2 (any arity written in preorder with escape
// pointers) into the 4-wide quantized layout.  Leaves of up to
// `maxLeafTris` triangles are formed from subtrees that small.  rawVertices:
// triangles store {p0, slot}{p1, mat}{p2, 0} (watertight test) instead of the
// reference test's operands {p0, slot}{p1 - p0, mat}{p0 - p2, 0}.
std::string BuildWideBVH(const gpugi_scene_desc& desc, uint32_t maxLeafTris, WideBVHHost* out,
                         bool rawVertices = false);

// Launch shape of the wide kernel: threads per block and LDS stack entries per thread.
constexpr int kWideBlock = 256;
constexpr int kWideLdsStack = 16;

#if defined(__HIPCC__)
struct WideScene {
  const uint4* __restrict__ nodes;    // 4 x uint4 per node
  const float4* __restrict__ tris;    // 3 x float4 per triangle
  uint2* spill;                       // per-thread overflow stack (global memory)
  uint32_t spillDepth;                // entries per thread in `spill`
  uint32_t numThreads;                // persistent threads the spill buffer is sized for
  uint32_t* errorFlag;                // set to nonzero if a stack would overflow
  uint32_t* queue;                    // kWideQueueWords: the trace kernel's per-XCD batch counters
};
constexpr int kWideQueueStride = 32;  // counters 128 bytes apart
constexpr int kWideQueueWords = 8 * kWideQueueStride;
#endif

}  // namespace gpugi
