// Importance and material passes of the HierarchyImportance renderer
// (hierarchy.hip) and the tree description they run over.
#pragma once
#include <hip/hip_runtime.h>

#include <vector>

#include "integrators.h"

namespace gpugi {

constexpr int kHierPathLength = 8;  // MAX_PATHLENGTH of acquisition2.comp

// Per node the parent (-1 for the root) and the nodes grouped by tree level:
// levelOrder[levelStart[d] .. levelStart[d + 1]) are the nodes at depth d.
// parent and levelOrder are device arrays; levelStart stays on the host.
struct HierTree {
  uint32_t numNodes = 0;
  const int32_t* parent = nullptr;
  const uint32_t* levelOrder = nullptr;
  std::vector<uint32_t> levelStart;
};

// Builds the parent and level arrays on the host from the reference hierarchy
// (8 floats per node: {min, firstChild}{max, escape}): an inner node's children
// are its first child and the siblings its escape chain reaches inside its own
// preorder range.  Returns false on a malformed tree.
bool BuildHierTree(const float* hierarchy, uint32_t numNodes, std::vector<int32_t>* parent,
                   std::vector<uint32_t>* levelOrder, std::vector<uint32_t>* levelStart);

// One acquisition pass (iteration: NumImportanceIterations, 1-based) adding into
// importance (numNodes + triangle slots 64-bit fixed-point sums, zeroed before
// the first pass).
hipError_t LaunchHierAcquire(const DrawParams& p, int iteration, unsigned long long* importance, uint32_t numNodes,
                             hipStream_t stream);
// After the acquisition passes: each node's subtree sum of count_hits' node
// terms (one launch per tree level, deepest first), in place.
hipError_t LaunchHierSubtreeSum(const RefScene& s, const HierTree& tree, unsigned long long* importance,
                                hipStream_t stream);
// The fixed-point sums of `entries` entries to FP32 {sum, 0} pairs, in place.
hipError_t LaunchHierToFloat(unsigned long long* importance, size_t entries, hipStream_t stream);
// Propagation over the tree: importance {x, y} per node (x divided by the
// iteration count), subtree importance per node.
hipError_t LaunchHierPropagate(const RefScene& s, const HierTree& tree, int iterations, float2* importance,
                               float* subtree, hipStream_t stream);
// Area-averaged node materials (12 words per node: six pairs of packed halves).
hipError_t LaunchHierMaterials(const DrawParams& p, const HierTree& tree, uint32_t* materials, hipStream_t stream);

}  // namespace gpugi
