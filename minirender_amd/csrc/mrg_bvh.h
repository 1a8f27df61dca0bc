// Host-side SAH BVH builder and flattening into the device layout (mrg_layout.h).
#pragma once

#include <cstdint>
#include <vector>

#include "mrg_layout.h"

namespace mrg {

struct BvhBuildResult {
    std::vector<GpuNode4> nodes;      // exact child boxes (host)
    std::vector<GpuNodeQ> qnodes;     // the same nodes with quantized child boxes (uploaded)
    std::vector<uint32_t> order;      // leaf order -> scene-order primitive
    uint64_t              leaves = 0;
    uint64_t              max_depth = 0;     // binary-tree depth
    uint64_t              max_stack = 0;     // worst-case traversal stack entries of the 4-wide tree
    uint64_t              binary_nodes = 0;
};

// Binned SAH build over world-space FP32 triangles ([n][3][3]) into a binary tree (leaves <= kMaxLeaf
// triangles, depth <= kMaxDepth: median splits take over where SAH would go deeper), collapsed into
// four-wide nodes whose worst-case traversal stack stays within kStackTotal.
BvhBuildResult build_bvh(const float* positions, uint64_t count, unsigned threads);

// Quantizes a node's child boxes (GpuNodeQ): every decoded box contains the exact one.
GpuNodeQ quantize_node(const GpuNode4& n);

// Child k's decoded box (origin + q * scale per plane, exact float arithmetic), as the device sees it.
void decode_child(const GpuNodeQ& q, int k, float lo[3], float hi[3]);

}  // namespace mrg
