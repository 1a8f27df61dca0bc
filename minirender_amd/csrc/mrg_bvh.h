// Host-side SAH BVH builder and flattening into the device layout (mrg_layout.h).
#pragma once

#include <cstdint>
#include <vector>

#include "mrg_layout.h"

namespace mrg {

struct BvhBuildResult {
    std::vector<GpuNode>  nodes;
    std::vector<uint32_t> order;      // leaf order -> scene-order primitive
    uint64_t              leaves = 0;
    uint64_t              max_depth = 0;
};

// Binned SAH build over world-space FP32 triangles ([n][3][3]). Leaves hold <= kMaxLeaf triangles
// and the depth never exceeds kMaxDepth (median splits take over where SAH would go deeper).
BvhBuildResult build_bvh(const float* positions, uint64_t count, unsigned threads);

}  // namespace mrg
