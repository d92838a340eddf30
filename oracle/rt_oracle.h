// TEST INFRASTRUCTURE - the oracle.  Only tests/, __graft_entry__.smoke() and
// bench.py's cpu_baseline leg may load it, and only as the checker.
#pragma once
#include <vector>
#include <cstdint>

#include "rt_oracle_math.h"

namespace oracle {

// Volume::BVHNode (BVH.h:14-24) as flattened by BVHBuilder::Flatten.
struct BVHNodeOut {
    V3 leftMin{0, 0, 0}, leftMax{0, 0, 0};
    V3 rightMin{0, 0, 0}, rightMax{0, 0, 0};
    int32_t leftPtr = 0;
    int32_t rightPtr = 0;
};

struct BuildResult {
    std::vector<BVHNodeOut> nodes;
    std::vector<uint32_t> refIdx;    // BVH slot -> input index (refs[i].idx)
    std::vector<uint8_t> endOfNode;  // refs[i].endOfNode
};

bool build_triangle_bvh(const std::vector<Box>& aabbs, const float* verts, uint32_t n, BuildResult& res);
bool build_aabb_bvh(const std::vector<Box>& aabbs, BuildResult& res);

}  // namespace oracle
