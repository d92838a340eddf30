// Host-side BVH builder of the product library.
//
// Produces the engine's BLAS / TLAS trees bit-for-bit:
//  - triangle BVH: Volume::BVH::BVH(aabbs, data, parallelBuild) (src/engine/volume/BVH.cpp:14-56):
//    the root tries a binned-SAH object split and an SBVH spatial split
//    (BVH.cpp:249-341); every other node uses the AABB-only rules of
//    BVH.cpp:343-406 (leaf iff one reference, SAH object split else median split).
//  - AABB BVH (TLAS): Volume::BVH::BVH(aabbs, parallelBuild) (BVH.cpp:58-101), 64 bins.
//  - flattening: BVHBuilder::Flatten (BVH.cpp:408-441), pre-order, larger surface
//    area child first; leaf pointer = ~first reference slot.
// Arithmetic uses glm 0.9.8's scalar min/max/clamp/mix (written out in
// rt_hostmath.h) and no FMA contraction, so the result does not depend on the
// thread count.
#pragma once

#include <cstdint>
#include <vector>

#include "../../include/atlas_rt.h"

namespace atlas_rt {

struct BuildOutput {
    std::vector<atlas_rt_gpu_bvh_node> nodes;  // engine GPUBVHNode layout
    std::vector<uint32_t> ref_idx;             // BVH slot -> caller index
    std::vector<uint8_t> end_of_node;          // last reference of each leaf
    uint32_t max_depth = 0;                    // deepest leaf (root = depth 0)
};

// verts: `count` triangles as 9 floats (v0, v1, v2).  boxes: 6 floats per
// triangle (min, max).  threads <= 1 builds serially.
bool build_triangle_bvh(const float* boxes, const float* verts, uint32_t count, int threads, BuildOutput& out);

// boxes: 6 floats per item.
bool build_aabb_bvh(const float* boxes, uint32_t count, int threads, BuildOutput& out);

// Depth of a packed pre-order tree; returns false if the tree is malformed
// (pointer out of range, child index not after its parent, or a leaf run
// without an end-of-leaf marker before `leaf_items`).  end_flags may be NULL
// (then every leaf holds one item).
bool validate_tree(const atlas_rt_gpu_bvh_node* nodes, uint32_t node_count, uint32_t leaf_items,
                   const uint8_t* end_flags, uint32_t* max_depth);

}  // namespace atlas_rt
