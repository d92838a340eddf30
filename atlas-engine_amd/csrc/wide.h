// Wide (4- and 8-child) BVH nodes collapsed from the engine's binary BVH,
// for the wide traversal kernels.
//
// Node layout for width W (G = W / 4 groups of four children), all float4:
//   [b * G + q], b = 0..5 : bound b (min x, min y, min z, max x, max y, max z)
//                           of children 4q .. 4q+3
//   [6G + q]              : child pointers of children 4q .. 4q+3 (int bits):
//                           >= 0 wide node index, < 0 a leaf of `count`
//                           consecutive triangle slots from `first`,
//                           encoded ~((first << 3) | (count - 1)), count <= 7;
//                           code 7 instead of count - 1: the slots run to the
//                           engine's end-of-leaf flag (a larger binary leaf)
//   [7G]                  : x = number of children (int bits); children are
//                           packed first, unused slots hold zero boxes
// W = 4: 8 float4 (128 B); W = 8: 15 float4 (240 B).
// Every child box is the box of a node of the binary tree, so a wide
// traversal tests a subset of the boxes the binary one tests on the way to
// each triangle.  A binary subtree with at most max_leaf triangles becomes
// one wide leaf over its slots (contiguous: the engine's flattening assigns
// slots in pre-order).
#pragma once

#include <cstdint>
#include <vector>

#include "../../include/atlas_rt.h"

namespace atlas_rt {

constexpr int wide_node_float4s(int width) { return 7 * (width / 4) + 1; }

struct WideTree {
    std::vector<float> data;  // node_count * wide_node_float4s(width) * 4 floats
    uint32_t node_count = 0;
    uint32_t max_stack = 0;   // largest number of entries a traversal can hold
};

// Collapses the binary tree (pre-order, root at 0; empty when the BLAS is a
// single leaf) into width-4 or width-8 nodes: starting from a node's two
// children, the inner child with the largest surface area is replaced by its
// own two children until the node has `width` children or only leaves.
// end_of_leaf[slot]: the engine's end-of-leaf flag of each triangle slot.
WideTree collapse_wide(const std::vector<atlas_rt_gpu_bvh_node>& nodes, const std::vector<uint8_t>& end_of_leaf,
                       int width, int max_leaf);

}  // namespace atlas_rt
