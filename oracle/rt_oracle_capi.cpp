// TEST INFRASTRUCTURE - the oracle.  Only tests/, __graft_entry__.smoke() and
// bench.py's cpu_baseline leg may load it, and only as the checker.
//
// C entry points of the oracle builder, for ctypes.  The packing into the GPU
// layout follows MeshData::BuildBVH (src/engine/mesh/MeshData.cpp:172-269) and
// RayTracingWorld::UpdateForSoftwareRayTracing (src/engine/raytracing/RayTracingWorld.cpp:267-307).
#include <cstring>

#include "../include/atlas_rt.h"
#include "rt_oracle.h"

using namespace oracle;

struct OracleBVH {
    BuildResult res;
};

static void pack_node(const BVHNodeOut& n, atlas_rt_gpu_bvh_node& g) {
    std::memset(&g, 0, sizeof(g));
    g.left_min[0] = n.leftMin.x; g.left_min[1] = n.leftMin.y; g.left_min[2] = n.leftMin.z;
    g.left_max[0] = n.leftMax.x; g.left_max[1] = n.leftMax.y; g.left_max[2] = n.leftMax.z;
    g.right_min[0] = n.rightMin.x; g.right_min[1] = n.rightMin.y; g.right_min[2] = n.rightMin.z;
    g.right_max[0] = n.rightMax.x; g.right_max[1] = n.rightMax.y; g.right_max[2] = n.rightMax.z;
    g.left_ptr = n.leftPtr;
    g.right_ptr = n.rightPtr;
}

extern "C" {

// verts: n x 9 floats (v0,v1,v2).  aabbs: n x 6 floats (min,max) or NULL, in
// which case the per-triangle vertex bounds are used (MeshData.cpp:143-153).
OracleBVH* oracle_bvh_build_tris(const float* aabbs, const float* verts, uint32_t n) {
    std::vector<Box> boxes(n);
    for (uint32_t i = 0; i < n; i++) {
        if (aabbs) {
            const float* a = aabbs + 6 * size_t(i);
            boxes[i] = Box{V3{a[0], a[1], a[2]}, V3{a[3], a[4], a[5]}};
        } else {
            const float* v = verts + 9 * size_t(i);
            V3 v0{v[0], v[1], v[2]}, v1{v[3], v[4], v[5]}, v2{v[6], v[7], v[8]};
            boxes[i] = Box{gmin(gmin(v0, v1), v2), gmax(gmax(v0, v1), v2)};
        }
    }
    auto* o = new OracleBVH();
    if (!build_triangle_bvh(boxes, verts, n, o->res)) {
        delete o;
        return nullptr;
    }
    return o;
}

OracleBVH* oracle_bvh_build_aabbs(const float* aabbs, uint32_t n) {
    std::vector<Box> boxes(n);
    for (uint32_t i = 0; i < n; i++) {
        const float* a = aabbs + 6 * size_t(i);
        boxes[i] = Box{V3{a[0], a[1], a[2]}, V3{a[3], a[4], a[5]}};
    }
    auto* o = new OracleBVH();
    if (!build_aabb_bvh(boxes, o->res)) {
        delete o;
        return nullptr;
    }
    return o;
}

uint32_t oracle_bvh_node_count(const OracleBVH* o) { return uint32_t(o->res.nodes.size()); }
uint32_t oracle_bvh_ref_count(const OracleBVH* o) { return uint32_t(o->res.refIdx.size()); }

// Writes the GPUBVHNode array, the BVH-order reference indices and end-of-leaf
// flags.  If verts != NULL, also writes GPUBVHTriangle records in BVH order
// (v0.w = +1/-1 end of leaf, v1.w = material index bits, v2.w = opacity).
void oracle_bvh_export(const OracleBVH* o, atlas_rt_gpu_bvh_node* nodes, uint32_t* ref_idx, uint8_t* end_of_node,
                       const float* verts, atlas_rt_gpu_bvh_triangle* tris, int32_t material_idx, float opacity) {
    const auto& r = o->res;
    if (nodes)
        for (size_t i = 0; i < r.nodes.size(); i++) pack_node(r.nodes[i], nodes[i]);
    if (ref_idx) std::memcpy(ref_idx, r.refIdx.data(), r.refIdx.size() * sizeof(uint32_t));
    if (end_of_node) std::memcpy(end_of_node, r.endOfNode.data(), r.endOfNode.size());
    if (verts && tris) {
        for (size_t i = 0; i < r.refIdx.size(); i++) {
            const float* v = verts + 9 * size_t(r.refIdx[i]);
            auto& t = tris[i];
            t.v0[0] = v[0]; t.v0[1] = v[1]; t.v0[2] = v[2];
            t.v0[3] = r.endOfNode[i] ? 1.0f : -1.0f;
            t.v1[0] = v[3]; t.v1[1] = v[4]; t.v1[2] = v[5];
            std::memcpy(&t.v1[3], &material_idx, 4);
            t.v2[0] = v[6]; t.v2[1] = v[7]; t.v2[2] = v[8];
            t.v2[3] = opacity;
        }
    }
}

void oracle_bvh_free(OracleBVH* o) { delete o; }

// Maximum depth of a packed BVH (root = depth 0; leaves count as one more level).
uint32_t oracle_bvh_depth(const atlas_rt_gpu_bvh_node* nodes, uint32_t count) {
    if (!count) return 0;
    std::vector<uint32_t> depth(count, 0);
    uint32_t maxd = 0;
    for (uint32_t i = 0; i < count; i++) {
        uint32_t d = depth[i];
        int32_t ptrs[2] = {nodes[i].left_ptr, nodes[i].right_ptr};
        for (int k = 0; k < 2; k++) {
            if (ptrs[k] >= 0 && uint32_t(ptrs[k]) < count) depth[ptrs[k]] = d + 1;
            else if (ptrs[k] < 0 && d + 1 > maxd) maxd = d + 1;
        }
    }
    return maxd;
}

}  // extern "C"
