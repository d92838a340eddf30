// TEST INFRASTRUCTURE - the oracle.  Only tests/, __graft_entry__.smoke() and
// bench.py's cpu_baseline leg may load it, and only as the checker.
#pragma once
#include <vector>
#include <cstdint>

#include "../include/atlas_rt.h"
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

// Traversal over the engine's packed arrays (rt_oracle_trace.cpp).
struct TraceStats {
    uint32_t node_visits = 0;  // inner nodes fetched (BLAS + TLAS)
    uint32_t tri_tests = 0;    // triangles intersected
    uint32_t inst_visits = 0;  // instances entered (CheckInstance calls)
    // Equal-t ties: a triangle that passed every accept test except the
    // strict t < hitDistance because its t equals the current best (so the
    // traversal order picked the winner).  tie_t is the distance of the last
    // such event; the ray's final hit is a tie iff tie && tie_t == hit t.
    bool tie = false;
    float tie_t = 0.0f;
};

struct HitRec {
    float t = 0, u = 0, v = 0;
    int32_t slot = -1;      // engine hitID (BVH slot in the BLAS)
    int32_t instance = -1;  // engine hitInstanceID (TLAS instance slot)
    float ho[3] = {0, 0, 0}, hd[3] = {0, 0, 0};  // object-space ray of the accepted test
};

struct BlasView {
    const atlas_rt_gpu_bvh_node* nodes;
    uint32_t node_count;
    const atlas_rt_gpu_bvh_triangle* tris;
    uint32_t tri_count;
};

struct World {
    const atlas_rt_gpu_bvh_node* tlas;
    uint32_t tlas_count;
    const atlas_rt_gpu_bvh_instance* inst;
    uint32_t inst_count;
    const BlasView* blas;
    uint32_t blas_count;
};

// Opacity of a candidate hit (instance slot, BVH slot, object-space u, v):
// CheckLeafClosestTransparency / CheckLeafTransparency (bvh.hsh:106-170).
struct Opacity {
    const void* ctx;
    float (*fn)(const void* ctx, int32_t instance, int32_t slot, float u, float v);
};

// op = nullptr: HitClosest; otherwise HitClosestTransparency (bvh.hsh:275-357),
// which drops candidates whose opacity is <= 0.
void closest_rgpu_world(const World& w, V3 o0, V3 d0, float tmin, float tmax, uint32_t mask, HitRec& h,
                        TraceStats& st, std::vector<int32_t>& stack, const Opacity* op = nullptr);
// HitAnyTransparency (bvh.hsh:443-524): the transparency left along the ray.
float transparency_rgpu_world(const World& w, V3 o0, V3 d0, float tmin, float tmax, uint32_t mask,
                              TraceStats& st, std::vector<int32_t>& stack, const Opacity& op);
bool any_rgpu_world(const World& w, V3 o0, V3 d0, float tmin, float tmax, uint32_t mask, TraceStats& st,
                    std::vector<int32_t>& stack, bool restore_always);

}  // namespace oracle
