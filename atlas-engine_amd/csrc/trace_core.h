// Per-lane BVH2 traversal state machine shared by the BLAS, world and batched
// kernels.
//
// Semantics are the engine's HitClosest / HitAny (data/shader/raytracer/bvh.hsh:191-273,
// 359-441), restated as a step function so that one lane can be refilled with
// a new ray while the rest of its wave is still traversing:
//   closest: at an inner node test both child boxes against the current hit
//            distance (IntersectAABB, intersections.hsh:18-34); go to the
//            nearer child (dist of a missed box = current t; ties go left),
//            push the other if both were hit, pop if none;
//            at a leaf test its triangles in order, accepting t > tmin and
//            t < current t (CheckLeafClosest, bvh.hsh:44-72: its extra
//            t < tmax-at-entry test is implied since t only decreases).
//   any:     left child first regardless of distance, stop at the first
//            accepted triangle (CheckLeaf, bvh.hsh:74-104).
// One step processes one inner node or one triangle.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "rt_device.h"
#include "rt_internal.h"

namespace atlas_rt {

// Stack: lane-contiguous LDS rows [LDS_DEPTH][BLOCK], then global spill rows
// [cap - LDS_DEPTH][spill_stride].
template <int BLOCK, int LDS_DEPTH>
struct LaneStack {
    int32_t* lds;
    int32_t* spill;
    uint32_t lane;
    uint32_t column;
    uint32_t stride;

    __device__ __forceinline__ void put(uint32_t sp, int32_t v) const {
        if (sp < LDS_DEPTH)
            lds[sp * BLOCK + lane] = v;
        else
            spill[size_t(sp - LDS_DEPTH) * stride + column] = v;
    }
    __device__ __forceinline__ int32_t get(uint32_t sp) const {
        if (sp < LDS_DEPTH) return lds[sp * BLOCK + lane];
        return spill[size_t(sp - LDS_DEPTH) * stride + column];
    }
};

struct LaneState {
    DevRay R;
    float hitT;
    int32_t node;   // >= 0 inner node, < 0: ~(current triangle slot)
    int32_t slot;   // best hit slot (-1 none)
    int32_t prim;   // caller index of the best hit
    uint32_t sp;    // stack entries
};

enum StepResult : int { kContinue = 0, kDone = 1, kOverflow = 2 };

template <int BLOCK, int LDS_DEPTH, bool STATS>
__device__ __forceinline__ int step_closest(const BlasDev& B, LaneState& L, const LaneStack<BLOCK, LDS_DEPTH>& S,
                                            uint32_t cap, uint32_t& n_nodes, uint32_t& n_tris) {
    if (L.node >= 0) {
        const float4* np = B.nodes + 4 * size_t(L.node);
        const float4 a = np[0], b = np[1], c = np[2], d = np[3];
        if (STATS) n_nodes++;
        float dl, dr;
        const bool hl = slab_dist(L.R, a.x, a.y, a.z, a.w, b.x, b.y, L.hitT, dl);
        const bool hr = slab_dist(L.R, b.z, b.w, c.x, c.y, c.z, c.w, L.hitT, dr);
        const int32_t lp = __float_as_int(d.x), rp = __float_as_int(d.y);
        const bool leftFirst = dl <= dr;
        if (hl && hr) {
            if (L.sp >= cap) return kOverflow;
            S.put(L.sp, leftFirst ? rp : lp);
            L.sp++;
            L.node = leftFirst ? lp : rp;
        } else if (hl || hr) {
            L.node = leftFirst ? lp : rp;
        } else {
            if (L.sp == 0) return kDone;
            L.sp--;
            L.node = S.get(L.sp);
        }
        return kContinue;
    }
    const int32_t tp = ~L.node;
    const float4* tr = B.tris + 3 * size_t(tp);
    const float4 a = tr[0], b = tr[1], c = tr[2];
    if (STATS) n_tris++;
    float t, u, v;
    const bool geo = tri_uvt(L.R, a.x, a.y, a.z, b.x, b.y, b.z, c.x, c.y, c.z, t, u, v);
    if (geo && t > L.R.tmin && t < L.hitT) {
        L.hitT = t;
        L.slot = tp;
        L.prim = __float_as_int(a.w);
    }
    if (b.w > 0.0f) {  // end of leaf
        if (L.sp == 0) return kDone;
        L.sp--;
        L.node = S.get(L.sp);
    } else {
        L.node = ~(tp + 1);
    }
    return kContinue;
}

// Any-hit step; sets L.slot >= 0 and returns kDone on the first accepted triangle.
template <int BLOCK, int LDS_DEPTH>
__device__ __forceinline__ int step_any(const BlasDev& B, LaneState& L, const LaneStack<BLOCK, LDS_DEPTH>& S,
                                        uint32_t cap, float tmax) {
    if (L.node >= 0) {
        const float4* np = B.nodes + 4 * size_t(L.node);
        const float4 a = np[0], b = np[1], c = np[2], d = np[3];
        float dl, dr;
        const bool hl = slab_dist(L.R, a.x, a.y, a.z, a.w, b.x, b.y, tmax, dl);
        const bool hr = slab_dist(L.R, b.z, b.w, c.x, c.y, c.z, c.w, tmax, dr);
        const int32_t lp = __float_as_int(d.x), rp = __float_as_int(d.y);
        if (hl && hr) {
            if (L.sp >= cap) return kOverflow;
            S.put(L.sp, rp);
            L.sp++;
            L.node = lp;
        } else if (hl || hr) {
            L.node = hl ? lp : rp;
        } else {
            if (L.sp == 0) return kDone;
            L.sp--;
            L.node = S.get(L.sp);
        }
        return kContinue;
    }
    const int32_t tp = ~L.node;
    const float4* tr = B.tris + 3 * size_t(tp);
    const float4 a = tr[0], b = tr[1], c = tr[2];
    float t, u, v;
    const bool geo = tri_uvt(L.R, a.x, a.y, a.z, b.x, b.y, b.z, c.x, c.y, c.z, t, u, v);
    if (geo && t > L.R.tmin && t < tmax) {
        L.slot = tp;
        L.hitT = t;
        return kDone;
    }
    if (b.w > 0.0f) {
        if (L.sp == 0) return kDone;
        L.sp--;
        L.node = S.get(L.sp);
    } else {
        L.node = ~(tp + 1);
    }
    return kContinue;
}

__device__ __forceinline__ void lane_begin(LaneState& L, const BlasDev& B, float4 r0, float4 r1) {
    make_dev_ray(L.R, r0.x, r0.y, r0.z, r1.x, r1.y, r1.z, r0.w);
    L.hitT = r1.w;
    L.node = B.node_count ? 0 : ~0;
    L.slot = -1;
    L.prim = -1;
    L.sp = 0;
}

// Final hit record (t, u, v, caller index) for the compact 16 B hit format.
__device__ __forceinline__ float4 lane_hit_record(const BlasDev& B, const LaneState& L, float tmax) {
    if (L.slot < 0) return make_float4(tmax, 0.0f, 0.0f, __int_as_float(-1));
    const float4* vp = B.vtx + 3 * size_t(L.slot);
    float t, u, v;
    refine_tuv(L.R.ox, L.R.oy, L.R.oz, L.R.dx, L.R.dy, L.R.dz, vp[0], vp[1], vp[2], t, u, v);
    return make_float4(t, u, v, __int_as_float(L.prim));
}

}  // namespace atlas_rt
