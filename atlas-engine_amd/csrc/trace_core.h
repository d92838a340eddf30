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

// Entries of the register-resident top of the stack (register stack kernels).
constexpr int kRegTop = 6;

// The XCD (accelerator complex) the wave runs on.
__device__ __forceinline__ uint32_t xcc_id() {
    uint32_t v;
    asm volatile("s_getreg_b32 %0, hwreg(HW_REG_XCC_ID)" : "=s"(v));
    return v & 0xfu;
}

// A wave's position in the kPullRanges ranges of [0, total): counter r sits
// at counters[r * stride] (32-bit words).  take() is called by one lane: it
// reserves n items from the current range, moving to the next range while
// the current one is dry, and returns the reserved [base, end) clipped to the
// range (base == end when every range is dry: then dry == kPullRanges).
struct RangePull {
    uint32_t span, cur, dry;
    __device__ __forceinline__ void init(uint32_t total) {
        span = (total + kPullRanges - 1) / kPullRanges;
        cur = xcc_id() % kPullRanges;
        dry = 0;
    }
    __device__ __forceinline__ void take(uint32_t* counters, uint32_t stride, uint32_t total, uint32_t n,
                                         uint32_t& base, uint32_t& end) {
        base = end = 0;
        while (dry < kPullRanges) {
            const uint32_t lo = cur * span, hi = min(lo + span, total);
            const uint32_t b = lo < hi ? lo + atomicAdd(&counters[size_t(cur) * stride], n) : hi;
            if (b < hi) {
                base = b;
                end = hi;
                return;
            }
            dry++;
            cur = (cur + 1) % kPullRanges;
        }
    }
    // after take() on one lane: every lane of the wave takes its state
    __device__ __forceinline__ void broadcast(int from) {
        cur = __shfl(cur, from, 64);
        dry = __shfl(dry, from, 64);
    }
};

// A BLAS record read from a table in global memory (see ldg4).
__device__ __forceinline__ BlasDev ld_blas(const BlasDev* p) {
    const __attribute__((address_space(1))) BlasDev* g = (const __attribute__((address_space(1))) BlasDev*)p;
    BlasDev b;
    b.nodes = g->nodes;
    b.tris = g->tris;
    b.vtx = g->vtx;
    b.nodes4 = g->nodes4;
    b.nodes8 = g->nodes8;
    b.node_count = g->node_count;
    b.tri_count = g->tri_count;
    return b;
}

struct LaneState {
    DevRay R;
    float hitT;
    int32_t node;   // >= 0 inner node, < 0: ~(current triangle slot)
    int32_t slot;   // best hit slot (-1 none)
    int32_t prim;   // caller index of the best hit
    uint32_t sp;    // stack entries
    int32_t top[kRegTop];  // register stack: top[0] is the top entry
};

// Stack: lane-contiguous LDS rows [LDS_DEPTH][BLOCK], then global spill rows
// [cap - LDS_DEPTH][spill_stride].
template <int BLOCK, int LDS_DEPTH>
struct LaneStack {
    int32_t* lds;
    int32_t* spill;
    uint32_t lane;
    uint32_t column;
    uint32_t stride;

    // The LDS rows are addressed through an LDS (address space 3) pointer:
    // with two generic pointers the compiler merges the two loads of get()
    // into one flat load of a selected address, which waits for every
    // outstanding vector memory access as well as the LDS.
    using LdsPtr = __attribute__((address_space(3))) int32_t*;
    __device__ __forceinline__ LdsPtr rows() const { return (LdsPtr)lds; }
    __device__ __forceinline__ void put(uint32_t sp, int32_t v) const {
        if (sp < LDS_DEPTH)
            rows()[sp * BLOCK + lane] = v;
        else
            spill[size_t(sp - LDS_DEPTH) * stride + column] = v;
    }
    __device__ __forceinline__ int32_t get(uint32_t sp) const {
        if (sp < LDS_DEPTH) return rows()[sp * BLOCK + lane];
        return spill[size_t(sp - LDS_DEPTH) * stride + column];
    }
};

// Stack placements for the generic step functions (push / pop on L.sp).
// MemStack: every entry in the LDS rows / global spill of LaneStack.
template <int BLOCK, int LDS_DEPTH>
struct MemStack {
    LaneStack<BLOCK, LDS_DEPTH> m;
    __device__ __forceinline__ void push(LaneState& L, int32_t v) const {
        m.put(L.sp, v);
        L.sp++;
    }
    __device__ __forceinline__ int32_t pop(LaneState& L) const {
        L.sp--;
        return m.get(L.sp);
    }
};

// RegStack: the top kRegTop entries live in VGPRs (shifted on push / pop, so
// every register index is static); deeper entries go to LaneStack memory.
template <int BLOCK, int LDS_DEPTH>
struct RegStack {
    LaneStack<BLOCK, LDS_DEPTH> m;
    __device__ __forceinline__ void push(LaneState& L, int32_t v) const {
        if (L.sp >= uint32_t(kRegTop)) m.put(L.sp - kRegTop, L.top[kRegTop - 1]);
#pragma unroll
        for (int i = kRegTop - 1; i > 0; i--) L.top[i] = L.top[i - 1];
        L.top[0] = v;
        L.sp++;
    }
    __device__ __forceinline__ int32_t pop(LaneState& L) const {
        const int32_t v = L.top[0];
#pragma unroll
        for (int i = 0; i < kRegTop - 1; i++) L.top[i] = L.top[i + 1];
        L.sp--;
        if (L.sp >= uint32_t(kRegTop)) L.top[kRegTop - 1] = m.get(L.sp - kRegTop);
        return v;
    }
};

enum StepResult : int { kContinue = 0, kDone = 1, kOverflow = 2 };

template <int BLOCK, int LDS_DEPTH, bool STATS>
__device__ __forceinline__ int step_closest(const BlasDev& B, LaneState& L, const LaneStack<BLOCK, LDS_DEPTH>& S,
                                            uint32_t cap, uint32_t& n_nodes, uint32_t& n_tris) {
    if (L.node >= 0) {
        const float4* np = B.nodes + 4 * size_t(L.node);
        const float4 a = ldg4(np), b = ldg4(np + 1), c = ldg4(np + 2), d = ldg4(np + 3);
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
    const float4 a = ldg4(tr), b = ldg4(tr + 1), c = ldg4(tr + 2);
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
        const float4 a = ldg4(np), b = ldg4(np + 1), c = ldg4(np + 2), d = ldg4(np + 3);
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
    const float4 a = ldg4(tr), b = ldg4(tr + 1), c = ldg4(tr + 2);
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

// ---------------------------------------------------------------------------
// Generic-stack and wide-node variants (the autotuner's other candidates).
// Nearest hit distance and the set of reachable triangles are those of the
// binary traversal (every wide child box is a binary node's box); the
// winner among triangles hit at exactly the same t may differ.
// ---------------------------------------------------------------------------

// Leaf step shared by every variant: test triangle ~L.node, advance within
// the leaf or pop.
template <class STK>
__device__ __forceinline__ int leaf_step_closest(const BlasDev& B, LaneState& L, const STK& S, uint32_t& n_tris,
                                                 bool stats) {
    const int32_t tp = ~L.node;
    const float4* tr = B.tris + 3 * size_t(tp);
    const float4 a = ldg4(tr), b = ldg4(tr + 1), c = ldg4(tr + 2);
    if (stats) n_tris++;
    float t, u, v;
    const bool geo = tri_uvt(L.R, a.x, a.y, a.z, b.x, b.y, b.z, c.x, c.y, c.z, t, u, v);
    if (geo && t > L.R.tmin && t < L.hitT) {
        L.hitT = t;
        L.slot = tp;
        L.prim = __float_as_int(a.w);
    }
    if (b.w > 0.0f) {
        if (L.sp == 0) return kDone;
        L.node = S.pop(L);
    } else {
        L.node = ~(tp + 1);
    }
    return kContinue;
}

// Inner-node step of a binary tree (L.node >= 0), engine order, any stack
// placement.  FLOAT_SLAB: the widened float slab test (node_test_float)
// instead of the engine-exact one.
template <bool STATS, bool FLOAT_SLAB, class STK>
__device__ __forceinline__ int node_step_closest(const BlasDev& B, LaneState& L, const STK& S, uint32_t cap,
                                                 uint32_t& n_nodes) {
    const float4* np = B.nodes + 4 * size_t(L.node);
    const float4 a = ldg4(np), b = ldg4(np + 1), c = ldg4(np + 2), d = ldg4(np + 3);
    if (STATS) n_nodes++;
    bool hl, hr, leftFirst;
    if (FLOAT_SLAB) {
        node_test_float(L.R, a, b, c, L.hitT, hl, hr, leftFirst);
    } else {
        float dl, dr;
        hl = slab_dist(L.R, a.x, a.y, a.z, a.w, b.x, b.y, L.hitT, dl);
        hr = slab_dist(L.R, b.z, b.w, c.x, c.y, c.z, c.w, L.hitT, dr);
        leftFirst = dl <= dr;
    }
    const int32_t lp = __float_as_int(d.x), rp = __float_as_int(d.y);
    if (hl && hr) {
        if (L.sp >= cap) return kOverflow;
        S.push(L, leftFirst ? rp : lp);
        L.node = leftFirst ? lp : rp;
    } else if (hl || hr) {
        L.node = leftFirst ? lp : rp;
    } else {
        if (L.sp == 0) return kDone;
        L.node = S.pop(L);
    }
    return kContinue;
}

// One step of a binary tree: an inner node or one triangle.
template <bool STATS, bool FLOAT_SLAB, class STK>
__device__ __forceinline__ int step_closest_s(const BlasDev& B, LaneState& L, const STK& S, uint32_t cap,
                                              uint32_t& n_nodes, uint32_t& n_tris) {
    if (L.node < 0) return leaf_step_closest(B, L, S, n_tris, STATS);
    return node_step_closest<STATS, FLOAT_SLAB>(B, L, S, cap, n_nodes);
}

__device__ __forceinline__ float comp4(const float4& v, int k) {
    return k == 0 ? v.x : (k == 1 ? v.y : (k == 2 ? v.z : v.w));
}

// Wide leaf step: L.node = ~((slot << 3) | code), code = triangles left - 1,
// or 7: continue to the end-of-leaf flag.
template <class STK>
__device__ __forceinline__ int wide_leaf_step_closest(const BlasDev& B, LaneState& L, const STK& S, uint32_t& n_tris,
                                                      bool stats) {
    const uint32_t enc = uint32_t(~L.node);
    const uint32_t tp = enc >> 3, code = enc & 7u;
    const float4* tr = B.tris + 3 * size_t(tp);
    const float4 a = ldg4(tr), b = ldg4(tr + 1), c = ldg4(tr + 2);
    if (stats) n_tris++;
    float t, u, v;
    const bool geo = tri_uvt(L.R, a.x, a.y, a.z, b.x, b.y, b.z, c.x, c.y, c.z, t, u, v);
    if (geo && t > L.R.tmin && t < L.hitT) {
        L.hitT = t;
        L.slot = int32_t(tp);
        L.prim = __float_as_int(a.w);
    }
    const bool last = code == 7u ? b.w > 0.0f : code == 0u;
    if (last) {
        if (L.sp == 0) return kDone;
        L.node = S.pop(L);
    } else {
        L.node = ~int32_t(((tp + 1) << 3) | (code == 7u ? 7u : code - 1));
    }
    return kContinue;
}

// W-wide nodes (W = 4 or 8): test every child box, visit hit children
// nearest first (equal entry distance: lower child index first).
// FLOAT_SLAB: the widened float box test (slab_dist_float).
template <int W, bool STATS, class STK, bool FLOAT_SLAB = false>
__device__ __forceinline__ int step_closest_w(const BlasDev& B, LaneState& L, const STK& S, uint32_t cap,
                                              uint32_t& n_nodes, uint32_t& n_tris) {
    if (L.node < 0) return wide_leaf_step_closest(B, L, S, n_tris, STATS);
    constexpr int G = W / 4, F = 7 * G + 1;
    const float4* np = (W == 4 ? B.nodes4 : B.nodes8) + F * size_t(L.node);
    float4 f[F];
#pragma unroll
    for (int i = 0; i < F; i++) f[i] = ldg4(np + i);
    if (STATS) n_nodes++;
    const int cnt = __float_as_int(f[7 * G].x);
    float d[W];
    int32_t ptr[W];
    bool h[W];
    int hits = 0;
#pragma unroll
    for (int c = 0; c < W; c++) {
        const int q = c >> 2, k = c & 3;
        ptr[c] = __float_as_int(comp4(f[6 * G + q], k));
        d[c] = 0.0f;
        if (FLOAT_SLAB)
            h[c] = c < cnt && slab_dist_float(L.R, comp4(f[q], k), comp4(f[G + q], k), comp4(f[2 * G + q], k),
                                              comp4(f[3 * G + q], k), comp4(f[4 * G + q], k),
                                              comp4(f[5 * G + q], k), L.hitT, d[c]);
        else
            h[c] = c < cnt && slab_dist(L.R, comp4(f[q], k), comp4(f[G + q], k), comp4(f[2 * G + q], k),
                                        comp4(f[3 * G + q], k), comp4(f[4 * G + q], k), comp4(f[5 * G + q], k),
                                        L.hitT, d[c]);
        hits += h[c] ? 1 : 0;
    }
    if (hits == 0) {
        if (L.sp == 0) return kDone;
        L.node = S.pop(L);
        return kContinue;
    }
    if (L.sp + uint32_t(hits) - 1u > cap) return kOverflow;
    int32_t srt[W];
#pragma unroll
    for (int r = 0; r < W; r++) srt[r] = 0;
#pragma unroll
    for (int c = 0; c < W; c++) {
        int rank = 0;
#pragma unroll
        for (int j = 0; j < W; j++) rank += (h[j] && (d[j] < d[c] || (d[j] == d[c] && j < c))) ? 1 : 0;
#pragma unroll
        for (int r = 0; r < W; r++) srt[r] = (h[c] && rank == r) ? ptr[c] : srt[r];
    }
#pragma unroll
    for (int r = W - 1; r >= 1; r--)
        if (r < hits) S.push(L, srt[r]);
    L.node = srt[0];
    return kContinue;
}

// One closest-hit step of the variant (node width NODE_W in {2, 4, 8};
// REG: register-topped stack; FLOAT_SLAB: widened float slab test).
// NODE_W = 2 without REG or FLOAT_SLAB is the engine-order binary step above.
template <int NODE_W, bool REG, int BLOCK, int LDS_DEPTH, bool STATS, bool FLOAT_SLAB = false>
__device__ __forceinline__ int step_variant(const BlasDev& B, LaneState& L, const LaneStack<BLOCK, LDS_DEPTH>& S,
                                            uint32_t cap, uint32_t& n_nodes, uint32_t& n_tris) {
    if constexpr (NODE_W == 2 && !REG && !FLOAT_SLAB)
        return step_closest<BLOCK, LDS_DEPTH, STATS>(B, L, S, cap, n_nodes, n_tris);
    else if constexpr (NODE_W == 2 && REG)
        return step_closest_s<STATS, FLOAT_SLAB>(B, L, RegStack<BLOCK, LDS_DEPTH>{S}, cap, n_nodes, n_tris);
    else if constexpr (NODE_W == 2)
        return step_closest_s<STATS, FLOAT_SLAB>(B, L, MemStack<BLOCK, LDS_DEPTH>{S}, cap, n_nodes, n_tris);
    else if constexpr (REG)
        return step_closest_w<NODE_W, STATS, RegStack<BLOCK, LDS_DEPTH>, FLOAT_SLAB>(B, L, RegStack<BLOCK, LDS_DEPTH>{S},
                                                                                  cap, n_nodes, n_tris);
    else
        return step_closest_w<NODE_W, STATS, MemStack<BLOCK, LDS_DEPTH>, FLOAT_SLAB>(B, L, MemStack<BLOCK, LDS_DEPTH>{S},
                                                                                  cap, n_nodes, n_tris);
}

// The two halves of a step, for kernels that batch leaf work across a wave:
// an inner-node step (L.node >= 0) and a triangle step (L.node < 0), for
// binary or wide nodes.
template <int NODE_W, bool REG, int BLOCK, int LDS_DEPTH, bool FLOAT_SLAB>
__device__ __forceinline__ int node_step_variant(const BlasDev& B, LaneState& L, const LaneStack<BLOCK, LDS_DEPTH>& S,
                                                 uint32_t cap) {
    uint32_t n_nodes = 0, n_tris = 0;
    if constexpr (NODE_W > 2)
        return step_variant<NODE_W, REG, BLOCK, LDS_DEPTH, false, FLOAT_SLAB>(B, L, S, cap, n_nodes, n_tris);
    else if constexpr (REG)
        return node_step_closest<false, FLOAT_SLAB>(B, L, RegStack<BLOCK, LDS_DEPTH>{S}, cap, n_nodes);
    else
        return node_step_closest<false, FLOAT_SLAB>(B, L, MemStack<BLOCK, LDS_DEPTH>{S}, cap, n_nodes);
}

template <int NODE_W, bool REG, int BLOCK, int LDS_DEPTH>
__device__ __forceinline__ int leaf_step_variant(const BlasDev& B, LaneState& L, const LaneStack<BLOCK, LDS_DEPTH>& S) {
    uint32_t n_tris = 0;
    if constexpr (NODE_W > 2 && REG)
        return wide_leaf_step_closest(B, L, RegStack<BLOCK, LDS_DEPTH>{S}, n_tris, false);
    else if constexpr (NODE_W > 2)
        return wide_leaf_step_closest(B, L, MemStack<BLOCK, LDS_DEPTH>{S}, n_tris, false);
    else if constexpr (REG)
        return leaf_step_closest(B, L, RegStack<BLOCK, LDS_DEPTH>{S}, n_tris, false);
    else
        return leaf_step_closest(B, L, MemStack<BLOCK, LDS_DEPTH>{S}, n_tris, false);
}

// One wave iteration of a resident-wave kernel for an active lane.  LEAF_DIV
// = 0: every active lane takes one step of its variant.  LEAF_DIV > 0
// (binary nodes): the wave takes triangle steps only when at least
// 1 / LEAF_DIV of its live lanes wait at a leaf, and inner-node steps for the
// other lanes otherwise, so the triangle test is not paid by the whole wave
// whenever one lane reaches a leaf.  Each lane's own sequence of steps is
// unchanged.  Every lane of the wave must call it; returns kContinue for a
// lane that did not step.
template <int NODE_W, bool REG, int BLOCK, int LDS_DEPTH, bool FLOAT_SLAB, int LEAF_DIV>
__device__ __forceinline__ int wave_step(const BlasDev& B, LaneState& L, const LaneStack<BLOCK, LDS_DEPTH>& S,
                                         uint32_t cap, bool active) {
    if constexpr (LEAF_DIV > 0) {
        const bool at_leaf = active && L.node < 0;
        const uint32_t nl = uint32_t(__popcll(__ballot(at_leaf)));
        const uint32_t na = uint32_t(__popcll(__ballot(active)));
        if (nl * LEAF_DIV >= na) {  // wave-uniform
            if (at_leaf) return leaf_step_variant<NODE_W, REG, BLOCK, LDS_DEPTH>(B, L, S);
        } else if (active && !at_leaf) {
            return node_step_variant<NODE_W, REG, BLOCK, LDS_DEPTH, FLOAT_SLAB>(B, L, S, cap);
        }
        return kContinue;
    } else {
        uint32_t dn = 0, dt = 0;
        if (!active) return kContinue;
        return step_variant<NODE_W, REG, BLOCK, LDS_DEPTH, false, FLOAT_SLAB>(B, L, S, cap, dn, dt);
    }
}

// Stack bound of a variant for one BLAS (entries).
__host__ __device__ __forceinline__ uint32_t variant_cap(int node_w, uint32_t cap2, uint32_t cap4, uint32_t cap8) {
    return node_w == 2 ? cap2 : (node_w == 4 ? cap4 : cap8);
}

// wide: the root of a single-leaf BLAS in the wide leaf encoding (run to
// the end-of-leaf flag from slot 0).
__device__ __forceinline__ void lane_begin(LaneState& L, const BlasDev& B, float4 r0, float4 r1, bool wide = false) {
    make_dev_ray(L.R, r0.x, r0.y, r0.z, r1.x, r1.y, r1.z, r0.w);
    L.hitT = r1.w;
    L.node = B.node_count ? 0 : (wide ? ~7 : ~0);
    L.slot = -1;
    L.prim = -1;
    L.sp = 0;
}

// Final hit record (t, u, v, caller index) for the compact 16 B hit format.
__device__ __forceinline__ float4 lane_hit_record(const BlasDev& B, const LaneState& L, float tmax) {
    if (L.slot < 0) return make_float4(tmax, 0.0f, 0.0f, __int_as_float(-1));
    const float4* vp = B.vtx + 3 * size_t(L.slot);
    float t, u, v;
    refine_tuv(L.R.ox, L.R.oy, L.R.oz, L.R.dx, L.R.dy, L.R.dz, ldg4(vp), ldg4(vp + 1), ldg4(vp + 2), t, u, v);
    return make_float4(t, u, v, __int_as_float(L.prim));
}

}  // namespace atlas_rt
