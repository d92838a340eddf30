// Per-lane two-level (TLAS + instanced BLAS) traversal state machine, shared
// by the world trace kernels and the path tracer.
//
// Reference: HitClosest / HitAny with CheckInstance (data/shader/raytracer/bvh.hsh:172-273,
// 359-441).  One stack serves both levels; `tlas_index` (the engine's tlasIndex)
// records the stack depth at which the current instance was entered, so
// popping below it returns to the TLAS with the world-space ray restored.
// Entering an instance transforms the ray with the instance's inverse matrix
// as a row vector times mat3x4 (unnormalised direction, so t stays in world
// units, bvh.hsh:177-180) and selects that instance's BLAS.
#pragma once

#include <hip/hip_runtime.h>

#include "trace_core.h"
#include "world.h"

namespace atlas_rt {

constexpr uint32_t kTlasInvalid = 0xFFFFFFFFu;

struct WorldLane {
    float o0x, o0y, o0z, d0x, d0y, d0z;  // world ray
    double wrx, wry, wrz;                // world reciprocals
    DevRay R;                            // ray of the current level
    float hitT;
    int32_t node;
    int32_t slot;
    int32_t inst;      // instance slot of the best hit
    int32_t cur_inst;  // instance being traversed
    uint32_t sp;
    uint32_t tlas_index;
    BlasDev B;
    // object-space ray of the best hit (for the reported t, u, v)
    float hox, hoy, hoz, hdx, hdy, hdz;
    float trans;  // HitAnyTransparency's transparency
};

// Candidate filter of the closest-hit walk: HitClosest accepts every
// geometric hit; HitClosestTransparency passes an opacity test instead.
struct AcceptAll {
    __device__ __forceinline__ bool operator()(int32_t, int32_t, float, float) const { return true; }
};

__device__ __forceinline__ float dot4(float x, float y, float z, float w, float4 m) {
    return ((x * m.x + y * m.y) + z * m.z) + w * m.w;
}

__device__ __forceinline__ void restore_world(WorldLane& L) {
    L.R.ox = L.o0x;
    L.R.oy = L.o0y;
    L.R.oz = L.o0z;
    L.R.dx = L.d0x;
    L.R.dy = L.d0y;
    L.R.dz = L.d0z;
    L.R.rx = L.wrx;
    L.R.ry = L.wry;
    L.R.rz = L.wrz;
}

__device__ __forceinline__ void world_begin(WorldLane& L, float4 r0, float4 r1) {
    make_dev_ray(L.R, r0.x, r0.y, r0.z, r1.x, r1.y, r1.z, r0.w);
    L.o0x = r0.x;
    L.o0y = r0.y;
    L.o0z = r0.z;
    L.d0x = r1.x;
    L.d0y = r1.y;
    L.d0z = r1.z;
    L.wrx = L.R.rx;
    L.wry = L.R.ry;
    L.wrz = L.R.rz;
    L.hitT = r1.w;
    L.node = 0;
    L.slot = -1;
    L.inst = -1;
    L.cur_inst = -1;
    L.sp = 0;
    L.tlas_index = kTlasInvalid;
    L.trans = 1.0f;
}

// CheckInstance (bvh.hsh:172-189); returns whether the mask admits it.
__device__ __forceinline__ bool enter_instance(WorldLane& L, const WorldDev& W, uint32_t mask) {
    const int32_t ip = ~L.node;
    const float4* m = W.instances + 4 * size_t(ip);
    const float4 c0 = ldg4(m), c1 = ldg4(m + 1), c2 = ldg4(m + 2), meta = ldg4(m + 3);
    const float ox = dot4(L.o0x, L.o0y, L.o0z, 1.0f, c0);
    const float oy = dot4(L.o0x, L.o0y, L.o0z, 1.0f, c1);
    const float oz = dot4(L.o0x, L.o0y, L.o0z, 1.0f, c2);
    const float dx = dot4(L.d0x, L.d0y, L.d0z, 0.0f, c0);
    const float dy = dot4(L.d0x, L.d0y, L.d0z, 0.0f, c1);
    const float dz = dot4(L.d0x, L.d0y, L.d0z, 0.0f, c2);
    make_dev_ray(L.R, ox, oy, oz, dx, dy, dz, L.R.tmin);
    L.cur_inst = ip;
    const int32_t mesh = __float_as_int(meta.x);
    L.B = ld_blas(W.blases + mesh);
    L.node = L.B.node_count ? 0 : ~0;
    return (uint32_t(__float_as_int(meta.w)) & mask) > 0;
}

// One step of HitClosest (bvh.hsh:191-273); with an opacity filter it is
// HitClosestTransparency (bvh.hsh:275-357), whose leaf test additionally
// needs opacity > 0 (CheckLeafClosestTransparency, bvh.hsh:106-137).
template <int BLOCK, int LDS_DEPTH, class Accept = AcceptAll>
__device__ __forceinline__ int world_step_closest(WorldLane& L, const WorldDev& W, uint32_t mask,
                                                  const LaneStack<BLOCK, LDS_DEPTH>& S, uint32_t cap,
                                                  const Accept& accept = Accept()) {
    const uint32_t spE = L.sp + 1;  // the engine's stackPtr
    if (spE < L.tlas_index) {
        restore_world(L);
        L.tlas_index = kTlasInvalid;
        if (L.node < 0) {
            if (enter_instance(L, W, mask)) {
                L.tlas_index = spE;
            } else {
                if (L.sp == 0) return kDone;
                L.sp--;
                L.node = S.get(L.sp);
            }
            return kContinue;
        }
        const float4* np = W.tlas + 4 * size_t(L.node);
        const float4 a = ldg4(np), b = ldg4(np + 1), c = ldg4(np + 2), d = ldg4(np + 3);
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
    if (L.node >= 0) {
        const float4* np = L.B.nodes + 4 * size_t(L.node);
        const float4 a = ldg4(np), b = ldg4(np + 1), c = ldg4(np + 2), d = ldg4(np + 3);
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
    const float4* tr = L.B.tris + 3 * size_t(tp);
    const float4 a = ldg4(tr), b = ldg4(tr + 1), c = ldg4(tr + 2);
    float t, u, v;
    const bool geo = tri_uvt(L.R, a.x, a.y, a.z, b.x, b.y, b.z, c.x, c.y, c.z, t, u, v);
    if (geo && t > L.R.tmin && t < L.hitT && accept(L.cur_inst, tp, u, v)) {
        L.hitT = t;
        L.slot = tp;
        L.inst = L.cur_inst;
        L.hox = L.R.ox;
        L.hoy = L.R.oy;
        L.hoz = L.R.oz;
        L.hdx = L.R.dx;
        L.hdy = L.R.dy;
        L.hdz = L.R.dz;
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

// RESTORE_ALWAYS = false is HitAny (bvh.hsh:382-387): the world ray is restored
// only when returning from an entered instance, so after an instance the mask
// rejects, the TLAS walk continues with that instance's object-space ray.
// RESTORE_ALWAYS = true is HitAnyTransparency's traversal (bvh.hsh:466-470),
// used for the path tracer's shadow rays.
template <int BLOCK, int LDS_DEPTH, bool RESTORE_ALWAYS>
__device__ __forceinline__ int world_step_any(WorldLane& L, const WorldDev& W, uint32_t mask, float tmax,
                                              const LaneStack<BLOCK, LDS_DEPTH>& S, uint32_t cap) {
    const uint32_t spE = L.sp + 1;
    if (spE < L.tlas_index) {
        if (RESTORE_ALWAYS || L.tlas_index != kTlasInvalid) restore_world(L);
        L.tlas_index = kTlasInvalid;
        if (L.node < 0) {
            if (enter_instance(L, W, mask)) {
                L.tlas_index = spE;
            } else {
                if (L.sp == 0) return kDone;
                L.sp--;
                L.node = S.get(L.sp);
            }
            return kContinue;
        }
        const float4* np = W.tlas + 4 * size_t(L.node);
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
    if (L.node >= 0) {
        const float4* np = L.B.nodes + 4 * size_t(L.node);
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
    const float4* tr = L.B.tris + 3 * size_t(tp);
    const float4 a = ldg4(tr), b = ldg4(tr + 1), c = ldg4(tr + 2);
    float t, u, v;
    const bool geo = tri_uvt(L.R, a.x, a.y, a.z, b.x, b.y, b.z, c.x, c.y, c.z, t, u, v);
    if (geo && t > L.R.tmin && t < tmax) {
        L.slot = tp;
        L.inst = L.cur_inst;
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

// One step of either walk over opaque triangles, chosen per lane, so that a
// wave whose lanes trace closest hits and shadow rays at the same time runs
// one step body instead of both:
//   any = false: world_step_closest (HitClosest);
//   any = true:  world_step_any<RESTORE_ALWAYS = true> (HitAnyTransparency's
//                traversal with every triangle opaque) with tmax = L.hitT,
//                which world_begin set and which the walk never lowers before
//                it ends at its first accepted triangle.
// Both restore the world ray on every TLAS step.  Left-first order of the
// any-hit walk is the closest walk's order with "left is nearer" replaced by
// "left was hit", so the two share every load and slab test.
template <int BLOCK, int LDS_DEPTH>
__device__ __forceinline__ int world_step_mixed(WorldLane& L, const WorldDev& W, uint32_t mask, bool any,
                                                const LaneStack<BLOCK, LDS_DEPTH>& S, uint32_t cap) {
    const uint32_t spE = L.sp + 1;
    const bool top = spE < L.tlas_index;
    if (top) {
        restore_world(L);
        L.tlas_index = kTlasInvalid;
        if (L.node < 0) {
            if (enter_instance(L, W, mask)) {
                L.tlas_index = spE;
            } else {
                if (L.sp == 0) return kDone;
                L.sp--;
                L.node = S.get(L.sp);
            }
            return kContinue;
        }
    }
    if (top || L.node >= 0) {
        const float4* np = (top ? W.tlas : L.B.nodes) + 4 * size_t(L.node);
        const float4 a = ldg4(np), b = ldg4(np + 1), c = ldg4(np + 2), d = ldg4(np + 3);
        float dl, dr;
        const bool hl = slab_dist(L.R, a.x, a.y, a.z, a.w, b.x, b.y, L.hitT, dl);
        const bool hr = slab_dist(L.R, b.z, b.w, c.x, c.y, c.z, c.w, L.hitT, dr);
        const int32_t lp = __float_as_int(d.x), rp = __float_as_int(d.y);
        const bool leftFirst = any ? hl : dl <= dr;
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
    const float4* tr = L.B.tris + 3 * size_t(tp);
    const float4 a = ldg4(tr), b = ldg4(tr + 1), c = ldg4(tr + 2);
    float t, u, v;
    const bool geo = tri_uvt(L.R, a.x, a.y, a.z, b.x, b.y, b.z, c.x, c.y, c.z, t, u, v);
    if (geo && t > L.R.tmin && t < L.hitT) {
        L.hitT = t;
        L.slot = tp;
        L.inst = L.cur_inst;
        if (any) return kDone;
        L.hox = L.R.ox;
        L.hoy = L.R.oy;
        L.hoz = L.R.oz;
        L.hdx = L.R.dx;
        L.hdy = L.R.dy;
        L.hdz = L.R.dz;
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

// One step of HitAnyTransparency (bvh.hsh:443-524) with the world ray
// restored on every TLAS step.  A leaf is handled in one step, as
// CheckLeafTransparency does (bvh.hsh:139-170): starting from the current
// transparency it multiplies (1 - opacity) for every triangle hit in
// (tmin, tmax), and the caller multiplies that product into the transparency
// once more (bvh.hsh:496), so each leaf squares what came before it.  Below
// 1e-6 the transparency is 0 and the walk ends.  `opacity(instance, slot, u, v)`
// gives the triangle's opacity.
template <int BLOCK, int LDS_DEPTH, class Opacity>
__device__ __forceinline__ int world_step_transparency(WorldLane& L, const WorldDev& W, uint32_t mask, float tmax,
                                                       const LaneStack<BLOCK, LDS_DEPTH>& S, uint32_t cap,
                                                       const Opacity& opacity) {
    const uint32_t spE = L.sp + 1;
    if (spE < L.tlas_index) {
        restore_world(L);
        L.tlas_index = kTlasInvalid;
        if (L.node < 0) {
            if (enter_instance(L, W, mask)) {
                L.tlas_index = spE;
            } else {
                if (L.sp == 0) return kDone;
                L.sp--;
                L.node = S.get(L.sp);
            }
            return kContinue;
        }
        const float4* np = W.tlas + 4 * size_t(L.node);
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
    if (L.node >= 0) {
        const float4* np = L.B.nodes + 4 * size_t(L.node);
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
    float leaf = L.trans;
    for (int32_t tp = ~L.node;; tp++) {
        const float4* tr = L.B.tris + 3 * size_t(tp);
        const float4 a = ldg4(tr), b = ldg4(tr + 1), c = ldg4(tr + 2);
        float t, u, v;
        const bool geo = tri_uvt(L.R, a.x, a.y, a.z, b.x, b.y, b.z, c.x, c.y, c.z, t, u, v);
        if (geo && t > L.R.tmin && t < tmax) leaf *= (1.0f - opacity(L.cur_inst, tp, u, v));
        if (b.w > 0.0f) break;
    }
    L.trans *= leaf;
    if (L.trans < 0.000001f) L.trans = 0.0f;
    if (!(L.trans > 0.0f)) return kDone;
    if (L.sp == 0) return kDone;
    L.sp--;
    L.node = S.get(L.sp);
    return kContinue;
}

}  // namespace atlas_rt
