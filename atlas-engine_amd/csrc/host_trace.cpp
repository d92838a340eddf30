// CPU traversal of a BLAS with the engine's own CPU semantics (R-CPU):
// BVH::GetIntersection / GetIntersectionAny (src/engine/volume/BVH.cpp:103-217)
// over Ray::Intersects (src/engine/volume/Ray.cpp:28-76).  This is the
// reference's host API, not a stand-in for the GPU kernels, whose semantics
// (HitClosest / HitAny, bvh.hsh) differ:
//   * slab test with the float reciprocal 1 / d computed once per ray;
//   * closest hit: a stack of (node, entry distance); a popped entry farther
//     than the current hit is skipped; at an inner node both hit children
//     are pushed, the nearer one last (popped first), the left one last when
//     the distances tie; the leaf accepts t >= 0 (the ray's tMin is not
//     applied to triangles) and t < current t;
//   * any hit: right child pushed first, left popped first; the first
//     triangle with t < tMax ends the walk.
// The engine's GetIntersection compares its final t against the ray's tMax
// after the loop has overwritten tMax with the hit distance (BVH.cpp:141,172),
// so it reports no hit whenever the nearest triangle was found before the last
// inner-node test.  Here a hit is reported when its t is below the caller's
// tMax, which is what the engine's callers expect.
// Every operation rounds as written (no FMA contraction; glm's min / max /
// dot / cross evaluation order).
#include <algorithm>
#include <thread>
#include <utility>
#include <vector>

#include "rt_hostmath.h"
#include "rt_internal.h"

namespace atlas_rt {

namespace {

struct F3 {
    float x, y, z;
};
inline F3 sub(F3 a, F3 b) { return F3{a.x - b.x, a.y - b.y, a.z - b.z}; }
inline float dot3(F3 a, F3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
inline F3 cross3(F3 a, F3 b) {
    return F3{a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y};
}

struct CpuRay {
    F3 o, d, inv;
    float tmin, tmax;
};

// Ray::Intersects(const AABB&, float&) (Ray.cpp:28-43).
inline bool slab(const CpuRay& r, const float* lo, const float* hi, float& t) {
    const float t0x = (lo[0] - r.o.x) * r.inv.x, t0y = (lo[1] - r.o.y) * r.inv.y, t0z = (lo[2] - r.o.z) * r.inv.z;
    const float t1x = (hi[0] - r.o.x) * r.inv.x, t1y = (hi[1] - r.o.y) * r.inv.y, t1z = (hi[2] - r.o.z) * r.inv.z;
    const float sx = gl_min(t0x, t1x), sy = gl_min(t0y, t1y), sz = gl_min(t0z, t1z);
    const float bx = gl_max(t0x, t1x), by = gl_max(t0y, t1y), bz = gl_max(t0z, t1z);
    const float tminf = gl_max(r.tmin, gl_max(sx, gl_max(sy, sz)));
    const float tmaxf = gl_min(r.tmax, gl_min(bx, gl_min(by, bz)));
    t = tminf;
    return tminf <= tmaxf;
}

// Ray::Intersects(v0, v1, v2, vec3&) (Ray.cpp:53-76): (t, u, v) with three
// divides by q.e0; accepted when t, u, v >= 0 and u + v <= 1.
inline bool triangle(const CpuRay& r, const atlas_rt_gpu_bvh_triangle& tri, float& t, float& u, float& v) {
    const F3 v0{tri.v0[0], tri.v0[1], tri.v0[2]};
    const F3 e0 = sub(F3{tri.v1[0], tri.v1[1], tri.v1[2]}, v0);
    const F3 e1 = sub(F3{tri.v2[0], tri.v2[1], tri.v2[2]}, v0);
    const F3 s = sub(r.o, v0);
    const F3 p = cross3(s, e0);
    const F3 q = cross3(r.d, e1);
    const float den = dot3(q, e0);
    t = dot3(p, e1) / den;
    u = dot3(q, s) / den;
    v = dot3(p, r.d) / den;
    return t >= 0.0f && u >= 0.0f && v >= 0.0f && u + v <= 1.0f;
}

CpuRay make_ray(const atlas_rt_ray& in) {
    CpuRay r;
    r.o = F3{in.origin[0], in.origin[1], in.origin[2]};
    r.d = F3{in.direction[0], in.direction[1], in.direction[2]};
    r.inv = F3{1.0f / r.d.x, 1.0f / r.d.y, 1.0f / r.d.z};
    r.tmin = in.tmin;
    r.tmax = in.tmax;
    return r;
}

// The root entry: node 0, or the single leaf of a one-triangle BLAS.
inline bool root(const atlas_rt_blas& b, int32_t& ptr) {
    if (!b.nodes.empty()) {
        ptr = 0;
        return true;
    }
    ptr = ~0;
    return b.tris.size() == 1;
}

void closest_one(const atlas_rt_blas& b, const atlas_rt_ray& in, atlas_rt_hit& out,
                 std::vector<std::pair<int32_t, float>>& stack) {
    CpuRay r = make_ray(in);
    float best_t = r.tmax, best_u = 0.0f, best_v = 0.0f;
    int32_t best = -1;
    int32_t start;
    if (root(b, start)) {
        stack.clear();
        stack.emplace_back(start, 0.0f);
        while (!stack.empty()) {
            const auto [ptr, dist] = stack.back();
            stack.pop_back();
            if (dist > best_t) continue;
            if (ptr < 0) {
                for (int32_t k = ~ptr;; k++) {
                    const atlas_rt_gpu_bvh_triangle& tri = b.tris[size_t(k)];
                    float t, u, v;
                    if (triangle(r, tri, t, u, v) && t < best_t) {
                        best = k;
                        best_t = t;
                        best_u = u;
                        best_v = v;
                    }
                    if (tri.v0[3] > 0.0f) break;  // end of leaf
                }
                continue;
            }
            r.tmax = best_t;
            const atlas_rt_gpu_bvh_node& n = b.nodes[size_t(ptr)];
            float hl = 0.0f, hr = 0.0f;
            const bool il = slab(r, n.left_min, n.left_max, hl);
            const bool ir = slab(r, n.right_min, n.right_max, hr);
            if (il && ir) {
                if (hl < hr) {
                    stack.emplace_back(n.right_ptr, hr);
                    stack.emplace_back(n.left_ptr, hl);
                } else {
                    stack.emplace_back(n.left_ptr, hl);
                    stack.emplace_back(n.right_ptr, hr);
                }
            } else if (ir) {
                stack.emplace_back(n.right_ptr, hr);
            } else if (il) {
                stack.emplace_back(n.left_ptr, hl);
            }
        }
    }
    if (best >= 0 && best_t < in.tmax) {
        out.t = best_t;
        out.u = best_u;
        out.v = best_v;
        out.prim_id = int32_t(b.orig_idx[size_t(best)]);
    } else {
        out.t = in.tmax;
        out.u = out.v = 0.0f;
        out.prim_id = -1;
    }
}

bool any_one(const atlas_rt_blas& b, const atlas_rt_ray& in, std::vector<int32_t>& stack) {
    const CpuRay r = make_ray(in);
    int32_t start;
    if (!root(b, start)) return false;
    stack.clear();
    stack.push_back(start);
    while (!stack.empty()) {
        const int32_t ptr = stack.back();
        stack.pop_back();
        if (ptr < 0) {
            for (int32_t k = ~ptr;; k++) {
                const atlas_rt_gpu_bvh_triangle& tri = b.tris[size_t(k)];
                float t, u, v;
                if (triangle(r, tri, t, u, v) && t < r.tmax) return true;
                if (tri.v0[3] > 0.0f) break;
            }
            continue;
        }
        const atlas_rt_gpu_bvh_node& n = b.nodes[size_t(ptr)];
        float t;
        const bool il = slab(r, n.left_min, n.left_max, t);
        const bool ir = slab(r, n.right_min, n.right_max, t);
        if (ir) stack.push_back(n.right_ptr);
        if (il) stack.push_back(n.left_ptr);
    }
    return false;
}

// Contiguous chunks of rays over the host's hardware threads (a thread
// each, from 4,096 rays per thread up), each with its own stack.
template <class Fn>
void for_chunks(uint32_t count, Fn fn) {
    const uint32_t hw = std::max(1u, std::thread::hardware_concurrency());
    const uint32_t threads = std::min(hw, std::max(1u, count / 4096u));
    if (threads <= 1) {
        fn(0u, count);
        return;
    }
    std::vector<std::thread> pool;
    const uint32_t per = (count + threads - 1) / threads;
    for (uint32_t t = 0; t < threads; t++) {
        const uint32_t a = t * per, e = std::min(count, a + per);
        if (a < e) pool.emplace_back(fn, a, e);
    }
    for (std::thread& th : pool) th.join();
}

}  // namespace

void host_trace_closest(const atlas_rt_blas& b, const atlas_rt_ray* rays, atlas_rt_hit* hits, uint32_t count) {
    for_chunks(count, [&](uint32_t a, uint32_t e) {
        std::vector<std::pair<int32_t, float>> stack;
        stack.reserve(64);
        for (uint32_t i = a; i < e; i++) closest_one(b, rays[i], hits[i], stack);
    });
}

void host_trace_any(const atlas_rt_blas& b, const atlas_rt_ray* rays, uint32_t* hit, uint32_t count) {
    for_chunks(count, [&](uint32_t a, uint32_t e) {
        std::vector<int32_t> stack;
        stack.reserve(64);
        for (uint32_t i = a; i < e; i++) hit[i] = any_one(b, rays[i], stack) ? 1u : 0u;
    });
}

}  // namespace atlas_rt
