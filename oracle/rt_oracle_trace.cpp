// TEST INFRASTRUCTURE - the oracle.  Only tests/, __graft_entry__.smoke() and
// bench.py's cpu_baseline leg may load it, and only as the checker.
//
// Restatement of the engine's traversal semantics, on the engine's packed GPU
// layout (GPUBVHNode 64 B, GPUBVHTriangle 48 B, GPUBVHInstance 64 B):
//  R-GPU: HitClosest / HitAny (data/shader/raytracer/bvh.hsh:191-273, 359-441),
//         CheckLeafClosest / CheckLeaf (bvh.hsh:44-104), CheckInstance
//         (bvh.hsh:172-189), IntersectAABB / IntersectTriangle
//         (intersections.hsh:3-58).  This is what the engine runs.
//  R-CPU: BVH::GetIntersection / GetIntersectionAny (src/engine/volume/BVH.cpp:103-217)
//         with Ray::Intersects (src/engine/volume/Ray.cpp:82-130).
// Extensions (documented): unbounded stack (the shader's STACK_SIZE 32 has no
// overflow check), barycentrics returned from the accepted test, the caller's
// original triangle index returned next to the BVH slot, and a BLAS with no
// inner node (one triangle) starts at leaf slot 0.
#include <cstring>
#include <thread>
#include <vector>
#include <cmath>

#include "../include/atlas_rt.h"
#include "rt_oracle.h"

namespace oracle {

static inline V3 node_lmin(const atlas_rt_gpu_bvh_node& n) { return V3{n.left_min[0], n.left_min[1], n.left_min[2]}; }
static inline V3 node_lmax(const atlas_rt_gpu_bvh_node& n) { return V3{n.left_max[0], n.left_max[1], n.left_max[2]}; }
static inline V3 node_rmin(const atlas_rt_gpu_bvh_node& n) { return V3{n.right_min[0], n.right_min[1], n.right_min[2]}; }
static inline V3 node_rmax(const atlas_rt_gpu_bvh_node& n) { return V3{n.right_max[0], n.right_max[1], n.right_max[2]}; }

// IntersectAABB with dist (intersections.hsh:18-34).
static inline bool slab_gpu(V3 o, V3 d, V3 bmin, V3 bmax, float tmin, float tmax, float& dist) {
    V3 t0 = (bmin - o) / d;
    V3 t1 = (bmax - o) / d;
    V3 ts{smin(t0.x, t1.x), smin(t0.y, t1.y), smin(t0.z, t1.z)};
    V3 tb{smax(t0.x, t1.x), smax(t0.y, t1.y), smax(t0.z, t1.z)};
    float tminf = smax(smax(tmin, ts.x), smax(ts.y, ts.z));
    float tmaxf = smin(smin(tmax, tb.x), smin(tb.y, tb.z));
    bool hit = tminf <= tmaxf;
    dist = hit ? tminf : tmax;
    return hit;
}

// IntersectTriangle (intersections.hsh:36-58) == Ray::Intersects(v0,v1,v2,vec3&) (Ray.cpp:107-130).
static inline bool tri_test(V3 o, V3 d, V3 v0, V3 v1, V3 v2, V3& sol) {
    V3 e0 = v1 - v0;
    V3 e1 = v2 - v0;
    V3 s = o - v0;
    V3 p = cross(s, e0);
    V3 q = cross(d, e1);
    float den = dot(q, e0);
    sol = V3{dot(p, e1), dot(q, s), dot(p, d)} / den;
    return sol.x >= 0.0f && sol.y >= 0.0f && sol.z >= 0.0f && sol.y + sol.z <= 1.0f;
}

static inline V3 tri_v(const atlas_rt_gpu_bvh_triangle& t, int k) {
    const float* p = k == 0 ? t.v0 : (k == 1 ? t.v1 : t.v2);
    return V3{p[0], p[1], p[2]};
}

static inline bool is_nan3(V3 d) { return std::isnan(d.x) || std::isnan(d.y) || std::isnan(d.z); }

// ---------------------------------------------------------------------------
// R-GPU, one BLAS (HitClosest's BLAS-mode loop, bvh.hsh:244-267).
// ---------------------------------------------------------------------------
static void closest_rgpu_blas(const BlasView& b, V3 o, V3 d, float tmin, float tmax, HitRec& h,
                              TraceStats& st, std::vector<int32_t>& stack) {
    h = HitRec();
    h.t = tmax;  // ray.hitDistance = tMax (bvh.hsh:202)
    if (is_nan3(d)) return;
    stack.clear();
    int32_t nodePtr = b.node_count ? 0 : ~0;
    for (;;) {
        if (nodePtr < 0) {
            // CheckLeafClosest(ray, meshPtr, nodePtr, tMin, ray.hitDistance)
            const float leafTmax = h.t;
            int32_t triPtr = ~nodePtr;
            bool eon = false;
            while (!eon) {
                const auto& tri = b.tris[triPtr];
                eon = tri.v0[3] > 0.0f;
                V3 sol;
                st.tri_tests++;
                bool hit = tri_test(o, d, tri_v(tri, 0), tri_v(tri, 1), tri_v(tri, 2), sol);
                if (hit && sol.x > tmin && h.slot >= 0 && sol.x == h.t) {
                    st.tie = true;
                    st.tie_t = sol.x;
                }
                if (hit && sol.x > tmin && sol.x < leafTmax) {
                    if (sol.x < h.t) {
                        h.t = sol.x;
                        h.u = sol.y;
                        h.v = sol.z;
                        h.slot = triPtr;
                    }
                }
                triPtr++;
            }
            if (stack.empty()) break;
            nodePtr = stack.back();
            stack.pop_back();
        } else {
            const auto& n = b.nodes[nodePtr];
            st.node_visits++;
            float hitL = 0.0f, hitR = 0.0f;
            bool iL = slab_gpu(o, d, node_lmin(n), node_lmax(n), tmin, h.t, hitL);
            bool iR = slab_gpu(o, d, node_rmin(n), node_rmax(n), tmin, h.t, hitR);
            bool none = !iL && !iR;
            int32_t next = hitL <= hitR ? n.left_ptr : n.right_ptr;
            if (none) {
                if (stack.empty()) break;
                next = stack.back();
                stack.pop_back();
            }
            if (iR && iL) stack.push_back(hitL <= hitR ? n.right_ptr : n.left_ptr);
            nodePtr = next;
        }
    }
}

// R-GPU any hit, one BLAS (HitAny's BLAS-mode loop, bvh.hsh:411-433).
static bool any_rgpu_blas(const BlasView& b, V3 o, V3 d, float tmin, float tmax, TraceStats& st,
                          std::vector<int32_t>& stack) {
    if (is_nan3(d)) return false;
    stack.clear();
    int32_t nodePtr = b.node_count ? 0 : ~0;
    for (;;) {
        if (nodePtr < 0) {
            int32_t triPtr = ~nodePtr;
            bool eon = false;
            while (!eon) {
                const auto& tri = b.tris[triPtr];
                eon = tri.v0[3] > 0.0f;
                V3 sol;
                st.tri_tests++;
                bool hit = tri_test(o, d, tri_v(tri, 0), tri_v(tri, 1), tri_v(tri, 2), sol);
                if (hit && sol.x > tmin && sol.x < tmax) return true;
                triPtr++;
            }
            if (stack.empty()) return false;
            nodePtr = stack.back();
            stack.pop_back();
        } else {
            const auto& n = b.nodes[nodePtr];
            st.node_visits++;
            float dl, dr;
            bool iL = slab_gpu(o, d, node_lmin(n), node_lmax(n), tmin, tmax, dl);
            bool iR = slab_gpu(o, d, node_rmin(n), node_rmax(n), tmin, tmax, dr);
            int32_t next = iL ? n.left_ptr : n.right_ptr;
            if (!iR && !iL) {
                if (stack.empty()) return false;
                next = stack.back();
                stack.pop_back();
            }
            if (iR && iL) stack.push_back(n.right_ptr);
            nodePtr = next;
        }
    }
}

// ---------------------------------------------------------------------------
// R-GPU two-level (TLAS + instances + BLASes), HitClosest / HitAny with
// CheckInstance (bvh.hsh:172-273, 359-441).  The TLAS/BLAS mode switch uses the
// engine's tlasIndex marker on a single stack.
// ---------------------------------------------------------------------------
static inline float dot4(float ax, float ay, float az, float aw, const float* m) {
    return ((ax * m[0] + ay * m[1]) + az * m[2]) + aw * m[3];
}

static const uint32_t kTlasInvalid = 0xFFFFFFFFu;

void closest_rgpu_world(const World& w, V3 o0, V3 d0, float tmin, float tmax, uint32_t mask,
                        HitRec& h, TraceStats& st, std::vector<int32_t>& stack, const Opacity* op) {
    h = HitRec();
    h.t = tmax;
    if (is_nan3(d0)) return;
    stack.clear();
    V3 o = o0, d = d0;
    int32_t nodePtr = 0;
    int32_t meshPtr = 0;
    int32_t curInstance = -1;
    uint32_t tlasIndex = kTlasInvalid;
    // stack.size() + 1 plays the role of the engine's stackPtr.
    auto pop = [&](int32_t& np) -> bool {
        if (stack.empty()) return false;
        np = stack.back();
        stack.pop_back();
        return true;
    };
    for (;;) {
        uint32_t sp = uint32_t(stack.size()) + 1;
        if (sp < tlasIndex) {
            o = o0;
            d = d0;
            tlasIndex = kTlasInvalid;
            if (nodePtr < 0) {
                const auto& in = w.inst[~nodePtr];
                st.inst_visits++;
                o = V3{dot4(o0.x, o0.y, o0.z, 1.0f, in.inverse_matrix[0]), dot4(o0.x, o0.y, o0.z, 1.0f, in.inverse_matrix[1]),
                       dot4(o0.x, o0.y, o0.z, 1.0f, in.inverse_matrix[2])};
                d = V3{dot4(d0.x, d0.y, d0.z, 0.0f, in.inverse_matrix[0]), dot4(d0.x, d0.y, d0.z, 0.0f, in.inverse_matrix[1]),
                       dot4(d0.x, d0.y, d0.z, 0.0f, in.inverse_matrix[2])};
                curInstance = ~nodePtr;
                meshPtr = in.mesh_offset;
                nodePtr = w.blas[meshPtr].node_count ? 0 : ~0;
                if ((in.mask & mask) > 0)
                    tlasIndex = sp;
                else if (!pop(nodePtr))
                    break;
            } else {
                const auto& n = w.tlas[nodePtr];
                st.node_visits++;
                float hitL = 0.0f, hitR = 0.0f;
                bool iL = slab_gpu(o, d, node_lmin(n), node_lmax(n), tmin, h.t, hitL);
                bool iR = slab_gpu(o, d, node_rmin(n), node_rmax(n), tmin, h.t, hitR);
                int32_t next = hitL <= hitR ? n.left_ptr : n.right_ptr;
                if (!iL && !iR) {
                    if (!pop(next)) break;
                }
                if (iR && iL) stack.push_back(hitL <= hitR ? n.right_ptr : n.left_ptr);
                nodePtr = next;
            }
        } else {
            const BlasView& b = w.blas[meshPtr];
            if (nodePtr < 0) {
                const float leafTmax = h.t;
                int32_t triPtr = ~nodePtr;
                bool eon = false;
                while (!eon) {
                    const auto& tri = b.tris[triPtr];
                    eon = tri.v0[3] > 0.0f;
                    V3 sol;
                    st.tri_tests++;
                    bool hit = tri_test(o, d, tri_v(tri, 0), tri_v(tri, 1), tri_v(tri, 2), sol);
                    if (hit && sol.x > tmin && h.slot >= 0 && sol.x == h.t &&
                        (!op || op->fn(op->ctx, curInstance, triPtr, sol.y, sol.z) > 0.0f)) {
                        st.tie = true;
                        st.tie_t = sol.x;
                    }
                    if (hit && sol.x > tmin && sol.x < leafTmax && sol.x < h.t &&
                        (!op || op->fn(op->ctx, curInstance, triPtr, sol.y, sol.z) > 0.0f)) {
                        h.t = sol.x;
                        h.u = sol.y;
                        h.v = sol.z;
                        h.slot = triPtr;
                        h.instance = curInstance;
                        h.ho[0] = o.x; h.ho[1] = o.y; h.ho[2] = o.z;
                        h.hd[0] = d.x; h.hd[1] = d.y; h.hd[2] = d.z;
                    }
                    triPtr++;
                }
                if (!pop(nodePtr)) break;
            } else {
                const auto& n = b.nodes[nodePtr];
                st.node_visits++;
                float hitL = 0.0f, hitR = 0.0f;
                bool iL = slab_gpu(o, d, node_lmin(n), node_lmax(n), tmin, h.t, hitL);
                bool iR = slab_gpu(o, d, node_rmin(n), node_rmax(n), tmin, h.t, hitR);
                int32_t next = hitL <= hitR ? n.left_ptr : n.right_ptr;
                if (!iL && !iR) {
                    if (!pop(next)) break;
                }
                if (iR && iL) stack.push_back(hitL <= hitR ? n.right_ptr : n.left_ptr);
                nodePtr = next;
            }
        }
    }
}

// restore_always = false: HitAny (bvh.hsh:382-387) restores the world ray only
// when returning from an entered instance, so after an instance rejected by
// the mask the TLAS walk continues with that instance's object-space ray.
// restore_always = true: HitAnyTransparency's traversal (bvh.hsh:466-470),
// which restores on every TLAS step; with opaque triangles its result is this
// any-hit test.
bool any_rgpu_world(const World& w, V3 o0, V3 d0, float tmin, float tmax, uint32_t mask, TraceStats& st,
                    std::vector<int32_t>& stack, bool restore_always) {
    if (is_nan3(d0)) return false;
    stack.clear();
    V3 o = o0, d = d0;
    int32_t nodePtr = 0;
    int32_t meshPtr = 0;
    uint32_t tlasIndex = kTlasInvalid;
    auto pop = [&](int32_t& np) -> bool {
        if (stack.empty()) return false;
        np = stack.back();
        stack.pop_back();
        return true;
    };
    for (;;) {
        uint32_t sp = uint32_t(stack.size()) + 1;
        if (sp < tlasIndex) {
            if (restore_always || tlasIndex != kTlasInvalid) {
                o = o0;
                d = d0;
            }
            tlasIndex = kTlasInvalid;
            if (nodePtr < 0) {
                const auto& in = w.inst[~nodePtr];
                st.inst_visits++;
                o = V3{dot4(o0.x, o0.y, o0.z, 1.0f, in.inverse_matrix[0]), dot4(o0.x, o0.y, o0.z, 1.0f, in.inverse_matrix[1]),
                       dot4(o0.x, o0.y, o0.z, 1.0f, in.inverse_matrix[2])};
                d = V3{dot4(d0.x, d0.y, d0.z, 0.0f, in.inverse_matrix[0]), dot4(d0.x, d0.y, d0.z, 0.0f, in.inverse_matrix[1]),
                       dot4(d0.x, d0.y, d0.z, 0.0f, in.inverse_matrix[2])};
                meshPtr = in.mesh_offset;
                nodePtr = w.blas[meshPtr].node_count ? 0 : ~0;
                if ((in.mask & mask) > 0)
                    tlasIndex = sp;
                else if (!pop(nodePtr))
                    return false;
            } else {
                const auto& n = w.tlas[nodePtr];
                st.node_visits++;
                float dl, dr;
                bool iL = slab_gpu(o, d, node_lmin(n), node_lmax(n), tmin, tmax, dl);
                bool iR = slab_gpu(o, d, node_rmin(n), node_rmax(n), tmin, tmax, dr);
                int32_t next = iL ? n.left_ptr : n.right_ptr;
                if (!iR && !iL) {
                    if (!pop(next)) return false;
                }
                if (iR && iL) stack.push_back(n.right_ptr);
                nodePtr = next;
            }
        } else {
            const BlasView& b = w.blas[meshPtr];
            if (nodePtr < 0) {
                int32_t triPtr = ~nodePtr;
                bool eon = false;
                while (!eon) {
                    const auto& tri = b.tris[triPtr];
                    eon = tri.v0[3] > 0.0f;
                    V3 sol;
                    st.tri_tests++;
                    bool hit = tri_test(o, d, tri_v(tri, 0), tri_v(tri, 1), tri_v(tri, 2), sol);
                    if (hit && sol.x > tmin && sol.x < tmax) return true;
                    triPtr++;
                }
                if (!pop(nodePtr)) return false;
            } else {
                const auto& n = b.nodes[nodePtr];
                st.node_visits++;
                float dl, dr;
                bool iL = slab_gpu(o, d, node_lmin(n), node_lmax(n), tmin, tmax, dl);
                bool iR = slab_gpu(o, d, node_rmin(n), node_rmax(n), tmin, tmax, dr);
                int32_t next = iL ? n.left_ptr : n.right_ptr;
                if (!iR && !iL) {
                    if (!pop(next)) return false;
                }
                if (iR && iL) stack.push_back(n.right_ptr);
                nodePtr = next;
            }
        }
    }
}

// HitAnyTransparency (bvh.hsh:443-524).  The world ray is restored on every
// TLAS step.  At a leaf, CheckLeafTransparency starts from the current
// transparency and multiplies (1 - opacity) for every triangle the ray hits in
// (tmin, tmax); the caller then multiplies that product into the transparency
// again (bvh.hsh:496), so each leaf squares what came before it.  Below 1e-6
// the transparency becomes 0 and the walk ends.
float transparency_rgpu_world(const World& w, V3 o0, V3 d0, float tmin, float tmax, uint32_t mask,
                              TraceStats& st, std::vector<int32_t>& stack, const Opacity& op) {
    if (is_nan3(d0)) return 1.0f;
    float transparency = 1.0f;
    stack.clear();
    V3 o = o0, d = d0;
    int32_t nodePtr = 0;
    int32_t meshPtr = 0;
    int32_t curInstance = -1;
    uint32_t tlasIndex = kTlasInvalid;
    auto pop = [&](int32_t& np) -> bool {
        if (stack.empty()) return false;
        np = stack.back();
        stack.pop_back();
        return true;
    };
    while (transparency > 0.0f) {
        uint32_t sp = uint32_t(stack.size()) + 1;
        if (sp < tlasIndex) {
            o = o0;
            d = d0;
            tlasIndex = kTlasInvalid;
            if (nodePtr < 0) {
                const auto& in = w.inst[~nodePtr];
                st.inst_visits++;
                o = V3{dot4(o0.x, o0.y, o0.z, 1.0f, in.inverse_matrix[0]), dot4(o0.x, o0.y, o0.z, 1.0f, in.inverse_matrix[1]),
                       dot4(o0.x, o0.y, o0.z, 1.0f, in.inverse_matrix[2])};
                d = V3{dot4(d0.x, d0.y, d0.z, 0.0f, in.inverse_matrix[0]), dot4(d0.x, d0.y, d0.z, 0.0f, in.inverse_matrix[1]),
                       dot4(d0.x, d0.y, d0.z, 0.0f, in.inverse_matrix[2])};
                curInstance = ~nodePtr;
                meshPtr = in.mesh_offset;
                nodePtr = w.blas[meshPtr].node_count ? 0 : ~0;
                if ((in.mask & mask) > 0)
                    tlasIndex = sp;
                else if (!pop(nodePtr))
                    break;
            } else {
                const auto& n = w.tlas[nodePtr];
                st.node_visits++;
                float dl, dr;
                bool iL = slab_gpu(o, d, node_lmin(n), node_lmax(n), tmin, tmax, dl);
                bool iR = slab_gpu(o, d, node_rmin(n), node_rmax(n), tmin, tmax, dr);
                int32_t next = iL ? n.left_ptr : n.right_ptr;
                if (!iR && !iL) {
                    if (!pop(next)) break;
                }
                if (iR && iL) stack.push_back(n.right_ptr);
                nodePtr = next;
            }
        } else {
            const BlasView& b = w.blas[meshPtr];
            if (nodePtr < 0) {
                float leaf = transparency;
                int32_t triPtr = ~nodePtr;
                bool eon = false;
                while (!eon) {
                    const auto& tri = b.tris[triPtr];
                    eon = tri.v0[3] > 0.0f;
                    V3 sol;
                    st.tri_tests++;
                    bool hit = tri_test(o, d, tri_v(tri, 0), tri_v(tri, 1), tri_v(tri, 2), sol);
                    if (hit && sol.x > tmin && sol.x < tmax)
                        leaf *= (1.0f - op.fn(op.ctx, curInstance, triPtr, sol.y, sol.z));
                    triPtr++;
                }
                transparency *= leaf;
                if (transparency < 0.000001f) transparency = 0.0f;
                if (!pop(nodePtr)) break;
            } else {
                const auto& n = b.nodes[nodePtr];
                st.node_visits++;
                float dl, dr;
                bool iL = slab_gpu(o, d, node_lmin(n), node_lmax(n), tmin, tmax, dl);
                bool iR = slab_gpu(o, d, node_rmin(n), node_rmax(n), tmin, tmax, dr);
                int32_t next = iL ? n.left_ptr : n.right_ptr;
                if (!iR && !iL) {
                    if (!pop(next)) break;
                }
                if (iR && iL) stack.push_back(n.right_ptr);
                nodePtr = next;
            }
        }
    }
    return transparency;
}

// ---------------------------------------------------------------------------
// R-CPU (BVH::GetIntersection / GetIntersectionAny, BVH.cpp:103-217).
// ---------------------------------------------------------------------------
struct CpuRay {
    V3 origin, direction, inv;
    float tMin, tMax;
};

// Ray::Intersects(const AABB&, float&) (Ray.cpp:82-97), glm min/max.
static inline bool slab_cpu(const CpuRay& r, V3 bmin, V3 bmax, float& t) {
    V3 t0 = (bmin - r.origin) * r.inv;
    V3 t1 = (bmax - r.origin) * r.inv;
    V3 ts = gmin(t0, t1);
    V3 tb = gmax(t0, t1);
    float tminf = gmax(r.tMin, gmax(ts.x, gmax(ts.y, ts.z)));
    float tmaxf = gmin(r.tMax, gmin(tb.x, gmin(tb.y, tb.z)));
    t = tminf;
    return tminf <= tmaxf;
}

static void closest_rcpu_blas(const BlasView& b, V3 o, V3 d, float tmin, float tmax, HitRec& h,
                              TraceStats& st, std::vector<std::pair<int32_t, float>>& stack) {
    CpuRay ray{o, d, V3{1.0f / d.x, 1.0f / d.y, 1.0f / d.z}, tmin, tmax};
    h = HitRec();
    V3 inter{ray.tMax, 0.0f, 0.0f};
    int32_t closestPtr = -1;
    if (b.node_count || b.tri_count == 1) {
        stack.clear();
        stack.push_back({b.node_count ? 0 : ~0, 0.0f});
        while (!stack.empty()) {
            auto [nodePtr, distance] = stack.back();
            stack.pop_back();
            if (distance > inter.x) continue;
            if (nodePtr < 0) {
                int32_t triPtr = ~nodePtr;
                bool eon = false;
                while (!eon) {
                    int32_t ptr = triPtr++;
                    const auto& tri = b.tris[ptr];
                    eon = tri.v0[3] > 0.0f;
                    V3 sol;
                    st.tri_tests++;
                    bool hit = tri_test(ray.origin, ray.direction, tri_v(tri, 0), tri_v(tri, 1), tri_v(tri, 2), sol);
                    if (hit && sol.x < inter.x) {
                        closestPtr = ptr;
                        inter = sol;
                    }
                }
            } else {
                float hitL = 0.0f, hitR = 0.0f;
                ray.tMax = inter.x;
                const auto& n = b.nodes[nodePtr];
                st.node_visits++;
                bool iL = slab_cpu(ray, node_lmin(n), node_lmax(n), hitL);
                bool iR = slab_cpu(ray, node_rmin(n), node_rmax(n), hitR);
                if (iR && iL) {
                    if (hitL < hitR) {
                        stack.push_back({n.right_ptr, hitR});
                        stack.push_back({n.left_ptr, hitL});
                    } else {
                        stack.push_back({n.left_ptr, hitL});
                        stack.push_back({n.right_ptr, hitR});
                    }
                } else if (iR && !iL) {
                    stack.push_back({n.right_ptr, hitR});
                } else if (!iR && iL) {
                    stack.push_back({n.left_ptr, hitL});
                }
            }
        }
    }
    // The engine's return value compares against the mutated ray.tMax
    // (BVH.cpp:141,172); a hit is reported here when t < the caller's tMax.
    h.t = inter.x;
    if (closestPtr >= 0 && inter.x < tmax) {
        h.u = inter.y;
        h.v = inter.z;
        h.slot = closestPtr;
    } else {
        h.t = tmax;
    }
}

static bool any_rcpu_blas(const BlasView& b, V3 o, V3 d, float tmin, float tmax, TraceStats& st,
                          std::vector<std::pair<int32_t, float>>& stack) {
    CpuRay ray{o, d, V3{1.0f / d.x, 1.0f / d.y, 1.0f / d.z}, tmin, tmax};
    if (!(b.node_count || b.tri_count == 1)) return false;
    stack.clear();
    stack.push_back({b.node_count ? 0 : ~0, 0.0f});
    while (!stack.empty()) {
        int32_t nodePtr = stack.back().first;
        stack.pop_back();
        if (nodePtr < 0) {
            int32_t triPtr = ~nodePtr;
            bool eon = false;
            while (!eon) {
                const auto& tri = b.tris[triPtr++];
                eon = tri.v0[3] > 0.0f;
                V3 sol;
                st.tri_tests++;
                bool hit = tri_test(ray.origin, ray.direction, tri_v(tri, 0), tri_v(tri, 1), tri_v(tri, 2), sol);
                if (hit && sol.x < ray.tMax) return true;
            }
        } else {
            const auto& n = b.nodes[nodePtr];
            st.node_visits++;
            float t;
            bool iL = slab_cpu(ray, node_lmin(n), node_lmax(n), t);
            bool iR = slab_cpu(ray, node_rmin(n), node_rmax(n), t);
            if (iR) stack.push_back({n.right_ptr, 0.0f});
            if (iL) stack.push_back({n.left_ptr, 0.0f});
        }
    }
    return false;
}

// Brute force over every instance (TLAS slot order) and every triangle of its
// BLAS (slot order) in the instance's object space, with CheckInstance's ray
// transform and mask test: accept t > tmin, t < best.
static void closest_brute_world(const World& w, V3 o0, V3 d0, float tmin, float tmax, uint32_t mask, HitRec& h) {
    h = HitRec();
    h.t = tmax;
    if (is_nan3(d0)) return;
    for (uint32_t k = 0; k < w.inst_count; k++) {
        const auto& in = w.inst[k];
        if (!((in.mask & mask) > 0)) continue;
        V3 o{dot4(o0.x, o0.y, o0.z, 1.0f, in.inverse_matrix[0]), dot4(o0.x, o0.y, o0.z, 1.0f, in.inverse_matrix[1]),
             dot4(o0.x, o0.y, o0.z, 1.0f, in.inverse_matrix[2])};
        V3 d{dot4(d0.x, d0.y, d0.z, 0.0f, in.inverse_matrix[0]), dot4(d0.x, d0.y, d0.z, 0.0f, in.inverse_matrix[1]),
             dot4(d0.x, d0.y, d0.z, 0.0f, in.inverse_matrix[2])};
        const BlasView& b = w.blas[in.mesh_offset];
        for (uint32_t i = 0; i < b.tri_count; i++) {
            const auto& tri = b.tris[i];
            V3 sol;
            bool hit = tri_test(o, d, tri_v(tri, 0), tri_v(tri, 1), tri_v(tri, 2), sol);
            if (hit && sol.x > tmin && sol.x < h.t) {
                h.t = sol.x;
                h.u = sol.y;
                h.v = sol.z;
                h.slot = int32_t(i);
                h.instance = int32_t(k);
                h.ho[0] = o.x; h.ho[1] = o.y; h.ho[2] = o.z;
                h.hd[0] = d.x; h.hd[1] = d.y; h.hd[2] = d.z;
            }
        }
    }
}

// Brute force over every triangle in slot order: accept t > tmin, t < best.
static void closest_brute(const BlasView& b, V3 o, V3 d, float tmin, float tmax, HitRec& h) {
    h = HitRec();
    h.t = tmax;
    if (is_nan3(d)) return;
    for (uint32_t i = 0; i < b.tri_count; i++) {
        const auto& tri = b.tris[i];
        V3 sol;
        bool hit = tri_test(o, d, tri_v(tri, 0), tri_v(tri, 1), tri_v(tri, 2), sol);
        if (hit && sol.x > tmin && sol.x < h.t) {
            h.t = sol.x;
            h.u = sol.y;
            h.v = sol.z;
            h.slot = int32_t(i);
        }
    }
}

// ---------------------------------------------------------------------------
// Threaded drivers
// ---------------------------------------------------------------------------
template <class F>
static void parallel_rays(uint32_t n, int threads, F&& f) {
    if (threads <= 1 || n < 1024) {
        f(0u, n);
        return;
    }
    std::vector<std::thread> pool;
    uint32_t chunk = (n + uint32_t(threads) - 1) / uint32_t(threads);
    for (int t = 0; t < threads; t++) {
        uint32_t b = uint32_t(t) * chunk, e = std::min(n, b + chunk);
        if (b >= e) break;
        pool.emplace_back([&f, b, e]() { f(b, e); });
    }
    for (auto& th : pool) th.join();
}

}  // namespace oracle

using namespace oracle;

// Reported (t, u, v) of a hit: the engine's triangle formula evaluated in
// double from the float32 vertices and ray (fixed order, no contraction),
// rounded once to float.  Hit selection never uses it.
static inline void refine_tuv(const float* o, const float* dir, const atlas_rt_gpu_bvh_triangle& tri, float& t,
                              float& u, float& v) {
    const double ax = tri.v0[0], ay = tri.v0[1], az = tri.v0[2];
    const double e0x = (double)tri.v1[0] - ax, e0y = (double)tri.v1[1] - ay, e0z = (double)tri.v1[2] - az;
    const double e1x = (double)tri.v2[0] - ax, e1y = (double)tri.v2[1] - ay, e1z = (double)tri.v2[2] - az;
    const double sx = (double)o[0] - ax, sy = (double)o[1] - ay, sz = (double)o[2] - az;
    const double Dx = dir[0], Dy = dir[1], Dz = dir[2];
    const double px = sy * e0z - e0y * sz, py = sz * e0x - e0z * sx, pz = sx * e0y - e0x * sy;
    const double qx = Dy * e1z - e1y * Dz, qy = Dz * e1x - e1z * Dx, qz = Dx * e1y - e1x * Dy;
    const double den = (qx * e0x + qy * e0y) + qz * e0z;
    t = (float)(((px * e1x + py * e1y) + pz * e1z) / den);
    u = (float)(((qx * sx + qy * sy) + qz * sz) / den);
    v = (float)(((px * Dx + py * Dy) + pz * Dz) / den);
}

static inline void write_hit(const HitRec& h, const atlas_rt_gpu_bvh_triangle* tris, const uint32_t* orig,
                             const atlas_rt_ray& r, atlas_rt_hit* out, int32_t* slot_out, float* raw_t, uint32_t i) {
    if (h.slot >= 0) {
        refine_tuv(r.origin, r.direction, tris[h.slot], out[i].t, out[i].u, out[i].v);
        out[i].prim_id = orig ? int32_t(orig[h.slot]) : h.slot;
    } else {
        out[i].t = r.tmax;
        out[i].u = 0.0f;
        out[i].v = 0.0f;
        out[i].prim_id = -1;
    }
    if (slot_out) slot_out[i] = h.slot;
    if (raw_t) raw_t[i] = h.t;
}

extern "C" {

// mode: 0 = R-GPU closest, 1 = R-CPU closest, 2 = brute force.
// hits get the refined (t, u, v); raw_t (may be NULL) gets the engine's float32
// hit distance (tmax on a miss).
void oracle_trace_closest(const atlas_rt_gpu_bvh_node* nodes, uint32_t node_count,
                          const atlas_rt_gpu_bvh_triangle* tris, const uint32_t* orig_idx, uint32_t tri_count,
                          const atlas_rt_ray* rays, uint32_t n, atlas_rt_hit* hits, int32_t* slots, float* raw_t,
                          uint32_t* node_visits, uint32_t* tri_tests, uint8_t* ties, int32_t mode,
                          int32_t threads) {
    BlasView b{nodes, node_count, tris, tri_count};
    parallel_rays(n, threads, [&](uint32_t lo, uint32_t hi) {
        std::vector<int32_t> stack;
        std::vector<std::pair<int32_t, float>> cstack;
        stack.reserve(128);
        cstack.reserve(128);
        for (uint32_t i = lo; i < hi; i++) {
            const atlas_rt_ray& r = rays[i];
            V3 o{r.origin[0], r.origin[1], r.origin[2]};
            V3 d{r.direction[0], r.direction[1], r.direction[2]};
            HitRec h;
            TraceStats st;
            if (mode == 0)
                closest_rgpu_blas(b, o, d, r.tmin, r.tmax, h, st, stack);
            else if (mode == 1)
                closest_rcpu_blas(b, o, d, r.tmin, r.tmax, h, st, cstack);
            else
                closest_brute(b, o, d, r.tmin, r.tmax, h);
            write_hit(h, tris, orig_idx, r, hits, slots, raw_t, i);
            if (node_visits) node_visits[i] = st.node_visits;
            if (tri_tests) tri_tests[i] = st.tri_tests;
            if (ties) ties[i] = (h.slot >= 0 && st.tie && st.tie_t == h.t) ? 1 : 0;
        }
    });
}

// mode: 0 = R-GPU any hit, 1 = R-CPU any hit.
void oracle_trace_any(const atlas_rt_gpu_bvh_node* nodes, uint32_t node_count,
                      const atlas_rt_gpu_bvh_triangle* tris, uint32_t tri_count, const atlas_rt_ray* rays,
                      uint32_t n, uint32_t* hit, uint32_t* node_visits, uint32_t* tri_tests, int32_t mode,
                      int32_t threads) {
    BlasView b{nodes, node_count, tris, tri_count};
    parallel_rays(n, threads, [&](uint32_t lo, uint32_t hi) {
        std::vector<int32_t> stack;
        std::vector<std::pair<int32_t, float>> cstack;
        for (uint32_t i = lo; i < hi; i++) {
            const atlas_rt_ray& r = rays[i];
            V3 o{r.origin[0], r.origin[1], r.origin[2]};
            V3 d{r.direction[0], r.direction[1], r.direction[2]};
            TraceStats st;
            bool h = mode == 0 ? any_rgpu_blas(b, o, d, r.tmin, r.tmax, st, stack)
                               : any_rcpu_blas(b, o, d, r.tmin, r.tmax, st, cstack);
            hit[i] = h ? 1u : 0u;
            if (node_visits) node_visits[i] = st.node_visits;
            if (tri_tests) tri_tests[i] = st.tri_tests;
        }
    });
}

// Two-level trace.  blas_nodes/blas_tris/... are arrays of per-BLAS
// pointers/counts.  mode 0: R-GPU HitClosest, 1: R-GPU HitAny (writes only
// hits[i].prim_id = 1/0), 2: brute force over instances x triangles,
// 3: any hit with HitAnyTransparency's traversal (world ray restored on
// every TLAS step; opaque triangles).
// Closest outputs: hits (refined t, u, v from the object-space ray of the
// accepted test; prim_id = orig index in the hit BLAS), slots (engine hitID),
// instances (engine hitInstanceID = TLAS instance slot), raw_t (engine
// float32 hit distance, tmax on a miss).
void oracle_trace_world(const atlas_rt_gpu_bvh_node* tlas, uint32_t tlas_count,
                        const atlas_rt_gpu_bvh_instance* inst, uint32_t inst_count,
                        const atlas_rt_gpu_bvh_node* const* blas_nodes, const uint32_t* blas_node_counts,
                        const atlas_rt_gpu_bvh_triangle* const* blas_tris, const uint32_t* const* blas_orig,
                        const uint32_t* blas_tri_counts, uint32_t blas_count, const atlas_rt_ray* rays, uint32_t n,
                        uint32_t mask, int32_t mode, atlas_rt_hit* hits, int32_t* slots, int32_t* instances,
                        float* raw_t, uint32_t* node_visits, uint32_t* tri_tests, uint32_t* inst_visits,
                        uint8_t* ties, int32_t threads) {
    const int32_t any = mode == 1 || mode == 3;
    std::vector<BlasView> views(blas_count);
    for (uint32_t k = 0; k < blas_count; k++)
        views[k] = BlasView{blas_nodes[k], blas_node_counts[k], blas_tris[k], blas_tri_counts[k]};
    World w{tlas, tlas_count, inst, inst_count, views.data(), blas_count};
    parallel_rays(n, threads, [&](uint32_t lo, uint32_t hi) {
        std::vector<int32_t> stack;
        for (uint32_t i = lo; i < hi; i++) {
            const atlas_rt_ray& r = rays[i];
            V3 o{r.origin[0], r.origin[1], r.origin[2]};
            V3 d{r.direction[0], r.direction[1], r.direction[2]};
            TraceStats st;
            if (any) {
                bool h = any_rgpu_world(w, o, d, r.tmin, r.tmax, mask, st, stack, mode == 3);
                hits[i].t = 0.0f;
                hits[i].u = 0.0f;
                hits[i].v = 0.0f;
                hits[i].prim_id = h ? 1 : 0;
            } else {
                HitRec h;
                if (mode == 2)
                    closest_brute_world(w, o, d, r.tmin, r.tmax, mask, h);
                else
                    closest_rgpu_world(w, o, d, r.tmin, r.tmax, mask, h, st, stack);
                // t is in world units: the object-space direction is not
                // renormalised (bvh.hsh:177-180).
                if (h.slot >= 0) {
                    const auto& bt = views[inst[h.instance].mesh_offset].tris[h.slot];
                    refine_tuv(h.ho, h.hd, bt, hits[i].t, hits[i].u, hits[i].v);
                    const uint32_t* orig = blas_orig[inst[h.instance].mesh_offset];
                    hits[i].prim_id = orig ? int32_t(orig[h.slot]) : h.slot;
                } else {
                    hits[i].t = r.tmax;
                    hits[i].u = hits[i].v = 0.0f;
                    hits[i].prim_id = -1;
                }
                if (slots) slots[i] = h.slot;
                if (instances) instances[i] = h.instance;
                if (raw_t) raw_t[i] = h.t;
                if (ties) ties[i] = (h.slot >= 0 && st.tie && st.tie_t == h.t) ? 1 : 0;
            }
            if (node_visits) node_visits[i] = st.node_visits;
            if (tri_tests) tri_tests[i] = st.tri_tests;
            if (inst_visits) inst_visits[i] = st.inst_visits;
        }
    });
}

// Single triangle test and slab test, exposed for closed-form unit tests.
int32_t oracle_triangle_test(const float* o, const float* d, const float* v0, const float* v1, const float* v2,
                             float* sol) {
    V3 s;
    bool h = tri_test(V3{o[0], o[1], o[2]}, V3{d[0], d[1], d[2]}, V3{v0[0], v0[1], v0[2]},
                      V3{v1[0], v1[1], v1[2]}, V3{v2[0], v2[1], v2[2]}, s);
    sol[0] = s.x;
    sol[1] = s.y;
    sol[2] = s.z;
    return h ? 1 : 0;
}

int32_t oracle_slab_gpu(const float* o, const float* d, const float* bmin, const float* bmax, float tmin,
                        float tmax, float* dist) {
    return slab_gpu(V3{o[0], o[1], o[2]}, V3{d[0], d[1], d[2]}, V3{bmin[0], bmin[1], bmin[2]},
                    V3{bmax[0], bmax[1], bmax[2]}, tmin, tmax, *dist)
               ? 1
               : 0;
}

}  // extern "C"
