// Drop-in C++ facade for the engine's Volume::BVH family
// (src/engine/volume/BVH.h:14-138, AABB.h:15-105, Ray.h:21-51), backed by
// the atlas_rt C ABI (include/atlas_rt.h).  Header-only; link libatlas_rt.so.
//
// Signatures follow the engine so existing call sites compile unchanged:
//   BVH(const std::vector<AABB>&, const std::vector<BVHTriangle>&, bool parallelBuild = true)
//   BVH(const std::vector<AABB>&, bool parallelBuild = true)
//   bool GetIntersection(std::vector<std::pair<int32_t,float>>& stack, Ray, BVHTriangle& closest, vec3& intersection)
//   bool GetIntersectionAny(std::vector<std::pair<int32_t,float>>& stack, Ray)
//   std::vector<BVHNode>& GetTree(); public aabbs, data, refs, nodes.
// Behavioural notes:
//  * The trees are built by the library's builder, bit-identical to the
//    engine's BVHBuilder; `nodes`, `data` (BVH order, endOfNode set) and
//    `refs` hold what the engine's members would hold.
//  * GetIntersection / GetIntersectionAny run the engine's own CPU traversal
//    (BVH.cpp:103-217: float-reciprocal slab test, distance-culled stack,
//    triangles accepted from t = 0) on the host through
//    atlas_rt_trace_closest_host / _any_host, also for host-only builds; the
//    GPU kernels (atlas_rt_trace_closest_rays) have the shaders' HitClosest
//    semantics instead.  The stack argument is accepted for source
//    compatibility and unused.
//  * A size mismatch between aabbs and data leaves the BVH empty, as in the
//    engine (BVH.cpp:18-19); status() reports why.
//  * vec3 is glm::vec3 when ATLAS_RT_FACADE_USE_GLM is defined (the engine
//    build), else a minimal struct with x, y, z.
#pragma once

#include <cstdint>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <unordered_map>
#include <utility>
#include <vector>

#include "atlas_rt.h"

#ifdef ATLAS_RT_FACADE_USE_GLM
#include <glm/glm.hpp>
#endif

namespace Atlas {

#ifdef ATLAS_RT_FACADE_USE_GLM
using vec3 = glm::vec3;
#else
struct vec3 {
    float x = 0.0f, y = 0.0f, z = 0.0f;
    vec3() = default;
    explicit vec3(float s) : x(s), y(s), z(s) {}
    vec3(float a, float b, float c) : x(a), y(b), z(c) {}
    float& operator[](int i) { return i == 0 ? x : (i == 1 ? y : z); }
    float operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
};
#endif

namespace Volume {

class AABB {
  public:
    AABB() = default;
    AABB(vec3 min, vec3 max) : min(min), max(max) {}
    float GetSurfaceArea() const {
        const float dx = max.x - min.x, dy = max.y - min.y, dz = max.z - min.z;
        return 2.0f * (dx * dy + dy * dz + dz * dx);
    }
    vec3 min = vec3(0.0f);
    vec3 max = vec3(0.0f);
};

class Ray {
  public:
    Ray() = default;
    Ray(vec3 origin, vec3 direction, float tMin = 0.0f, float tMax = 2048.0f)
        : origin(origin), direction(direction), tMin(tMin), tMax(tMax) {}
    vec3 origin = vec3(0.0f);
    vec3 direction = vec3(0.0f, 1.0f, 0.0f);
    float tMin = 0.0f;
    float tMax = 2048.0f;
};

class BVHNode {
  public:
    AABB leftAABB;
    AABB rightAABB;
    int32_t leftPtr = 0;
    int32_t rightPtr = 0;
};

class BVHTriangle {
  public:
    vec3 v0;
    vec3 v1;
    vec3 v2;
    uint32_t idx = 0;
    bool endOfNode = false;
};

class BVH {
  public:
    struct Ref {
        uint32_t idx = 0;
        uint32_t nodeIdx = 0;
        bool endOfNode = false;
        AABB aabb;
    };

    BVH() = default;

    BVH(const std::vector<AABB>& aabbs, const std::vector<BVHTriangle>& data, bool parallelBuild = true) {
        if (aabbs.size() != data.size() || data.empty()) {
            status_ = aabbs.size() != data.size() ? ATLAS_RT_E_SIZE_MISMATCH : ATLAS_RT_E_EMPTY;
            return;
        }
        std::vector<atlas_rt_aabb> boxes(aabbs.size());
        std::vector<atlas_rt_bvh_triangle> tris(data.size());
        for (size_t i = 0; i < data.size(); i++) {
            put3(boxes[i].min, aabbs[i].min);
            put3(boxes[i].max, aabbs[i].max);
            put3(tris[i].v0, data[i].v0);
            put3(tris[i].v1, data[i].v1);
            put3(tris[i].v2, data[i].v2);
            tris[i].idx = uint32_t(i);
        }
        atlas_rt_blas* h = nullptr;
        const bool device = atlas_rt_device_count() > 0;
        const uint32_t flags = (parallelBuild ? ATLAS_RT_BUILD_PARALLEL : 0u) | (device ? 0u : ATLAS_RT_BUILD_HOST_ONLY);
        status_ = atlas_rt_blas_build(boxes.data(), tris.data(), uint32_t(tris.size()), flags, &h);
        if (status_ != ATLAS_RT_OK) return;
        blas_.reset(h);
        uint32_t nn = 0, nt = 0;
        atlas_rt_blas_info(h, &nn, &nt, nullptr);
        std::vector<atlas_rt_gpu_bvh_node> gn(nn);
        std::vector<atlas_rt_gpu_bvh_triangle> gt(nt);
        std::vector<uint32_t> order(nt);
        atlas_rt_blas_export(h, gn.data(), gt.data(), order.data());
        unpack_nodes(gn);
        this->aabbs.resize(nt);
        this->data.resize(nt);
        refs.resize(nt);
        for (uint32_t i = 0; i < nt; i++) {
            const uint32_t src = order[i];
            this->aabbs[i] = aabbs[src];
            this->data[i] = data[src];
            this->data[i].endOfNode = gt[i].v0[3] > 0.0f;
            refs[i].idx = src;
            refs[i].endOfNode = this->data[i].endOfNode;
            refs[i].aabb = aabbs[src];
            if (!slot_of_.count(src)) slot_of_[src] = i;  // hits name the input position
        }
    }

    // AABB-only tree (BVH.cpp:58-101), as RayTracingWorld builds its TLAS.
    // Produces nodes, refs (with the single-item duplicate) and the reordered
    // aabbs; tracing goes through a world (instances), not this object.
    explicit BVH(const std::vector<AABB>& aabbs, bool parallelBuild = true) {
        if (aabbs.empty()) {
            status_ = ATLAS_RT_E_EMPTY;
            return;
        }
        std::vector<atlas_rt_aabb> boxes(aabbs.size());
        for (size_t i = 0; i < aabbs.size(); i++) {
            put3(boxes[i].min, aabbs[i].min);
            put3(boxes[i].max, aabbs[i].max);
        }
        uint32_t nn = uint32_t(aabbs.size()), nr = uint32_t(aabbs.size()) + 1;
        std::vector<atlas_rt_gpu_bvh_node> gn(nn);
        std::vector<uint32_t> ref(nr);
        std::vector<uint8_t> eon(nr);
        status_ = atlas_rt_aabb_tree_build(boxes.data(), uint32_t(boxes.size()),
                                           parallelBuild ? ATLAS_RT_BUILD_PARALLEL : 0u, gn.data(), &nn, ref.data(),
                                           eon.data(), &nr);
        if (status_ != ATLAS_RT_OK) return;
        gn.resize(nn);
        unpack_nodes(gn);
        refs.resize(nr);
        this->aabbs.resize(nr);
        for (uint32_t i = 0; i < nr; i++) {
            refs[i].idx = ref[i];
            refs[i].endOfNode = eon[i] != 0;
            refs[i].aabb = aabbs[ref[i]];
            this->aabbs[i] = aabbs[ref[i]];
        }
    }

    bool GetIntersection(std::vector<std::pair<int32_t, float>>& stack, Ray ray, BVHTriangle& closest,
                         vec3& intersection) {
        (void)stack;
        intersection.x = ray.tMax;
        if (!blas_) return false;
        atlas_rt_ray r = to_ray(ray);
        atlas_rt_hit h{};
        if (atlas_rt_trace_closest_host(blas_.get(), &r, &h, 1) != ATLAS_RT_OK)
            throw std::runtime_error(atlas_rt_last_error());
        if (h.prim_id < 0) return false;
        closest = data[slot_of_.at(uint32_t(h.prim_id))];
        intersection = vec3(h.t, h.u, h.v);
        return true;
    }

    bool GetIntersectionAny(std::vector<std::pair<int32_t, float>>& stack, Ray ray) {
        (void)stack;
        if (!blas_) return false;
        atlas_rt_ray r = to_ray(ray);
        uint32_t hit = 0;
        if (atlas_rt_trace_any_host(blas_.get(), &r, &hit, 1) != ATLAS_RT_OK)
            throw std::runtime_error(atlas_rt_last_error());
        return hit != 0;
    }

    // Batched form of GetIntersection (the host's hardware threads share the rays).
    std::vector<atlas_rt_hit> GetIntersections(const std::vector<Ray>& rays) {
        std::vector<atlas_rt_ray> r(rays.size());
        for (size_t i = 0; i < rays.size(); i++) r[i] = to_ray(rays[i]);
        std::vector<atlas_rt_hit> h(rays.size());
        if (blas_ && atlas_rt_trace_closest_host(blas_.get(), r.data(), h.data(), uint32_t(r.size())) != ATLAS_RT_OK)
            throw std::runtime_error(atlas_rt_last_error());
        return h;
    }

    std::vector<BVHNode>& GetTree() { return nodes; }

    // Declared by the engine (BVH.h:130) but never defined there.
    void Clear() {
        blas_.reset();
        aabbs.clear();
        data.clear();
        refs.clear();
        nodes.clear();
        slot_of_.clear();
    }

    atlas_rt_status status() const { return status_; }
    const atlas_rt_blas* handle() const { return blas_.get(); }

    std::vector<AABB> aabbs;
    std::vector<BVHTriangle> data;
    std::vector<Ref> refs;
    std::vector<BVHNode> nodes;

  private:
    struct Deleter {
        void operator()(atlas_rt_blas* b) const { atlas_rt_blas_destroy(b); }
    };

    static void put3(float* dst, const vec3& v) {
        dst[0] = v.x;
        dst[1] = v.y;
        dst[2] = v.z;
    }
    static atlas_rt_ray to_ray(const Ray& ray) {
        atlas_rt_ray r;
        put3(r.origin, ray.origin);
        put3(r.direction, ray.direction);
        r.tmin = ray.tMin;
        r.tmax = ray.tMax;
        return r;
    }
    void unpack_nodes(const std::vector<atlas_rt_gpu_bvh_node>& gn) {
        nodes.resize(gn.size());
        for (size_t i = 0; i < gn.size(); i++) {
            nodes[i].leftAABB = AABB(vec3(gn[i].left_min[0], gn[i].left_min[1], gn[i].left_min[2]),
                                     vec3(gn[i].left_max[0], gn[i].left_max[1], gn[i].left_max[2]));
            nodes[i].rightAABB = AABB(vec3(gn[i].right_min[0], gn[i].right_min[1], gn[i].right_min[2]),
                                      vec3(gn[i].right_max[0], gn[i].right_max[1], gn[i].right_max[2]));
            nodes[i].leftPtr = gn[i].left_ptr;
            nodes[i].rightPtr = gn[i].right_ptr;
        }
    }

    std::unique_ptr<atlas_rt_blas, Deleter> blas_;
    std::unordered_map<uint32_t, uint32_t> slot_of_;  // input position -> first BVH slot holding it
    atlas_rt_status status_ = ATLAS_RT_OK;
};

}  // namespace Volume
}  // namespace Atlas
