// Drop-in check of the Volume::BVH facade: engine-style call sites compile
// unchanged against atlas_rt_volume.h.  GetIntersection (the engine's CPU
// traversal, on the host) must agree with the batched form and with
// GetIntersectionAny, and every hit must be a triangle the ray meets.  Runs
// with and without a device.  Prints "facade ok ...".
#include <cmath>
#include <cstdio>
#include <random>
#include <vector>

#include "atlas_rt_volume.h"

using Atlas::vec3;
using namespace Atlas::Volume;

int main() {
    std::mt19937 rng(7);
    std::uniform_real_distribution<float> U(-5.0f, 5.0f), E(0.1f, 0.6f);
    std::vector<BVHTriangle> tris(3000);
    std::vector<AABB> aabbs(tris.size());
    for (uint32_t i = 0; i < tris.size(); i++) {
        vec3 c(U(rng), U(rng), U(rng));
        tris[i].v0 = c;
        tris[i].v1 = vec3(c.x + E(rng), c.y, c.z + E(rng) * 0.3f);
        tris[i].v2 = vec3(c.x, c.y + E(rng), c.z - E(rng) * 0.3f);
        tris[i].idx = i;
        vec3 lo(std::fmin(std::fmin(tris[i].v0.x, tris[i].v1.x), tris[i].v2.x),
                std::fmin(std::fmin(tris[i].v0.y, tris[i].v1.y), tris[i].v2.y),
                std::fmin(std::fmin(tris[i].v0.z, tris[i].v1.z), tris[i].v2.z));
        vec3 hi(std::fmax(std::fmax(tris[i].v0.x, tris[i].v1.x), tris[i].v2.x),
                std::fmax(std::fmax(tris[i].v0.y, tris[i].v1.y), tris[i].v2.y),
                std::fmax(std::fmax(tris[i].v0.z, tris[i].v1.z), tris[i].v2.z));
        aabbs[i] = AABB(lo, hi);
    }
    BVH bvh(aabbs, tris, true);
    if (bvh.status() != ATLAS_RT_OK) {
        std::printf("build failed: %s\n", atlas_rt_last_error());
        return 1;
    }
    // Structure as the engine keeps it: nodes = refs - 1, pre-order children.
    if (bvh.GetTree().size() + 1 != bvh.data.size()) return 2;
    for (size_t i = 0; i < bvh.nodes.size(); i++)
        if ((bvh.nodes[i].leftPtr >= 0 && size_t(bvh.nodes[i].leftPtr) <= i) ||
            (bvh.nodes[i].rightPtr >= 0 && size_t(bvh.nodes[i].rightPtr) <= i))
            return 3;
    BVH tlas(std::vector<AABB>(aabbs.begin(), aabbs.begin() + 17));
    if (tlas.status() != ATLAS_RT_OK || tlas.nodes.size() != 16 || tlas.refs.size() != 17) return 4;
    BVH single(std::vector<AABB>(1, aabbs[0]));
    if (single.nodes.size() != 1 || single.refs.size() != 2) return 5;
    BVH mismatch(aabbs, std::vector<BVHTriangle>(3));
    if (mismatch.status() != ATLAS_RT_E_SIZE_MISMATCH || !mismatch.nodes.empty()) return 6;

    // GetIntersection runs the engine's CPU traversal on the host, with or
    // without a device.
    std::vector<Ray> rays;
    std::normal_distribution<float> N(0.0f, 1.0f);
    for (int i = 0; i < 512; i++) {
        vec3 d(N(rng), N(rng), N(rng));
        float l = std::sqrt(d.x * d.x + d.y * d.y + d.z * d.z);
        rays.emplace_back(vec3(U(rng), U(rng), U(rng)), vec3(d.x / l, d.y / l, d.z / l), 1e-4f, 1e6f);
    }
    std::vector<atlas_rt_hit> batch = bvh.GetIntersections(rays);
    std::vector<std::pair<int32_t, float>> stack(64);
    int hits = 0;
    for (size_t i = 0; i < rays.size(); i++) {
        BVHTriangle closest;
        vec3 inter;
        bool h = bvh.GetIntersection(stack, rays[i], closest, inter);
        bool any = bvh.GetIntersectionAny(stack, rays[i]);
        if (h != (batch[i].prim_id >= 0) || any != h) return 7;
        if (!h) continue;
        hits++;
        if (closest.idx != uint32_t(batch[i].prim_id) || inter.x != batch[i].t) return 8;
        // the hit point lies on the triangle: p = v0 + u e0 + v e1
        const BVHTriangle& t = tris[closest.idx];
        for (int k = 0; k < 3; k++) {
            float p = rays[i].origin[k] + inter.x * rays[i].direction[k];
            float q = t.v0[k] + inter.y * (t.v1[k] - t.v0[k]) + inter.z * (t.v2[k] - t.v0[k]);
            if (std::fabs(p - q) > 1e-3f) return 9;
        }
    }
    if (hits == 0) return 10;
    std::printf("facade ok%s: %zu nodes, %d/%zu hits\n", atlas_rt_device_count() == 0 ? " (host only)" : "",
                bvh.nodes.size(), hits, rays.size());
    return 0;
}
