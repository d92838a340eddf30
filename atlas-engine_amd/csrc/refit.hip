// Per-frame TLAS refit on the device, gfx950.
//
// The engine rebuilds its TLAS on the host every frame
// (RayTracingWorld::Update -> UpdateForSoftwareRayTracing,
// RayTracingWorld.cpp:32-162, 267-307): world boxes of the instances
// (AABB::Transform, AABB.cpp:35-60), a new BVH over them, new instance
// records.  A refit keeps the last build's tree and instance order and only
// moves the instances:
//   1. per instance slot: the new matrices into the instance record and the
//      world transform table, and its world box (the BLAS's object box, its 8
//      corners transformed in glm's mat4 * vec4 order);
//   2. per tree level, deepest first: each node's child boxes, the union of
//      the child node's two boxes or of the instance boxes of the child leaf.
// Boxes grow with glm 0.9.8's min/max ((b < a) ? b : a, (a < b) ? b : a), as
// the host builder's boxes do, so refitting with unchanged matrices
// reproduces the built tree's boxes.
#include "world.h"

#pragma clang fp contract(off)

namespace atlas_rt {

namespace {

constexpr int kB = 256;

__device__ __forceinline__ float gl_min(float a, float b) { return (b < a) ? b : a; }
__device__ __forceinline__ float gl_max(float a, float b) { return (a < b) ? b : a; }

struct Box3 {
    float lo[3], hi[3];
};

__device__ __forceinline__ Box3 empty3() {
    Box3 b;
    for (int k = 0; k < 3; k++) {
        b.lo[k] = 3.4028234663852886e+38f;
        b.hi[k] = -3.4028234663852886e+38f;
    }
    return b;
}

__device__ __forceinline__ void grow(Box3& b, const Box3& o) {
    for (int k = 0; k < 3; k++) {
        b.hi[k] = gl_max(b.hi[k], o.hi[k]);
        b.lo[k] = gl_min(b.lo[k], o.lo[k]);
    }
}

__global__ __launch_bounds__(kB) void k_refit_instances(uint32_t n, const atlas_rt_instance_desc* __restrict__ desc,
                                                        const int32_t* __restrict__ order,
                                                        const float4* __restrict__ blas_box, float4* __restrict__ inst,
                                                        float4* __restrict__ xform, float4* __restrict__ ibox) {
    const uint32_t i = blockIdx.x * kB + threadIdx.x;
    if (i >= n) return;
    const atlas_rt_instance_desc& d = desc[order[i]];
    float m[3][4];
    for (int r = 0; r < 3; r++) {
        for (int c = 0; c < 4; c++) m[r][c] = d.object_to_world[r][c];
        xform[3 * size_t(i) + r] = make_float4(m[r][0], m[r][1], m[r][2], m[r][3]);
        inst[4 * size_t(i) + r] = make_float4(d.world_to_object[r][0], d.world_to_object[r][1],
                                              d.world_to_object[r][2], d.world_to_object[r][3]);
    }
    const int32_t mesh = __float_as_int(inst[4 * size_t(i) + 3].x);
    const float4 lo = blas_box[2 * size_t(mesh)], hi = blas_box[2 * size_t(mesh) + 1];
    const float cx[8] = {lo.x, lo.x, hi.x, hi.x, lo.x, lo.x, hi.x, hi.x};
    const float cy[8] = {lo.y, lo.y, lo.y, lo.y, hi.y, hi.y, hi.y, hi.y};
    const float cz[8] = {lo.z, hi.z, lo.z, hi.z, lo.z, hi.z, lo.z, hi.z};
    Box3 b;
    for (int k = 0; k < 8; k++) {
        float p[3];
        for (int r = 0; r < 3; r++) p[r] = (m[r][0] * cx[k] + m[r][1] * cy[k]) + (m[r][2] * cz[k] + m[r][3] * 1.0f);
        for (int r = 0; r < 3; r++) {
            if (k == 0) {
                b.lo[r] = b.hi[r] = p[r];
            } else {
                b.lo[r] = gl_min(b.lo[r], p[r]);
                b.hi[r] = gl_max(b.hi[r], p[r]);
            }
        }
    }
    ibox[2 * size_t(i)] = make_float4(b.lo[0], b.lo[1], b.lo[2], 0.0f);
    ibox[2 * size_t(i) + 1] = make_float4(b.hi[0], b.hi[1], b.hi[2], 0.0f);
}

__device__ __forceinline__ Box3 child_box(const float4* __restrict__ tlas, const float4* __restrict__ ibox, int32_t ptr,
                                          int32_t first, int32_t count) {
    Box3 b = empty3();
    if (count == 0) {  // inner child: the union of its own two child boxes
        const float4* c = tlas + 4 * size_t(ptr);
        const float4 a = c[0], e = c[1], f = c[2];
        const Box3 l{{a.x, a.y, a.z}, {a.w, e.x, e.y}};
        const Box3 r{{e.z, e.w, f.x}, {f.y, f.z, f.w}};
        grow(b, l);
        grow(b, r);
    } else {
        for (int32_t s = first; s < first + count; s++) {
            const float4 lo = ibox[2 * size_t(s)], hi = ibox[2 * size_t(s) + 1];
            const Box3 o{{lo.x, lo.y, lo.z}, {hi.x, hi.y, hi.z}};
            grow(b, o);
        }
    }
    return b;
}

__global__ __launch_bounds__(kB) void k_refit_level(uint32_t count, const uint32_t* __restrict__ nodes,
                                                    const int4* __restrict__ leaves, const float4* __restrict__ ibox,
                                                    float4* tlas) {
    const uint32_t i = blockIdx.x * kB + threadIdx.x;
    if (i >= count) return;
    const uint32_t g = nodes[i];
    float4* np = tlas + 4 * size_t(g);
    const float4 d = np[3];
    const int4 lf = leaves[g];
    const Box3 l = child_box(tlas, ibox, __float_as_int(d.x), lf.x, lf.y);
    Box3 r;
    if (lf.w < 0) {  // the single-instance TLAS's second child: the builder's zero box stays
        const float4 e = np[1], f = np[2];
        r = Box3{{e.z, e.w, f.x}, {f.y, f.z, f.w}};
    } else {
        r = child_box(tlas, ibox, __float_as_int(d.y), lf.z, lf.w);
    }
    np[0] = make_float4(l.lo[0], l.lo[1], l.lo[2], l.hi[0]);
    np[1] = make_float4(l.hi[1], l.hi[2], r.lo[0], r.lo[1]);
    np[2] = make_float4(r.lo[2], r.hi[0], r.hi[1], r.hi[2]);
}

}  // namespace

atlas_rt_status launch_world_refit(atlas_rt_world& w, const atlas_rt_instance_desc* d_desc, hipStream_t s) {
    const uint32_t n = uint32_t(w.instances.size());
    const uint32_t blocks = (n + kB - 1) / kB;
    hipLaunchKernelGGL(k_refit_instances, dim3(blocks), dim3(kB), 0, s, n, d_desc,
                       reinterpret_cast<const int32_t*>(w.d_order), reinterpret_cast<const float4*>(w.d_blas_box),
                       reinterpret_cast<float4*>(w.d_instances), reinterpret_cast<float4*>(w.d_xform),
                       reinterpret_cast<float4*>(w.d_inst_box));
    const uint32_t* list = reinterpret_cast<const uint32_t*>(w.d_refit_nodes);
    for (size_t l = 0; l + 1 < w.refit_level_off.size(); l++) {
        const uint32_t first = w.refit_level_off[l], count = w.refit_level_off[l + 1] - first;
        if (count == 0) continue;
        hipLaunchKernelGGL(k_refit_level, dim3((count + kB - 1) / kB), dim3(kB), 0, s, count, list + first,
                           reinterpret_cast<const int4*>(w.d_refit_leaves),
                           reinterpret_cast<const float4*>(w.d_inst_box), reinterpret_cast<float4*>(w.d_tlas));
    }
    ATLAS_HIP_TRY(hipGetLastError());
    w.host_stale = true;
    return ATLAS_RT_OK;
}

}  // namespace atlas_rt
