// BVH2 traversal kernels over one BLAS, gfx950.
//
// Reference: HitClosest / HitAny in BLAS mode (data/shader/raytracer/bvh.hsh:244-267,
// 411-433) with CheckLeafClosest / CheckLeaf (bvh.hsh:44-104), launched by
// traceClosest.csh:12-36 one ray per invocation.
//
// MI355X mapping:
//  * one ray per lane, wave64, BLOCK lanes per workgroup;
//  * the traversal stack lives in LDS as stack[depth][lane] (lane-contiguous,
//    so lanes at different depths never share a bank); entries beyond
//    LDS_DEPTH spill to a per-thread global array sized from the tree depth;
//  * nodes are the engine's 64 B records (one cache line each, four dwordx4
//    loads); triangles use 48 B records with the two edge vectors
//    precomputed;
//  * traversal order and acceptance are the engine's exactly: near child first
//    by entry distance with ties to the left, first accepted triangle wins an
//    equal t, so the hit slot equals the reference's bit for bit.
#include <hip/hip_runtime.h>

#include "rt_device.h"
#include "rt_internal.h"

namespace atlas_rt {

namespace {

struct StackRef {
    int32_t* lds;          // [LDS_DEPTH][BLOCK]
    int32_t* spill;        // [(cap - LDS_DEPTH)][stride]
    uint32_t lane;         // threadIdx.x
    uint32_t gid;          // global thread index (spill column)
    uint32_t stride;       // spill row stride (threads in grid)
};

template <int BLOCK, int LDS_DEPTH>
__device__ __forceinline__ void stack_put(const StackRef& S, uint32_t sp, int32_t v) {
    if (sp < LDS_DEPTH)
        S.lds[sp * BLOCK + S.lane] = v;
    else
        S.spill[size_t(sp - LDS_DEPTH) * S.stride + S.gid] = v;
}

template <int BLOCK, int LDS_DEPTH>
__device__ __forceinline__ int32_t stack_get(const StackRef& S, uint32_t sp) {
    if (sp < LDS_DEPTH) return S.lds[sp * BLOCK + S.lane];
    return S.spill[size_t(sp - LDS_DEPTH) * S.stride + S.gid];
}

// Closest hit for one ray; returns the engine slot (-1 on a miss).
template <int BLOCK, int LDS_DEPTH, bool STATS>
__device__ __forceinline__ int32_t closest_one(const float4* __restrict__ nodes, uint32_t node_count,
                                               const float4* __restrict__ tris, const DevRay& R, float& hitT,
                                               float& hitU, float& hitV, int32_t& prim, const StackRef& S,
                                               uint32_t cap, uint32_t* overflow_flag, uint32_t& n_nodes,
                                               uint32_t& n_tris) {
    int32_t slot = -1;
    int32_t node = node_count ? 0 : ~0;
    uint32_t sp = 0;
    for (;;) {
        if (node >= 0) {
            const float4* np = nodes + 4 * size_t(node);
            const float4 a = np[0], b = np[1], c = np[2], d = np[3];
            if (STATS) n_nodes++;
            float dl, dr;
            const bool hl = slab_dist(R, a.x, a.y, a.z, a.w, b.x, b.y, hitT, dl);
            const bool hr = slab_dist(R, b.z, b.w, c.x, c.y, c.z, c.w, hitT, dr);
            const int32_t lp = __float_as_int(d.x), rp = __float_as_int(d.y);
            const bool leftFirst = dl <= dr;
            if (hl && hr) {
                if (sp >= cap) {
                    atomicOr(overflow_flag, 1u);
                    return -1;
                }
                stack_put<BLOCK, LDS_DEPTH>(S, sp, leftFirst ? rp : lp);
                sp++;
                node = leftFirst ? lp : rp;
            } else if (hl || hr) {
                node = leftFirst ? lp : rp;
            } else {
                if (sp == 0) break;
                sp--;
                node = stack_get<BLOCK, LDS_DEPTH>(S, sp);
            }
        } else {
            int32_t tp = ~node;
            const float leafMax = hitT;
            for (;;) {
                const float4* tr = tris + 3 * size_t(tp);
                const float4 a = tr[0], b = tr[1], c = tr[2];
                float t, u, v;
                if (STATS) n_tris++;
                const bool geo = tri_uvt(R, a.x, a.y, a.z, b.x, b.y, b.z, c.x, c.y, c.z, t, u, v);
                if (geo && t > R.tmin && t < leafMax && t < hitT) {
                    hitT = t;
                    hitU = u;
                    hitV = v;
                    slot = tp;
                    prim = __float_as_int(a.w);
                }
                if (b.w > 0.0f) break;
                tp++;
            }
            if (sp == 0) break;
            sp--;
            node = stack_get<BLOCK, LDS_DEPTH>(S, sp);
        }
    }
    return slot;
}

template <int BLOCK, int LDS_DEPTH>
__device__ __forceinline__ bool any_one(const float4* __restrict__ nodes, uint32_t node_count,
                                        const float4* __restrict__ tris, const DevRay& R, float tmax,
                                        const StackRef& S, uint32_t cap, uint32_t* overflow_flag) {
    int32_t node = node_count ? 0 : ~0;
    uint32_t sp = 0;
    for (;;) {
        if (node >= 0) {
            const float4* np = nodes + 4 * size_t(node);
            const float4 a = np[0], b = np[1], c = np[2], d = np[3];
            float dl, dr;
            const bool hl = slab_dist(R, a.x, a.y, a.z, a.w, b.x, b.y, tmax, dl);
            const bool hr = slab_dist(R, b.z, b.w, c.x, c.y, c.z, c.w, tmax, dr);
            const int32_t lp = __float_as_int(d.x), rp = __float_as_int(d.y);
            if (hl && hr) {
                if (sp >= cap) {
                    atomicOr(overflow_flag, 1u);
                    return false;
                }
                stack_put<BLOCK, LDS_DEPTH>(S, sp, rp);
                sp++;
                node = lp;
            } else if (hl || hr) {
                node = hl ? lp : rp;
            } else {
                if (sp == 0) return false;
                sp--;
                node = stack_get<BLOCK, LDS_DEPTH>(S, sp);
            }
        } else {
            int32_t tp = ~node;
            for (;;) {
                const float4* tr = tris + 3 * size_t(tp);
                const float4 a = tr[0], b = tr[1], c = tr[2];
                float t, u, v;
                const bool geo = tri_uvt(R, a.x, a.y, a.z, b.x, b.y, b.z, c.x, c.y, c.z, t, u, v);
                if (geo && t > R.tmin && t < tmax) return true;
                if (b.w > 0.0f) break;
                tp++;
            }
            if (sp == 0) return false;
            sp--;
            node = stack_get<BLOCK, LDS_DEPTH>(S, sp);
        }
    }
}

template <int BLOCK, int LDS_DEPTH, bool STATS>
__global__ __launch_bounds__(BLOCK) void k_blas_closest(const float4* __restrict__ nodes, uint32_t node_count,
                                                        const float4* __restrict__ tris,
                                                        const float4* __restrict__ rays, float4* __restrict__ hits,
                                                        int32_t* __restrict__ slots, uint32_t count,
                                                        int32_t* __restrict__ spill, uint32_t cap,
                                                        uint32_t* __restrict__ overflow_flag,
                                                        uint2* __restrict__ visits) {
    __shared__ int32_t lds[LDS_DEPTH * BLOCK];
    const uint32_t gid = blockIdx.x * BLOCK + threadIdx.x;
    if (gid >= count) return;
    const float4 r0 = rays[2 * size_t(gid)];
    const float4 r1 = rays[2 * size_t(gid) + 1];
    float hitT = r1.w, hitU = 0.0f, hitV = 0.0f;
    int32_t prim = -1, slot = -1;
    uint32_t n_nodes = 0, n_tris = 0;
    if (!any_nan3(r1.x, r1.y, r1.z)) {
        DevRay R;
        make_dev_ray(R, r0.x, r0.y, r0.z, r1.x, r1.y, r1.z, r0.w);
        StackRef S{lds, spill, threadIdx.x, gid, gridDim.x * BLOCK};
        slot = closest_one<BLOCK, LDS_DEPTH, STATS>(nodes, node_count, tris, R, hitT, hitU, hitV, prim, S, cap,
                                                    overflow_flag, n_nodes, n_tris);
        if (slot < 0) {
            hitT = r1.w;
            hitU = hitV = 0.0f;
            prim = -1;
        }
    }
    hits[gid] = make_float4(hitT, hitU, hitV, __int_as_float(prim));
    if (slots) slots[gid] = slot;
    if (STATS) visits[gid] = make_uint2(n_nodes, n_tris);
}

template <int BLOCK, int LDS_DEPTH>
__global__ __launch_bounds__(BLOCK) void k_blas_any(const float4* __restrict__ nodes, uint32_t node_count,
                                                    const float4* __restrict__ tris, const float4* __restrict__ rays,
                                                    uint32_t* __restrict__ out, uint32_t count,
                                                    int32_t* __restrict__ spill, uint32_t cap,
                                                    uint32_t* __restrict__ overflow_flag) {
    __shared__ int32_t lds[LDS_DEPTH * BLOCK];
    const uint32_t gid = blockIdx.x * BLOCK + threadIdx.x;
    if (gid >= count) return;
    const float4 r0 = rays[2 * size_t(gid)];
    const float4 r1 = rays[2 * size_t(gid) + 1];
    bool hit = false;
    if (!any_nan3(r1.x, r1.y, r1.z)) {
        DevRay R;
        make_dev_ray(R, r0.x, r0.y, r0.z, r1.x, r1.y, r1.z, r0.w);
        StackRef S{lds, spill, threadIdx.x, gid, gridDim.x * BLOCK};
        hit = any_one<BLOCK, LDS_DEPTH>(nodes, node_count, tris, R, r1.w, S, cap, overflow_flag);
    }
    out[gid] = hit ? 1u : 0u;
}

constexpr int kBlock = 128;
constexpr int kLdsDepth = 32;

}  // namespace

static atlas_rt_status prepare_stack(atlas_rt_blas& blas, uint32_t threads, uint32_t& cap) {
    // The stack never holds more entries than the deepest leaf's depth.
    cap = blas.max_depth + 1;
    if (cap > uint32_t(kLdsDepth)) {
        size_t need = size_t(cap - kLdsDepth) * threads * sizeof(int32_t);
        atlas_rt_status st = blas.spill.ensure(need);
        if (st != ATLAS_RT_OK) return st;
    }
    if (!blas.flags.ptr) {
        atlas_rt_status st = blas.flags.ensure(256);
        if (st != ATLAS_RT_OK) return st;
        ATLAS_HIP_TRY(hipMemset(blas.flags.ptr, 0, 256));
    }
    return ATLAS_RT_OK;
}

atlas_rt_status launch_blas_closest(atlas_rt_blas& blas, const TraceLaunch& L) {
    if (L.count == 0) return ATLAS_RT_OK;
    const uint32_t grid = (L.count + kBlock - 1) / kBlock;
    uint32_t cap = 0;
    atlas_rt_status st = prepare_stack(blas, grid * kBlock, cap);
    if (st != ATLAS_RT_OK) return st;
    st = blas.timing.begin(L.stream);
    if (st != ATLAS_RT_OK) return st;
    if (L.visits)
        hipLaunchKernelGGL((k_blas_closest<kBlock, kLdsDepth, true>), dim3(grid), dim3(kBlock), 0, L.stream,
                           blas.d_nodes, uint32_t(blas.nodes.size()), blas.d_tris,
                           reinterpret_cast<const float4*>(L.rays), reinterpret_cast<float4*>(L.hits), L.slots,
                           L.count, reinterpret_cast<int32_t*>(blas.spill.ptr), cap,
                           reinterpret_cast<uint32_t*>(blas.flags.ptr), reinterpret_cast<uint2*>(L.visits));
    else
        hipLaunchKernelGGL((k_blas_closest<kBlock, kLdsDepth, false>), dim3(grid), dim3(kBlock), 0, L.stream,
                           blas.d_nodes, uint32_t(blas.nodes.size()), blas.d_tris,
                           reinterpret_cast<const float4*>(L.rays), reinterpret_cast<float4*>(L.hits), L.slots,
                           L.count, reinterpret_cast<int32_t*>(blas.spill.ptr), cap,
                           reinterpret_cast<uint32_t*>(blas.flags.ptr), nullptr);
    ATLAS_HIP_TRY(hipGetLastError());
    return blas.timing.end(L.stream);
}

atlas_rt_status launch_blas_any(atlas_rt_blas& blas, const TraceLaunch& L) {
    if (L.count == 0) return ATLAS_RT_OK;
    const uint32_t grid = (L.count + kBlock - 1) / kBlock;
    uint32_t cap = 0;
    atlas_rt_status st = prepare_stack(blas, grid * kBlock, cap);
    if (st != ATLAS_RT_OK) return st;
    st = blas.timing.begin(L.stream);
    if (st != ATLAS_RT_OK) return st;
    hipLaunchKernelGGL((k_blas_any<kBlock, kLdsDepth>), dim3(grid), dim3(kBlock), 0, L.stream, blas.d_nodes,
                       uint32_t(blas.nodes.size()), blas.d_tris, reinterpret_cast<const float4*>(L.rays),
                       reinterpret_cast<uint32_t*>(L.hits), L.count, reinterpret_cast<int32_t*>(blas.spill.ptr),
                       cap, reinterpret_cast<uint32_t*>(blas.flags.ptr));
    ATLAS_HIP_TRY(hipGetLastError());
    return blas.timing.end(L.stream);
}

}  // namespace atlas_rt
