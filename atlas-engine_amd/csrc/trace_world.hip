// Two-level (TLAS + instanced BLAS) traversal kernels, gfx950.
//
// Reference: HitClosest / HitAny with CheckInstance (data/shader/raytracer/bvh.hsh:172-273,
// 359-441).  One stack serves both levels; `tlas_index` (the engine's tlasIndex)
// records the stack depth at which the current instance was entered, so
// popping below it returns to the TLAS with the world-space ray restored.
// Entering an instance transforms the ray with the instance's inverse matrix
// as a row vector times mat3x4 (unnormalised direction, so t stays in world
// units, bvh.hsh:177-180) and selects that instance's BLAS.
//
// The BLAS of the current instance is kept in registers; node/triangle steps
// are the same engine-exact arithmetic as the single-BLAS kernels.
#include <hip/hip_runtime.h>

#include "alpha.h"
#include "rt_internal.h"
#include "trace_world_core.h"

namespace atlas_rt {

namespace {

// Compact ray in (32 B) -> refined hit (16 B) + caller instance index.
template <int BLOCK, int LDS_DEPTH>
__global__ __launch_bounds__(BLOCK) void k_world_closest_rays(WorldDev W, const float4* __restrict__ rays,
                                                              float4* __restrict__ hits, int32_t* __restrict__ inst_out,
                                                              uint32_t count, uint32_t mask,
                                                              int32_t* __restrict__ spill, uint32_t cap,
                                                              uint32_t* __restrict__ ctl) {
    __shared__ int32_t lds[LDS_DEPTH * BLOCK];
    const uint32_t gid = blockIdx.x * BLOCK + threadIdx.x;
    if (gid >= count) return;
    const float4 r0 = rays[2 * size_t(gid)];
    const float4 r1 = rays[2 * size_t(gid) + 1];
    WorldLane L;
    L.slot = -1;
    if (!any_nan3(r1.x, r1.y, r1.z)) {
        world_begin(L, r0, r1);
        const LaneStack<BLOCK, LDS_DEPTH> S{lds, spill, threadIdx.x, gid, gridDim.x * BLOCK};
        int r;
        do {
            r = world_step_closest<BLOCK, LDS_DEPTH>(L, W, mask, S, cap);
        } while (r == kContinue);
        if (r == kOverflow) {
            atomicOr(&ctl[2], 1u);
            L.slot = -1;
        }
    }
    float4 out = make_float4(r1.w, 0.0f, 0.0f, __int_as_float(-1));
    int32_t caller_inst = -1;
    if (L.slot >= 0) {
        const float4* im = W.instances + 4 * size_t(L.inst);
        const BlasDev B = ld_blas(W.blases + __float_as_int(im[3].x));
        const float4* vp = B.vtx + 3 * size_t(L.slot);
        float t, u, v;
        refine_tuv(L.hox, L.hoy, L.hoz, L.hdx, L.hdy, L.hdz, vp[0], vp[1], vp[2], t, u, v);
        out = make_float4(t, u, v, B.tris[3 * size_t(L.slot)].w);
        caller_inst = W.inst_order[L.inst];
    }
    hits[gid] = out;
    if (inst_out) inst_out[gid] = caller_inst;
}

template <int BLOCK, int LDS_DEPTH, bool RESTORE_ALWAYS>
__global__ __launch_bounds__(BLOCK) void k_world_any_rays(WorldDev W, const float4* __restrict__ rays,
                                                          uint32_t* __restrict__ out, uint32_t count, uint32_t mask,
                                                          int32_t* __restrict__ spill, uint32_t cap,
                                                          uint32_t* __restrict__ ctl) {
    __shared__ int32_t lds[LDS_DEPTH * BLOCK];
    const uint32_t gid = blockIdx.x * BLOCK + threadIdx.x;
    if (gid >= count) return;
    const float4 r0 = rays[2 * size_t(gid)];
    const float4 r1 = rays[2 * size_t(gid) + 1];
    WorldLane L;
    L.slot = -1;
    if (!any_nan3(r1.x, r1.y, r1.z)) {
        world_begin(L, r0, r1);
        const LaneStack<BLOCK, LDS_DEPTH> S{lds, spill, threadIdx.x, gid, gridDim.x * BLOCK};
        int r;
        do {
            r = world_step_any<BLOCK, LDS_DEPTH, RESTORE_ALWAYS>(L, W, mask, r1.w, S, cap);
        } while (r == kContinue);
        if (r == kOverflow) {
            atomicOr(&ctl[2], 1u);
            L.slot = -1;
        }
    }
    out[gid] = L.slot >= 0 ? 1u : 0u;
}

// Engine-shaped PackedRay kernels (traceClosest.csh:12-36).  The live ray
// count is read on the device.
template <int BLOCK, int LDS_DEPTH>
__global__ __launch_bounds__(BLOCK) void k_world_closest_packed(WorldDev W, const float4* __restrict__ in,
                                                                float4* __restrict__ out,
                                                                const uint32_t* __restrict__ d_count,
                                                                uint32_t max_count, float tmin, float tmax,
                                                                uint32_t mask, int32_t* __restrict__ spill,
                                                                uint32_t cap, uint32_t* __restrict__ ctl) {
    __shared__ int32_t lds[LDS_DEPTH * BLOCK];
    const uint32_t gid = blockIdx.x * BLOCK + threadIdx.x;
    const uint32_t count = d_count ? min(*d_count, max_count) : max_count;
    if (gid >= count) return;
    const float4 p0 = in[3 * size_t(gid)];
    const float4 p1 = in[3 * size_t(gid) + 1];
    const float4 p2 = in[3 * size_t(gid) + 2];
    const int32_t id = __float_as_int(p0.w);
    float hitDistance = 0.0f;
    int32_t hitID = -1;
    int32_t hitInstance = __float_as_int(p2.z);
    if (id >= 0) {
        hitDistance = tmax;
        if (!any_nan3(p1.x, p1.y, p1.z)) {
            WorldLane L;
            world_begin(L, make_float4(p0.x, p0.y, p0.z, tmin), make_float4(p1.x, p1.y, p1.z, tmax));
            const LaneStack<BLOCK, LDS_DEPTH> S{lds, spill, threadIdx.x, gid, gridDim.x * BLOCK};
            int r;
            do {
                r = world_step_closest<BLOCK, LDS_DEPTH>(L, W, mask, S, cap);
            } while (r == kContinue);
            if (r == kOverflow) {
                atomicOr(&ctl[2], 1u);
            } else {
                hitDistance = L.hitT;
                if (L.slot >= 0) {
                    hitID = L.slot;
                    hitInstance = L.inst;
                }
            }
        }
    }
    out[3 * size_t(gid)] = p0;
    out[3 * size_t(gid) + 1] = p1;
    out[3 * size_t(gid) + 2] = make_float4(hitDistance, __int_as_float(hitID), __int_as_float(hitInstance), p2.w);
}

// Opacity-aware PackedRay traces: out_rays != null runs
// HitClosestTransparency (written like k_world_closest_packed), otherwise
// HitAnyTransparency writes the transparency left along each ray.
template <int BLOCK, int LDS_DEPTH>
__global__ __launch_bounds__(BLOCK) void k_world_alpha_packed(WorldDev W, pt::AlphaArgs A, const float4* __restrict__ in,
                                                              float4* __restrict__ out_rays,
                                                              float* __restrict__ out_trans,
                                                              const uint32_t* __restrict__ d_count, uint32_t max_count,
                                                              float tmin, float tmax, uint32_t mask,
                                                              int32_t* __restrict__ spill, uint32_t cap,
                                                              uint32_t* __restrict__ ctl) {
    __shared__ int32_t lds[LDS_DEPTH * BLOCK];
    const uint32_t gid = blockIdx.x * BLOCK + threadIdx.x;
    const uint32_t count = d_count ? min(*d_count, max_count) : max_count;
    if (gid >= count) return;
    const float4 p0 = in[3 * size_t(gid)];
    const float4 p1 = in[3 * size_t(gid) + 1];
    const pt::TriOpacity op{W, A};
    const LaneStack<BLOCK, LDS_DEPTH> S{lds, spill, threadIdx.x, gid, gridDim.x * BLOCK};
    if (!out_rays) {
        float trans = 1.0f;
        if (!any_nan3(p1.x, p1.y, p1.z)) {
            WorldLane L;
            world_begin(L, make_float4(p0.x, p0.y, p0.z, tmin), make_float4(p1.x, p1.y, p1.z, tmax));
            int r;
            do {
                r = world_step_transparency<BLOCK, LDS_DEPTH>(L, W, mask, tmax, S, cap, op);
            } while (r == kContinue);
            if (r == kOverflow) atomicOr(&ctl[2], 1u);
            trans = L.trans;
        }
        out_trans[gid] = trans;
        return;
    }
    const float4 p2 = in[3 * size_t(gid) + 2];
    const int32_t id = __float_as_int(p0.w);
    float hitDistance = 0.0f;
    int32_t hitID = -1;
    int32_t hitInstance = __float_as_int(p2.z);
    if (id >= 0) {
        hitDistance = tmax;
        if (!any_nan3(p1.x, p1.y, p1.z)) {
            WorldLane L;
            world_begin(L, make_float4(p0.x, p0.y, p0.z, tmin), make_float4(p1.x, p1.y, p1.z, tmax));
            const pt::AcceptOpaque acc{op};
            int r;
            do {
                r = world_step_closest<BLOCK, LDS_DEPTH>(L, W, mask, S, cap, acc);
            } while (r == kContinue);
            if (r == kOverflow) {
                atomicOr(&ctl[2], 1u);
            } else {
                hitDistance = L.hitT;
                if (L.slot >= 0) {
                    hitID = L.slot;
                    hitInstance = L.inst;
                }
            }
        }
    }
    out_rays[3 * size_t(gid)] = p0;
    out_rays[3 * size_t(gid) + 1] = p1;
    out_rays[3 * size_t(gid) + 2] = make_float4(hitDistance, __int_as_float(hitID), __int_as_float(hitInstance), p2.w);
}

template <int BLOCK, int LDS_DEPTH, bool RESTORE_ALWAYS>
__global__ __launch_bounds__(BLOCK) void k_world_any_packed(WorldDev W, const float4* __restrict__ in,
                                                            uint32_t* __restrict__ out,
                                                            const uint32_t* __restrict__ d_count, uint32_t max_count,
                                                            float tmin, float tmax, uint32_t mask,
                                                            int32_t* __restrict__ spill, uint32_t cap,
                                                            uint32_t* __restrict__ ctl) {
    __shared__ int32_t lds[LDS_DEPTH * BLOCK];
    const uint32_t gid = blockIdx.x * BLOCK + threadIdx.x;
    const uint32_t count = d_count ? min(*d_count, max_count) : max_count;
    if (gid >= count) return;
    const float4 p0 = in[3 * size_t(gid)];
    const float4 p1 = in[3 * size_t(gid) + 1];
    uint32_t hit = 0;
    if (!any_nan3(p1.x, p1.y, p1.z)) {
        WorldLane L;
        world_begin(L, make_float4(p0.x, p0.y, p0.z, tmin), make_float4(p1.x, p1.y, p1.z, tmax));
        const LaneStack<BLOCK, LDS_DEPTH> S{lds, spill, threadIdx.x, gid, gridDim.x * BLOCK};
        int r;
        do {
            r = world_step_any<BLOCK, LDS_DEPTH, RESTORE_ALWAYS>(L, W, mask, tmax, S, cap);
        } while (r == kContinue);
        if (r == kOverflow) atomicOr(&ctl[2], 1u);
        hit = (r == kDone && L.slot >= 0) ? 1u : 0u;
    }
    out[gid] = hit;
}

constexpr int kBlock = 128;
constexpr int kLdsDepth = 32;

}  // namespace

static atlas_rt_status world_prepare(atlas_rt_world& w, uint32_t threads, uint32_t& cap) {
    cap = w.stack_cap();
    if (cap > uint32_t(kLdsDepth)) {
        atlas_rt_status st = w.spill.ensure(size_t(cap - kLdsDepth) * threads * sizeof(int32_t));
        if (st != ATLAS_RT_OK) return st;
    }
    if (!w.flags.ptr) {
        atlas_rt_status st = w.flags.ensure(256);
        if (st != ATLAS_RT_OK) return st;
        ATLAS_HIP_TRY(hipMemset(w.flags.ptr, 0, 256));
    }
    return ATLAS_RT_OK;
}

atlas_rt_status launch_world_closest_rays(atlas_rt_world& w, const void* rays, void* hits, int32_t* inst,
                                          uint32_t count, uint32_t mask, hipStream_t stream) {
    if (!count) return ATLAS_RT_OK;
    const uint32_t grid = (count + kBlock - 1) / kBlock;
    uint32_t cap;
    atlas_rt_status st = world_prepare(w, grid * kBlock, cap);
    if (st != ATLAS_RT_OK) return st;
    st = w.timing.begin(stream);
    if (st != ATLAS_RT_OK) return st;
    hipLaunchKernelGGL((k_world_closest_rays<kBlock, kLdsDepth>), dim3(grid), dim3(kBlock), 0, stream, w.dev(),
                       reinterpret_cast<const float4*>(rays), reinterpret_cast<float4*>(hits), inst, count, mask,
                       reinterpret_cast<int32_t*>(w.spill.ptr), cap, reinterpret_cast<uint32_t*>(w.flags.ptr));
    ATLAS_HIP_TRY(hipGetLastError());
    return w.timing.end(stream);
}

atlas_rt_status launch_world_any_rays(atlas_rt_world& w, const void* rays, uint32_t* out, uint32_t count,
                                      uint32_t mask, uint32_t flags, hipStream_t stream) {
    if (!count) return ATLAS_RT_OK;
    const uint32_t grid = (count + kBlock - 1) / kBlock;
    uint32_t cap;
    atlas_rt_status st = world_prepare(w, grid * kBlock, cap);
    if (st != ATLAS_RT_OK) return st;
    if (flags & ATLAS_RT_ANY_RESTORE_ALWAYS)
        hipLaunchKernelGGL((k_world_any_rays<kBlock, kLdsDepth, true>), dim3(grid), dim3(kBlock), 0, stream, w.dev(),
                           reinterpret_cast<const float4*>(rays), out, count, mask,
                           reinterpret_cast<int32_t*>(w.spill.ptr), cap, reinterpret_cast<uint32_t*>(w.flags.ptr));
    else
        hipLaunchKernelGGL((k_world_any_rays<kBlock, kLdsDepth, false>), dim3(grid), dim3(kBlock), 0, stream,
                           w.dev(), reinterpret_cast<const float4*>(rays), out, count, mask,
                           reinterpret_cast<int32_t*>(w.spill.ptr), cap, reinterpret_cast<uint32_t*>(w.flags.ptr));
    ATLAS_HIP_TRY(hipGetLastError());
    return ATLAS_RT_OK;
}

atlas_rt_status launch_world_closest_packed(atlas_rt_world& w, const void* in, void* out, const uint32_t* d_count,
                                            uint32_t max_count, float tmin, float tmax, uint32_t mask,
                                            hipStream_t stream) {
    if (!max_count) return ATLAS_RT_OK;
    const uint32_t grid = (max_count + kBlock - 1) / kBlock;
    uint32_t cap;
    atlas_rt_status st = world_prepare(w, grid * kBlock, cap);
    if (st != ATLAS_RT_OK) return st;
    hipLaunchKernelGGL((k_world_closest_packed<kBlock, kLdsDepth>), dim3(grid), dim3(kBlock), 0, stream, w.dev(),
                       reinterpret_cast<const float4*>(in), reinterpret_cast<float4*>(out), d_count, max_count, tmin,
                       tmax, mask, reinterpret_cast<int32_t*>(w.spill.ptr), cap,
                       reinterpret_cast<uint32_t*>(w.flags.ptr));
    ATLAS_HIP_TRY(hipGetLastError());
    return ATLAS_RT_OK;
}

atlas_rt_status launch_world_alpha_packed(atlas_rt_world& w, const void* in, void* out_rays, float* out_trans,
                                          const uint32_t* d_count, uint32_t max_count, float tmin, float tmax,
                                          uint32_t mask, hipStream_t stream) {
    if (!max_count) return ATLAS_RT_OK;
    // per-BLAS shading and uv pointers, refreshed when a BLAS's buffers change
    const size_t nb = w.blas_handles.size();
    std::vector<const void*> tab(2 * nb);
    for (size_t k = 0; k < nb; k++) {
        if (!w.blas_handles[k]->d_shade)
            return fail(ATLAS_RT_E_INVALID_ARG, "opacity-aware trace: a BLAS has no shading data");
        tab[k] = w.blas_handles[k]->d_shade;
        tab[nb + k] = w.blas_handles[k]->d_uv;
    }
    if (tab != w.alpha_tab_host) {
        atlas_rt_status st = w.alpha_tab.ensure(tab.size() * sizeof(void*));
        if (st != ATLAS_RT_OK) return st;
        ATLAS_HIP_TRY(hipStreamSynchronize(stream));
        ATLAS_HIP_TRY(hipMemcpy(w.alpha_tab.ptr, tab.data(), tab.size() * sizeof(void*), hipMemcpyHostToDevice));
        w.alpha_tab_host = tab;
    }
    const float4* const* dtab = reinterpret_cast<const float4* const*>(w.alpha_tab.ptr);
    const pt::AlphaArgs A{dtab, dtab + nb, nullptr, nullptr, nullptr, 0u, 0u};
    const uint32_t grid = (max_count + kBlock - 1) / kBlock;
    uint32_t cap;
    atlas_rt_status st = world_prepare(w, grid * kBlock, cap);
    if (st != ATLAS_RT_OK) return st;
    hipLaunchKernelGGL((k_world_alpha_packed<kBlock, kLdsDepth>), dim3(grid), dim3(kBlock), 0, stream, w.dev(), A,
                       reinterpret_cast<const float4*>(in), reinterpret_cast<float4*>(out_rays), out_trans, d_count,
                       max_count, tmin, tmax, mask, reinterpret_cast<int32_t*>(w.spill.ptr), cap,
                       reinterpret_cast<uint32_t*>(w.flags.ptr));
    ATLAS_HIP_TRY(hipGetLastError());
    return ATLAS_RT_OK;
}

atlas_rt_status launch_world_any_packed(atlas_rt_world& w, const void* in, uint32_t* out, const uint32_t* d_count,
                                        uint32_t max_count, float tmin, float tmax, uint32_t mask, uint32_t flags,
                                        hipStream_t stream) {
    if (!max_count) return ATLAS_RT_OK;
    const uint32_t grid = (max_count + kBlock - 1) / kBlock;
    uint32_t cap;
    atlas_rt_status st = world_prepare(w, grid * kBlock, cap);
    if (st != ATLAS_RT_OK) return st;
    if (flags & ATLAS_RT_ANY_RESTORE_ALWAYS)
        hipLaunchKernelGGL((k_world_any_packed<kBlock, kLdsDepth, true>), dim3(grid), dim3(kBlock), 0, stream,
                           w.dev(), reinterpret_cast<const float4*>(in), out, d_count, max_count, tmin, tmax, mask,
                           reinterpret_cast<int32_t*>(w.spill.ptr), cap, reinterpret_cast<uint32_t*>(w.flags.ptr));
    else
        hipLaunchKernelGGL((k_world_any_packed<kBlock, kLdsDepth, false>), dim3(grid), dim3(kBlock), 0, stream,
                           w.dev(), reinterpret_cast<const float4*>(in), out, d_count, max_count, tmin, tmax, mask,
                           reinterpret_cast<int32_t*>(w.spill.ptr), cap, reinterpret_cast<uint32_t*>(w.flags.ptr));
    ATLAS_HIP_TRY(hipGetLastError());
    return ATLAS_RT_OK;
}

}  // namespace atlas_rt
