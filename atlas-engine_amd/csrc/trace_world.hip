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

#include "rt_internal.h"
#include "trace_core.h"
#include "world.h"

namespace atlas_rt {

namespace {

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
}

// CheckInstance (bvh.hsh:172-189); returns whether the mask admits it.
__device__ __forceinline__ bool enter_instance(WorldLane& L, const WorldDev& W, uint32_t mask) {
    const int32_t ip = ~L.node;
    const float4* m = W.instances + 4 * size_t(ip);
    const float4 c0 = m[0], c1 = m[1], c2 = m[2], meta = m[3];
    const float ox = dot4(L.o0x, L.o0y, L.o0z, 1.0f, c0);
    const float oy = dot4(L.o0x, L.o0y, L.o0z, 1.0f, c1);
    const float oz = dot4(L.o0x, L.o0y, L.o0z, 1.0f, c2);
    const float dx = dot4(L.d0x, L.d0y, L.d0z, 0.0f, c0);
    const float dy = dot4(L.d0x, L.d0y, L.d0z, 0.0f, c1);
    const float dz = dot4(L.d0x, L.d0y, L.d0z, 0.0f, c2);
    make_dev_ray(L.R, ox, oy, oz, dx, dy, dz, L.R.tmin);
    L.cur_inst = ip;
    const int32_t mesh = __float_as_int(meta.x);
    L.B = W.blases[mesh];
    L.node = L.B.node_count ? 0 : ~0;
    return (uint32_t(__float_as_int(meta.w)) & mask) > 0;
}

template <int BLOCK, int LDS_DEPTH>
__device__ __forceinline__ int world_step_closest(WorldLane& L, const WorldDev& W, uint32_t mask,
                                                  const LaneStack<BLOCK, LDS_DEPTH>& S, uint32_t cap) {
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
        const float4 a = np[0], b = np[1], c = np[2], d = np[3];
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
        const float4 a = np[0], b = np[1], c = np[2], d = np[3];
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
    const float4 a = tr[0], b = tr[1], c = tr[2];
    float t, u, v;
    const bool geo = tri_uvt(L.R, a.x, a.y, a.z, b.x, b.y, b.z, c.x, c.y, c.z, t, u, v);
    if (geo && t > L.R.tmin && t < L.hitT) {
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
    if (L.node >= 0) {
        const float4* np = L.B.nodes + 4 * size_t(L.node);
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
    const float4* tr = L.B.tris + 3 * size_t(tp);
    const float4 a = tr[0], b = tr[1], c = tr[2];
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
        const BlasDev B = W.blases[__float_as_int(im[3].x)];
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
    cap = w.max_depth + 1;
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
