// BVH2 traversal kernels over one BLAS, gfx950.
//
// Reference: HitClosest / HitAny in BLAS mode (data/shader/raytracer/bvh.hsh:244-267,
// 411-433), launched one ray per invocation by traceClosest.csh:12-36.
//
// Two launch shapes share the per-lane step function of trace_core.h:
//  * grid: one ray per lane, the whole batch in one grid (the engine's shape);
//  * persistent: a resident grid of waves pulls rays from a global counter.
//    Whenever fewer than `refill` lanes of a wave are still traversing, the
//    idle lanes are refilled with new rays (ballot + popcount give each idle
//    lane its index from one atomicAdd per wave), so divergent incoherent
//    rays do not leave most of a wave idle.  The last wave to finish resets
//    the counter, so consecutive launches need no host-side reset.
// The stack lives in LDS rows [depth][lane] (bank-conflict free since lanes
// at different depths still sit in different banks) and spills to a global
// array sized from the tree depth.  Results equal the engine's bit for bit.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "rt_internal.h"
#include "trace_core.h"

namespace atlas_rt {

namespace {

constexpr int kWave = 64;

template <int BLOCK, int LDS_DEPTH, bool STATS, int NODE_W = 2, bool REG = false, bool FLOAT_SLAB = false>
__global__ __launch_bounds__(BLOCK) void k_blas_closest(BlasDev B, const float4* __restrict__ rays,
                                                        float4* __restrict__ hits, int32_t* __restrict__ slots,
                                                        uint2* __restrict__ visits, uint32_t count,
                                                        int32_t* __restrict__ spill, uint32_t cap,
                                                        uint32_t* __restrict__ ctl) {
    __shared__ int32_t lds[LDS_DEPTH * BLOCK];
    const uint32_t gid = blockIdx.x * BLOCK + threadIdx.x;
    if (gid >= count) return;
    const float4 r0 = rays[2 * size_t(gid)];
    const float4 r1 = rays[2 * size_t(gid) + 1];
    LaneState L;
    L.slot = -1;
    uint32_t n_nodes = 0, n_tris = 0;
    if (!any_nan3(r1.x, r1.y, r1.z)) {
        lane_begin(L, B, r0, r1, NODE_W > 2);
        const LaneStack<BLOCK, LDS_DEPTH> S{lds, spill, threadIdx.x, gid, gridDim.x * BLOCK};
        int r;
        do {
            r = step_variant<NODE_W, REG, BLOCK, LDS_DEPTH, STATS, FLOAT_SLAB>(B, L, S, cap, n_nodes, n_tris);
        } while (r == kContinue);
        if (r == kOverflow) {
            atomicOr(&ctl[2], 1u);
            L.slot = -1;
        }
    }
    hits[gid] = lane_hit_record(B, L, r1.w);
    if (slots) slots[gid] = L.slot;
    if (STATS) visits[gid] = make_uint2(n_nodes, n_tris);
}

// LEAF_DIV > 0: leaf-batched waves (wave_step).
template <int BLOCK, int LDS_DEPTH, int NODE_W = 2, bool REG = false, bool FLOAT_SLAB = false, int LEAF_DIV = 0>
__global__ __launch_bounds__(BLOCK) void k_blas_closest_persistent(BlasDev B, const float4* __restrict__ rays,
                                                                   float4* __restrict__ hits,
                                                                   int32_t* __restrict__ slots, uint32_t count,
                                                                   int32_t* __restrict__ spill, uint32_t cap,
                                                                   uint32_t* __restrict__ ctl, uint32_t refill) {
    __shared__ int32_t lds[LDS_DEPTH * BLOCK];
    const uint32_t lane = threadIdx.x & (kWave - 1);
    const LaneStack<BLOCK, LDS_DEPTH> S{lds, spill, threadIdx.x, blockIdx.x * BLOCK + threadIdx.x, gridDim.x * BLOCK};
    LaneState L;
    L.slot = -1;
    bool active = false;
    bool exhausted = false;
    uint32_t idx = 0;
    float tmax = 0.0f;
    bool pending = false;  // the lane's ray has ended; its hit record is not written yet
    RangePull rp;  // rays from kPullRanges ranges, counters ctl[kRangeWord * (r + 1)]
    rp.init(count);
    for (;;) {
        // The hit records of the rays that ended since the last refill, all
        // at once: the FP64 refinement (refine_tuv) then runs once per refill
        // round instead of in every traversal iteration in which some lane
        // of the wave ended its ray.
        if (pending) {
            hits[idx] = lane_hit_record(B, L, tmax);
            if (slots) slots[idx] = L.slot;
            pending = false;
        }
        if (!exhausted) {
            const uint64_t idle = __ballot(!active);
            if (idle) {
                const uint32_t n = __popcll(idle);
                uint32_t base = 0, end = 0;
                if (lane == 0) rp.take(ctl + kRangeWord, kRangeWord, count, n, base, end);
                base = __shfl(base, 0);
                end = __shfl(end, 0);
                rp.broadcast(0);
                if (rp.dry >= kPullRanges) exhausted = true;
                if (!active) {
                    idx = base + __popcll(idle & ((1ull << lane) - 1));
                    if (idx < end) {
                        const float4 r0 = rays[2 * size_t(idx)];
                        const float4 r1 = rays[2 * size_t(idx) + 1];
                        tmax = r1.w;
                        if (any_nan3(r1.x, r1.y, r1.z)) {
                            hits[idx] = make_float4(tmax, 0.0f, 0.0f, __int_as_float(-1));
                            if (slots) slots[idx] = -1;
                        } else {
                            lane_begin(L, B, r0, r1, NODE_W > 2);
                            active = true;
                        }
                    }
                }
            }
        }
        if (!__any(active)) {
            if (exhausted) break;
            continue;
        }
        const uint32_t threshold = exhausted ? 1u : refill;
        for (;;) {
            const int r = wave_step<NODE_W, REG, BLOCK, LDS_DEPTH, FLOAT_SLAB, LEAF_DIV>(B, L, S, cap, active);
            if (r != kContinue) {
                if (r == kOverflow) {
                    atomicOr(&ctl[2], 1u);
                    L.slot = -1;
                }
                pending = true;
                active = false;
            }
            if (uint32_t(__popcll(__ballot(active))) < threshold) break;
        }
    }
    // Every wave has taken its last ray; the last one out resets the counters.
    if (lane == 0) {
        __threadfence();
        const uint32_t waves = gridDim.x * (BLOCK / kWave);
        if (atomicAdd(&ctl[1], 1u) == waves - 1) {
            for (uint32_t r = 0; r < kPullRanges; r++) atomicExch(&ctl[kRangeWord * (r + 1)], 0u);
            atomicExch(&ctl[1], 0u);
        }
    }
}

template <int BLOCK, int LDS_DEPTH>
__global__ __launch_bounds__(BLOCK) void k_blas_any(BlasDev B, const float4* __restrict__ rays,
                                                    uint32_t* __restrict__ out, uint32_t count,
                                                    int32_t* __restrict__ spill, uint32_t cap,
                                                    uint32_t* __restrict__ ctl) {
    __shared__ int32_t lds[LDS_DEPTH * BLOCK];
    const uint32_t gid = blockIdx.x * BLOCK + threadIdx.x;
    if (gid >= count) return;
    const float4 r0 = rays[2 * size_t(gid)];
    const float4 r1 = rays[2 * size_t(gid) + 1];
    LaneState L;
    L.slot = -1;
    if (!any_nan3(r1.x, r1.y, r1.z)) {
        lane_begin(L, B, r0, r1);
        const LaneStack<BLOCK, LDS_DEPTH> S{lds, spill, threadIdx.x, gid, gridDim.x * BLOCK};
        int r;
        do {
            r = step_any<BLOCK, LDS_DEPTH>(B, L, S, cap, r1.w);
        } while (r == kContinue);
        if (r == kOverflow) {
            atomicOr(&ctl[2], 1u);
            L.slot = -1;
        }
    }
    out[gid] = L.slot >= 0 ? 1u : 0u;
}

constexpr int kBlock = 128;
constexpr int kLdsDepth = 32;
constexpr int kPBlock = 256;
constexpr int kPLdsDepth = 24;

BlasDev blas_dev(const atlas_rt_blas& b) {
    BlasDev d;
    d.nodes = b.d_nodes;
    d.tris = b.d_tris;
    d.vtx = b.d_vtx;
    d.nodes4 = b.d_nodes4;
    d.nodes8 = b.d_nodes8;
    d.node_count = uint32_t(b.nodes.size());
    d.tri_count = uint32_t(b.tris.size());
    return d;
}

}  // namespace

int device_cu_count() {
    static int cached = -1;
    if (cached < 0) {
        int dev = 0;
        hipDeviceProp_t p;
        if (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&p, dev) == hipSuccess &&
            p.multiProcessorCount > 0)
            cached = p.multiProcessorCount;
        else
            cached = 256;
    }
    return cached;
}

static atlas_rt_status prepare(atlas_rt_blas& blas, uint32_t threads, uint32_t lds_depth, uint32_t& cap,
                               int node_w = 2) {
    // The binary stack never holds more entries than the deepest leaf's
    // depth; the wide bounds come from the collapse.
    cap = limit_cap(variant_cap(node_w, blas.max_depth + 1, blas.cap4, blas.cap8), blas.stack_limit);
    if (cap > lds_depth) {
        const size_t need = size_t(cap - lds_depth) * threads * sizeof(int32_t);
        atlas_rt_status st = blas.spill.ensure(need);
        if (st != ATLAS_RT_OK) return st;
    }
    if (!blas.flags.ptr) {
        atlas_rt_status st = blas.flags.ensure(kFlagBytes);
        if (st != ATLAS_RT_OK) return st;
        ATLAS_HIP_TRY(hipMemset(blas.flags.ptr, 0, kFlagBytes));
    }
    return ATLAS_RT_OK;
}

template <int NODE_W, bool REG, bool FLOAT_SLAB = false, int LEAF_DIV = 0>
static void launch_persistent(const BlasDev& B, const TraceLaunch& L, uint32_t groups, int32_t* spill, uint32_t cap,
                              uint32_t* ctl) {
    hipLaunchKernelGGL((k_blas_closest_persistent<kPBlock, kPLdsDepth, NODE_W, REG, FLOAT_SLAB, LEAF_DIV>), dim3(groups),
                       dim3(kPBlock), 0,
                       L.stream, B, reinterpret_cast<const float4*>(L.rays), reinterpret_cast<float4*>(L.hits),
                       L.slots, L.count, spill, cap, ctl, L.refill ? L.refill : 40u);
}

// Leaf batching: triangle steps once a quarter of a wave's live lanes wait at
// a leaf.
constexpr int kLeafDiv = 4;

// ATLAS_RT_KERNEL_AUTO: resident leaf-batched waves (the fastest
// engine-exact variant on C0) for large batches, one ray per lane for small
// ones, where a resident grid would mostly idle.
constexpr uint32_t kAutoPersistentRays = 65536;

atlas_rt_status launch_blas_closest(atlas_rt_blas& blas, const TraceLaunch& L0) {
    if (L0.count == 0) return ATLAS_RT_OK;
    TraceLaunch L = L0;
    if (L.kernel == ATLAS_RT_KERNEL_AUTO && !L.visits)
        L.kernel = L.count >= kAutoPersistentRays ? (ATLAS_RT_KERNEL_BVH2_PERSISTENT | ATLAS_RT_KERNEL_LEAF_BATCH)
                                                  : ATLAS_RT_KERNEL_BVH2_EXACT;
    const BlasDev B = blas_dev(blas);
    uint32_t* ctl = reinterpret_cast<uint32_t*>(blas.flags.ptr);
    const int base = L.kernel & ATLAS_RT_KERNEL_VARIANT_MASK;
    const bool reg = (L.kernel & ATLAS_RT_KERNEL_STACK_REGISTER) != 0;
    const bool fslab = (L.kernel & ATLAS_RT_KERNEL_SLAB_FLOAT) != 0;
    const bool leaf_batch = (L.kernel & ATLAS_RT_KERNEL_LEAF_BATCH) != 0;
    const bool persistent = !L.visits && base >= ATLAS_RT_KERNEL_BVH2_PERSISTENT;
    const int node_w = base == ATLAS_RT_KERNEL_BVH4 ? 4 : (base == ATLAS_RT_KERNEL_BVH8 ? 8 : 2);
    atlas_rt_status st;
    uint32_t cap = 0;
    if (persistent) {
        // Resident grid: 256-lane groups per CU.  Four in general (measured
        // on C0: 0.98 ms against 1.06 ms with eight; fewer resident waves
        // keep more of the tree in each XCD's L2); five for the binary exact
        // leaf-batched kernel, whose 96 VGPRs leave room for a fifth wave per
        // SIMD (C0: 0.824 against 0.852 ms, where the float-slab variants
        // slow down with five; six, built for 80 VGPRs, took 0.84 ms).
        const bool five = node_w == 2 && leaf_batch && !reg && !fslab;
        const uint32_t groups = uint32_t(device_cu_count()) * (five ? 5u : 4u);
        st = prepare(blas, groups * kPBlock, kPLdsDepth, cap, node_w);
        if (st != ATLAS_RT_OK) return st;
        ctl = reinterpret_cast<uint32_t*>(blas.flags.ptr);
        st = blas.timing.begin(L.stream);
        if (st != ATLAS_RT_OK) return st;
        int32_t* spill = reinterpret_cast<int32_t*>(blas.spill.ptr);
        if (node_w == 2 && leaf_batch) {
            if (fslab)
                reg ? launch_persistent<2, true, true, kLeafDiv>(B, L, groups, spill, cap, ctl)
                    : launch_persistent<2, false, true, kLeafDiv>(B, L, groups, spill, cap, ctl);
            else
                reg ? launch_persistent<2, true, false, kLeafDiv>(B, L, groups, spill, cap, ctl)
                    : launch_persistent<2, false, false, kLeafDiv>(B, L, groups, spill, cap, ctl);
        } else if (node_w == 2 && fslab)
            reg ? launch_persistent<2, true, true>(B, L, groups, spill, cap, ctl)
                : launch_persistent<2, false, true>(B, L, groups, spill, cap, ctl);
        else if (node_w == 2)
            reg ? launch_persistent<2, true>(B, L, groups, spill, cap, ctl)
                : launch_persistent<2, false>(B, L, groups, spill, cap, ctl);
        else if (node_w == 4 && leaf_batch && fslab)
            reg ? launch_persistent<4, true, true, kLeafDiv>(B, L, groups, spill, cap, ctl)
                : launch_persistent<4, false, true, kLeafDiv>(B, L, groups, spill, cap, ctl);
        else if (node_w == 4 && leaf_batch)
            reg ? launch_persistent<4, true, false, kLeafDiv>(B, L, groups, spill, cap, ctl)
                : launch_persistent<4, false, false, kLeafDiv>(B, L, groups, spill, cap, ctl);
        else if (node_w == 4 && fslab)
            reg ? launch_persistent<4, true, true>(B, L, groups, spill, cap, ctl)
                : launch_persistent<4, false, true>(B, L, groups, spill, cap, ctl);
        else if (node_w == 4)
            reg ? launch_persistent<4, true>(B, L, groups, spill, cap, ctl)
                : launch_persistent<4, false>(B, L, groups, spill, cap, ctl);
        else
            reg ? launch_persistent<8, true>(B, L, groups, spill, cap, ctl)
                : launch_persistent<8, false>(B, L, groups, spill, cap, ctl);
    } else if ((reg || fslab) && !L.visits) {
        const uint32_t grid = (L.count + kBlock - 1) / kBlock;
        st = prepare(blas, grid * kBlock, kLdsDepth, cap);
        if (st != ATLAS_RT_OK) return st;
        ctl = reinterpret_cast<uint32_t*>(blas.flags.ptr);
        st = blas.timing.begin(L.stream);
        if (st != ATLAS_RT_OK) return st;
        const auto* rays = reinterpret_cast<const float4*>(L.rays);
        auto* hits = reinterpret_cast<float4*>(L.hits);
        int32_t* spill = reinterpret_cast<int32_t*>(blas.spill.ptr);
        if (reg && fslab)
            hipLaunchKernelGGL((k_blas_closest<kBlock, kLdsDepth, false, 2, true, true>), dim3(grid), dim3(kBlock), 0,
                               L.stream, B, rays, hits, L.slots, nullptr, L.count, spill, cap, ctl);
        else if (reg)
            hipLaunchKernelGGL((k_blas_closest<kBlock, kLdsDepth, false, 2, true>), dim3(grid), dim3(kBlock), 0,
                               L.stream, B, rays, hits, L.slots, nullptr, L.count, spill, cap, ctl);
        else
            hipLaunchKernelGGL((k_blas_closest<kBlock, kLdsDepth, false, 2, false, true>), dim3(grid), dim3(kBlock), 0,
                               L.stream, B, rays, hits, L.slots, nullptr, L.count, spill, cap, ctl);
    } else {
        const uint32_t grid = (L.count + kBlock - 1) / kBlock;
        st = prepare(blas, grid * kBlock, kLdsDepth, cap);
        if (st != ATLAS_RT_OK) return st;
        ctl = reinterpret_cast<uint32_t*>(blas.flags.ptr);
        st = blas.timing.begin(L.stream);
        if (st != ATLAS_RT_OK) return st;
        if (L.visits)
            hipLaunchKernelGGL((k_blas_closest<kBlock, kLdsDepth, true>), dim3(grid), dim3(kBlock), 0, L.stream, B,
                               reinterpret_cast<const float4*>(L.rays), reinterpret_cast<float4*>(L.hits), L.slots,
                               reinterpret_cast<uint2*>(L.visits), L.count,
                               reinterpret_cast<int32_t*>(blas.spill.ptr), cap, ctl);
        else
            hipLaunchKernelGGL((k_blas_closest<kBlock, kLdsDepth, false>), dim3(grid), dim3(kBlock), 0, L.stream, B,
                               reinterpret_cast<const float4*>(L.rays), reinterpret_cast<float4*>(L.hits), L.slots,
                               nullptr, L.count, reinterpret_cast<int32_t*>(blas.spill.ptr), cap, ctl);
    }
    ATLAS_HIP_TRY(hipGetLastError());
    return blas.timing.end(L.stream);
}

// ---------------------------------------------------------------------------
// Closest hit without allocation: the persistent kernels' counters, overflow
// flag and stack spill live in one file-scope device array instead of the
// BLAS's scratch, so a trace never calls hipMalloc.  Every call on a device
// shares that device's copy; calls that overlap on one device race.
// ---------------------------------------------------------------------------
namespace {

constexpr size_t kNoAllocCtlBytes = 2048;
static_assert(kFlagBytes <= kNoAllocCtlBytes, "the flag block fits the control bytes");
constexpr size_t kNoAllocSpillBytes = size_t(16) << 20;

__device__ __align__(256) uint8_t scratch[kNoAllocCtlBytes + kNoAllocSpillBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// The device address of the current device's `scratch`, looked up once per
// device with that device current.
atlas_rt_status scratch_address(void** out) {
    int dev = 0;
    ATLAS_HIP_TRY(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return fail(ATLAS_RT_E_UNSUPPORTED, "no-alloc trace: device index out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return hip_fail(e, "hipGetSymbolAddress(scratch)");
        if (!p) return fail(ATLAS_RT_E_HIP, "hipGetSymbolAddress(scratch) returned NULL");
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return ATLAS_RT_OK;
}

__global__ void k_zero_words(uint32_t* __restrict__ w, uint32_t n) {
    for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) w[i] = 0u;
}

}  // namespace

atlas_rt_status launch_blas_closest_noalloc(const atlas_rt_blas& blas, const TraceLaunch& L0) {
    TraceLaunch L = L0;
    if (L.kernel == ATLAS_RT_KERNEL_AUTO) L.kernel = ATLAS_RT_KERNEL_BVH2_PERSISTENT | ATLAS_RT_KERNEL_LEAF_BATCH;
    const int base = L.kernel & ATLAS_RT_KERNEL_VARIANT_MASK;
    if (base != ATLAS_RT_KERNEL_BVH2_PERSISTENT)
        return fail(ATLAS_RT_E_UNSUPPORTED, "atlas_rt_trace_closest_rays_noalloc: only the binary persistent kernels "
                                            "(ATLAS_RT_KERNEL_BVH2_PERSISTENT with the register-stack, float-slab and "
                                            "leaf-batch bits, or AUTO)");
    void* sp = nullptr;
    atlas_rt_status st = scratch_address(&sp);
    if (st != ATLAS_RT_OK) return st;
    if (L.count == 0) return ATLAS_RT_OK;
    const uint32_t groups = uint32_t(device_cu_count()) * 4;
    const size_t lanes = size_t(groups) * kPBlock;
    const uint32_t cap = limit_cap(blas.max_depth + 1, blas.stack_limit);
    const size_t max_cap = kPLdsDepth + kNoAllocSpillBytes / (lanes * sizeof(int32_t));
    if (cap > max_cap)
        return fail(ATLAS_RT_E_STACK_OVERFLOW, "atlas_rt_trace_closest_rays_noalloc: the tree needs " +
                                                   std::to_string(cap) + " stack entries per ray; the fixed scratch "
                                                   "holds " + std::to_string(max_cap) +
                                                   " (use atlas_rt_trace_closest_rays)");
    uint32_t* ctl = reinterpret_cast<uint32_t*>(sp);
    int32_t* spill = reinterpret_cast<int32_t*>(static_cast<uint8_t*>(sp) + kNoAllocCtlBytes);
    // the range counters, the wave counter and the overflow flag start at zero
    hipLaunchKernelGGL(k_zero_words, dim3(1), dim3(256), 0, L.stream, ctl, uint32_t(kNoAllocCtlBytes / 4));
    const BlasDev B = blas_dev(blas);
    const bool reg = (L.kernel & ATLAS_RT_KERNEL_STACK_REGISTER) != 0;
    const bool fslab = (L.kernel & ATLAS_RT_KERNEL_SLAB_FLOAT) != 0;
    const bool leaf_batch = (L.kernel & ATLAS_RT_KERNEL_LEAF_BATCH) != 0;
    if (leaf_batch) {
        if (fslab)
            reg ? launch_persistent<2, true, true, kLeafDiv>(B, L, groups, spill, cap, ctl)
                : launch_persistent<2, false, true, kLeafDiv>(B, L, groups, spill, cap, ctl);
        else
            reg ? launch_persistent<2, true, false, kLeafDiv>(B, L, groups, spill, cap, ctl)
                : launch_persistent<2, false, false, kLeafDiv>(B, L, groups, spill, cap, ctl);
    } else if (fslab) {
        reg ? launch_persistent<2, true, true>(B, L, groups, spill, cap, ctl)
            : launch_persistent<2, false, true>(B, L, groups, spill, cap, ctl);
    } else {
        reg ? launch_persistent<2, true>(B, L, groups, spill, cap, ctl)
            : launch_persistent<2, false>(B, L, groups, spill, cap, ctl);
    }
    ATLAS_HIP_TRY(hipGetLastError());
    return ATLAS_RT_OK;
}

atlas_rt_status launch_blas_any(atlas_rt_blas& blas, const TraceLaunch& L) {
    if (L.count == 0) return ATLAS_RT_OK;
    const uint32_t grid = (L.count + kBlock - 1) / kBlock;
    uint32_t cap = 0;
    atlas_rt_status st = prepare(blas, grid * kBlock, kLdsDepth, cap);
    if (st != ATLAS_RT_OK) return st;
    st = blas.timing.begin(L.stream);
    if (st != ATLAS_RT_OK) return st;
    hipLaunchKernelGGL((k_blas_any<kBlock, kLdsDepth>), dim3(grid), dim3(kBlock), 0, L.stream, blas_dev(blas),
                       reinterpret_cast<const float4*>(L.rays), reinterpret_cast<uint32_t*>(L.hits), L.count,
                       reinterpret_cast<int32_t*>(blas.spill.ptr), cap, reinterpret_cast<uint32_t*>(blas.flags.ptr));
    ATLAS_HIP_TRY(hipGetLastError());
    return blas.timing.end(L.stream);
}

}  // namespace atlas_rt
