// Batched traversal: many (BLAS, ray range) pairs in one launch, gfx950.
//
// Work item w in [0, total) belongs to the batch b with prefix[b] <= w <
// prefix[b+1] (binary search over the exclusive prefix of the batch sizes);
// its ray is rays[batch[b].ray_offset + (w - prefix[b])] and it is traced
// against blases[batch[b].blas] with one of the step functions of
// trace_core.h: binary nodes in engine order, or 4- / 8-wide nodes, each
// with the stack in LDS or with its top in registers.
//
// Launch shapes (picked by the autotuner):
//  * persistent<BLOCK>: resident waves pull work items from a global counter
//    and refill idle lanes (ballot + popcount) once fewer than `refill` lanes
//    of the wave are live; each lane carries its own BLAS descriptor, so a
//    wave may mix batches;
//  * grid<BLOCK>: one work item per lane.
// The total may be capped by a device-side count so that one captured graph
// serves uneven per-bounce ray counts.
#include <hip/hip_runtime.h>

#include "batch.h"
#include "trace_core.h"

namespace atlas_rt {

namespace {

constexpr int kWave = 64;

__device__ __forceinline__ uint32_t find_batch(const uint32_t* __restrict__ prefix, uint32_t n, uint32_t w) {
    uint32_t lo = 0, hi = n;  // invariant: prefix[lo] <= w < prefix[hi]
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (prefix[mid] <= w)
            lo = mid;
        else
            hi = mid;
    }
    return lo;
}

template <int BLOCK, int LDS_DEPTH, int NODE_W, bool REG, bool FLOAT_SLAB, int LEAF_DIV>
__global__ __launch_bounds__(BLOCK) void k_batch_persistent(BatchDev D, const uint32_t* __restrict__ d_count,
                                                            const float4* __restrict__ rays,
                                                            float4* __restrict__ hits, int32_t* __restrict__ spill,
                                                            uint32_t cap, uint32_t* __restrict__ ctl, uint32_t refill) {
    __shared__ int32_t lds[LDS_DEPTH * BLOCK];
    const uint32_t count = d_count ? min(*d_count, D.total) : D.total;
    const uint32_t lane = threadIdx.x & (kWave - 1);
    const LaneStack<BLOCK, LDS_DEPTH> S{lds, spill, threadIdx.x, blockIdx.x * BLOCK + threadIdx.x, gridDim.x * BLOCK};
    LaneState L;
    L.slot = -1;
    BlasDev B;
    bool active = false;
    bool exhausted = false;
    uint32_t ray = 0;
    float tmax = 0.0f;
    bool pending = false;  // the lane's ray has ended; its hit record is not written yet
    // The lane's current batch: work items [cur_lo, cur_hi) are its rays from
    // cur_ray0, traced against B.  A wave pulls consecutive work items, so a
    // new ray is nearly always in the batch of the lane's last one, and the
    // binary search over the batch table (a chain of dependent loads) and the
    // BLAS record load are skipped.
    uint32_t cur_lo = 1, cur_hi = 0, cur_ray0 = 0;
    RangePull rp;  // work items from kPullRanges ranges, counters ctl[kRangeWord * (r + 1)]
    rp.init(count);
    for (;;) {
        // hit records of the rays ended since the last refill, all at once
        // (the FP64 refinement runs once per refill round)
        if (pending) {
            hits[ray] = lane_hit_record(B, L, tmax);
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
                    const uint32_t w = base + __popcll(idle & ((1ull << lane) - 1));
                    if (w < end) {
                        if (w < cur_lo || w >= cur_hi) {
                            const uint32_t b = find_batch(D.prefix, D.batch_count, w);
                            const uint4 desc = D.batches[b];
                            cur_lo = D.prefix[b];
                            cur_hi = D.prefix[b + 1];
                            cur_ray0 = desc.y;
                            B = ld_blas(D.blases + desc.x);
                        }
                        ray = cur_ray0 + (w - cur_lo);
                        const float4 r0 = rays[2 * size_t(ray)];
                        const float4 r1 = rays[2 * size_t(ray) + 1];
                        tmax = r1.w;
                        if (any_nan3(r1.x, r1.y, r1.z)) {
                            hits[ray] = make_float4(tmax, 0.0f, 0.0f, __int_as_float(-1));
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
    if (lane == 0) {
        __threadfence();
        const uint32_t waves = gridDim.x * (BLOCK / kWave);
        if (atomicAdd(&ctl[1], 1u) == waves - 1) {
            for (uint32_t r = 0; r < kPullRanges; r++) atomicExch(&ctl[kRangeWord * (r + 1)], 0u);
            atomicExch(&ctl[1], 0u);
        }
    }
}

template <int BLOCK, int LDS_DEPTH, int NODE_W, bool REG, bool FLOAT_SLAB>
__global__ __launch_bounds__(BLOCK) void k_batch_grid(BatchDev D, const uint32_t* __restrict__ d_count,
                                                      const float4* __restrict__ rays, float4* __restrict__ hits,
                                                      int32_t* __restrict__ spill, uint32_t cap,
                                                      uint32_t* __restrict__ ctl) {
    __shared__ int32_t lds[LDS_DEPTH * BLOCK];
    const uint32_t count = d_count ? min(*d_count, D.total) : D.total;
    const uint32_t w = blockIdx.x * BLOCK + threadIdx.x;
    if (w >= count) return;
    const uint32_t b = find_batch(D.prefix, D.batch_count, w);
    const uint4 desc = D.batches[b];
    const uint32_t ray = desc.y + (w - D.prefix[b]);
    const BlasDev B = ld_blas(D.blases + desc.x);
    const float4 r0 = rays[2 * size_t(ray)];
    const float4 r1 = rays[2 * size_t(ray) + 1];
    LaneState L;
    L.slot = -1;
    if (!any_nan3(r1.x, r1.y, r1.z)) {
        lane_begin(L, B, r0, r1, NODE_W > 2);
        const LaneStack<BLOCK, LDS_DEPTH> S{lds, spill, threadIdx.x, w, gridDim.x * BLOCK};
        uint32_t dn = 0, dt = 0;
        int r;
        do {
            r = step_variant<NODE_W, REG, BLOCK, LDS_DEPTH, false, FLOAT_SLAB>(B, L, S, cap, dn, dt);
        } while (r == kContinue);
        if (r == kOverflow) {
            atomicOr(&ctl[2], 1u);
            L.slot = -1;
        }
    }
    hits[ray] = lane_hit_record(B, L, r1.w);
}

// Leaf-batched waves take triangle steps once a quarter of their live lanes
// wait at a leaf.
constexpr int kLeafDiv = 4;

// LDS stack rows per block size (LDS bytes = rows * BLOCK * 4).
template <int BLOCK>
struct Rows {
    static constexpr int value = BLOCK >= 256 ? 24 : 32;
};

template <int BLOCK, int NODE_W, bool REG, bool FLOAT_SLAB = false, int LEAF_DIV = 0>
void launch_kernel(atlas_rt_batch_set& s, const BatchLaunch& L, uint32_t grid, uint32_t cap) {
    constexpr int D = Rows<BLOCK>::value;
    uint32_t* ctl = reinterpret_cast<uint32_t*>(s.flags.ptr);
    if (L.persistent)
        hipLaunchKernelGGL((k_batch_persistent<BLOCK, D, NODE_W, REG, FLOAT_SLAB, LEAF_DIV>), dim3(grid), dim3(BLOCK), 0,
                           L.stream, s.dev(),
                           L.d_count, reinterpret_cast<const float4*>(L.rays), reinterpret_cast<float4*>(L.hits),
                           reinterpret_cast<int32_t*>(s.spill.ptr), cap, ctl, L.refill);
    else
        hipLaunchKernelGGL((k_batch_grid<BLOCK, D, NODE_W, REG, FLOAT_SLAB>), dim3(grid), dim3(BLOCK), 0, L.stream,
                           s.dev(),
                           L.d_count, reinterpret_cast<const float4*>(L.rays), reinterpret_cast<float4*>(L.hits),
                           reinterpret_cast<int32_t*>(s.spill.ptr), cap, ctl);
}

template <int BLOCK>
atlas_rt_status launch_one(atlas_rt_batch_set& s, const BatchLaunch& L) {
    constexpr int D = Rows<BLOCK>::value;
    uint32_t threads, grid;
    if (L.persistent) {
        grid = uint32_t(device_cu_count()) * (1024u / BLOCK);  // 16 resident waves per CU
        threads = grid * BLOCK;
    } else {
        grid = (s.total + BLOCK - 1) / BLOCK;
        threads = grid * BLOCK;
    }
    const uint32_t cap = variant_cap(L.node_w, s.max_depth + 1, s.max_cap4, s.max_cap8);
    if (cap > uint32_t(D)) {
        atlas_rt_status st = s.spill.ensure(size_t(cap - D) * threads * sizeof(int32_t));
        if (st != ATLAS_RT_OK) return st;
    }
    if (!L.persistent && L.node_w != 2)
        return fail(ATLAS_RT_E_UNSUPPORTED, "batched trace: wide nodes run in resident waves only");
    if (L.node_w == 2 && L.leaf_batch && L.persistent) {
        if (L.fslab)
            L.reg ? launch_kernel<BLOCK, 2, true, true, kLeafDiv>(s, L, grid, cap)
                  : launch_kernel<BLOCK, 2, false, true, kLeafDiv>(s, L, grid, cap);
        else
            L.reg ? launch_kernel<BLOCK, 2, true, false, kLeafDiv>(s, L, grid, cap)
                  : launch_kernel<BLOCK, 2, false, false, kLeafDiv>(s, L, grid, cap);
    } else if (L.node_w == 2 && L.fslab)
        L.reg ? launch_kernel<BLOCK, 2, true, true>(s, L, grid, cap)
              : launch_kernel<BLOCK, 2, false, true>(s, L, grid, cap);
    else if (L.node_w == 2)
        L.reg ? launch_kernel<BLOCK, 2, true>(s, L, grid, cap) : launch_kernel<BLOCK, 2, false>(s, L, grid, cap);
    else if (L.node_w == 4)
        L.reg ? launch_kernel<BLOCK, 4, true>(s, L, grid, cap) : launch_kernel<BLOCK, 4, false>(s, L, grid, cap);
    else
        L.reg ? launch_kernel<BLOCK, 8, true>(s, L, grid, cap) : launch_kernel<BLOCK, 8, false>(s, L, grid, cap);
    ATLAS_HIP_TRY(hipGetLastError());
    return ATLAS_RT_OK;
}

}  // namespace

atlas_rt_status launch_batch(atlas_rt_batch_set& s, const BatchLaunch& L) {
    if (s.total == 0) return ATLAS_RT_OK;
    switch (L.block) {
        case 64:
            return launch_one<64>(s, L);
        case 128:
            return launch_one<128>(s, L);
        case 256:
            return launch_one<256>(s, L);
        default:
            return fail(ATLAS_RT_E_INVALID_ARG, "batched trace: block must be 64, 128 or 256");
    }
}

}  // namespace atlas_rt
