// Ray coherence sorting (direction bins), gfx950: one counting-sort pass over
// 64 bins.  A block owns a tile of 4096 consecutive records: per-tile bin
// counts, one exclusive scan of the (bin, tile) counts in bin-major order, then
// a stable scatter that ranks the lanes of a wave holding the same bin with six
// ballots.
#include "binning.h"

#pragma clang fp contract(off)

namespace atlas_rt {

namespace {

__device__ __forceinline__ float bin_saturate(float x) { return fminf(fmaxf(x, 0.0f), 1.0f); }

// The direction's bin (tracing.hsh:18-21 with the octahedral map of
// octahedron.hsh:23-37); cells clamped to the 8x8 grid, NaN -> bin 0.
__device__ __forceinline__ uint32_t ray_bin(float x, float y, float z) {
    const float l1 = (fabsf(x) + fabsf(y)) + fabsf(z);
    float ox = x / l1, oz = z / l1;
    if (y < 0.0f) {
        const float a = ox, b = oz;
        ox = (a >= 0.0f ? 1.0f : -1.0f) * (1.0f - fabsf(b));
        oz = (b >= 0.0f ? 1.0f : -1.0f) * (1.0f - fabsf(a));
    }
    const float u = bin_saturate(0.5f * ox + 0.5f), v = bin_saturate(0.5f * oz + 0.5f);
    const uint32_t cx = uint32_t(u * 8.0f), cy = uint32_t(v * 8.0f);
    return (cy < 7u ? cy : 7u) * 8u + (cx < 7u ? cx : 7u);
}

constexpr int kB = 256;
constexpr int kItems = 16;
constexpr uint32_t kTile = kB * kItems;
constexpr int kScanThreads = 1024;

__device__ __forceinline__ uint32_t record_count(const uint32_t* d_count, uint32_t count, uint32_t cap) {
    const uint32_t n = d_count ? *d_count : count;
    return n < cap ? n : cap;
}

// hist[bin * tiles + tile]
__global__ __launch_bounds__(kB) void k_bin_hist(const float4* __restrict__ in, int rec, const uint32_t* d_count,
                                                  uint32_t count, uint32_t cap, uint32_t tiles,
                                                  uint32_t* __restrict__ hist) {
    __shared__ uint32_t h[kRayBins];
    if (threadIdx.x < kRayBins) h[threadIdx.x] = 0;
    __syncthreads();
    const uint32_t n = record_count(d_count, count, cap);
    const size_t base = size_t(blockIdx.x) * kTile;
    for (int r = 0; r < kItems; r++) {
        const size_t i = base + size_t(r) * kB + threadIdx.x;
        if (i < n) {
            const float4 d = in[size_t(rec) * i + 1];
            atomicAdd(&h[ray_bin(d.x, d.y, d.z)], 1u);
        }
    }
    __syncthreads();
    if (threadIdx.x < kRayBins) hist[size_t(threadIdx.x) * tiles + blockIdx.x] = h[threadIdx.x];
}

// In-place exclusive scan of the 64 * tiles counts by one block; the bin
// offsets are the scanned values at each bin's first tile.
__global__ __launch_bounds__(kScanThreads) void k_bin_scan(uint32_t* __restrict__ hist, uint32_t tiles,
                                                           const uint32_t* d_count, uint32_t count, uint32_t cap,
                                                           uint32_t* __restrict__ bin_offsets) {
    __shared__ uint32_t part[kScanThreads];
    const uint32_t total = kRayBins * tiles;
    const uint32_t per = (total + kScanThreads - 1) / kScanThreads;
    const uint32_t lo = min(total, threadIdx.x * per), hi = min(total, lo + per);
    uint32_t sum = 0;
    for (uint32_t i = lo; i < hi; i++) sum += hist[i];
    part[threadIdx.x] = sum;
    __syncthreads();
    for (uint32_t off = 1; off < kScanThreads; off <<= 1) {
        const uint32_t add = threadIdx.x >= off ? part[threadIdx.x - off] : 0u;
        __syncthreads();
        part[threadIdx.x] += add;
        __syncthreads();
    }
    uint32_t run = part[threadIdx.x] - sum;
    for (uint32_t i = lo; i < hi; i++) {
        const uint32_t x = hist[i];
        hist[i] = run;
        if (bin_offsets && i % tiles == 0) bin_offsets[i / tiles] = run;
        run += x;
    }
    if (bin_offsets && threadIdx.x == 0) bin_offsets[kRayBins] = record_count(d_count, count, cap);
}

__global__ __launch_bounds__(kB) void k_bin_scatter(const float4* __restrict__ in, float4* __restrict__ out, int rec,
                                                    const uint32_t* d_count, uint32_t count, uint32_t cap,
                                                    uint32_t tiles, const uint32_t* __restrict__ offsets,
                                                    uint32_t* __restrict__ perm) {
    constexpr int kWaves = kB / 64;
    __shared__ uint32_t running[kRayBins];  // records of each bin already placed from this tile
    __shared__ uint32_t wave_cnt[kWaves][kRayBins];
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    const uint32_t n = record_count(d_count, count, cap);
    if (threadIdx.x < kRayBins) running[threadIdx.x] = offsets[size_t(threadIdx.x) * tiles + blockIdx.x];
    const size_t base = size_t(blockIdx.x) * kTile;
    for (int r = 0; r < kItems; r++) {
        const size_t i = base + size_t(r) * kB + threadIdx.x;
        const bool valid = i < n;
        uint32_t bin = 0;
        if (valid) {
            const float4 d = in[size_t(rec) * i + 1];
            bin = ray_bin(d.x, d.y, d.z);
        }
        // lanes of this wave holding the same bin
        uint64_t same = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 6; b++) {
            const uint64_t ones = __ballot(valid && ((bin >> b) & 1u));
            same &= ((bin >> b) & 1u) ? ones : ~ones;
        }
        for (int d = threadIdx.x; d < kWaves * int(kRayBins); d += kB) (&wave_cnt[0][0])[d] = 0;
        __syncthreads();
        const uint32_t rank = __popcll(same & ((1ull << lane) - 1ull));
        if (valid && rank == 0) wave_cnt[wave][bin] = __popcll(same);
        __syncthreads();
        if (valid) {
            uint32_t dst = running[bin] + rank;
            for (uint32_t w = 0; w < wave; w++) dst += wave_cnt[w][bin];
            for (int k = 0; k < rec; k++) out[size_t(rec) * dst + k] = in[size_t(rec) * i + k];
            if (perm) perm[dst] = uint32_t(i);
        }
        __syncthreads();
        if (threadIdx.x < kRayBins) {
            uint32_t add = 0;
            for (int w = 0; w < kWaves; w++) add += wave_cnt[w][threadIdx.x];
            running[threadIdx.x] += add;
        }
        __syncthreads();
    }
}

uint32_t tiles_for(uint32_t cap) { return uint32_t((size_t(cap) + kTile - 1) / kTile); }

}  // namespace

size_t bin_sort_workspace(uint32_t cap) { return size_t(kRayBins) * (tiles_for(cap) ? tiles_for(cap) : 1) * 4; }

// Bin offsets of an empty sort: all zero.  A kernel rather than a memset, so
// that a captured graph of the call holds no memset node.
__global__ void k_bin_offsets_zero(uint32_t* __restrict__ bin_offsets) {
    if (threadIdx.x <= kRayBins) bin_offsets[threadIdx.x] = 0;
}

hipError_t launch_bin_sort(const float4* in, float4* out, int rec, const uint32_t* d_count, uint32_t count,
                           uint32_t cap, void* workspace, uint32_t* perm, uint32_t* bin_offsets, hipStream_t s) {
    const uint32_t tiles = tiles_for(cap);
    if (tiles == 0) {
        if (bin_offsets) hipLaunchKernelGGL(k_bin_offsets_zero, dim3(1), dim3(128), 0, s, bin_offsets);
        return hipGetLastError();
    }
    uint32_t* hist = static_cast<uint32_t*>(workspace);
    hipLaunchKernelGGL(k_bin_hist, dim3(tiles), dim3(kB), 0, s, in, rec, d_count, count, cap, tiles, hist);
    hipLaunchKernelGGL(k_bin_scan, dim3(1), dim3(kScanThreads), 0, s, hist, tiles, d_count, count, cap, bin_offsets);
    hipLaunchKernelGGL(k_bin_scatter, dim3(tiles), dim3(kB), 0, s, in, out, rec, d_count, count, cap, tiles, hist,
                       perm);
    return hipGetLastError();
}

}  // namespace atlas_rt
