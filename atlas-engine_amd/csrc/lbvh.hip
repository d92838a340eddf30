// GPU linear-BVH builder, gfx950.
//
// Per mesh (all meshes of a batch in the same launches):
//   1. centroids ((v0 + v1) + v2) / 3 and their bounds (ordered-integer atomics);
//   2. 30-bit Morton codes on a 1024^3 grid over the centroid bounds, keys
//      (mesh << 54) | (morton << 24) | triangle index, so every key is unique
//      and the sort is deterministic;
//   3. a least-significant-digit radix sort of the 64-bit keys (8 passes of
//      8 bits: block histograms, a three-launch scan, stable scatter ranked
//      by ballots);
//   4. the binary radix tree over each mesh's sorted keys (Karras, "Maximizing
//      parallelism in the construction of BVHs, octrees, and k-d trees", HPG
//      2012): one thread per inner node finds its key range and split, and
//      with it the node's leaf count;
//   5. boxes level by level: each inner node's depth by pointer jumping over
//      the parents, then one launch per depth from the deepest up, so no
//      launch reads a box another thread of it writes (no atomics, no fences);
//   6. the engine's flattening (BVH.cpp:408-441): the child with the larger
//      surface area goes left (ties keep the radix-tree order), inner nodes
//      numbered in pre-order, one triangle per leaf, leaf pointer ~slot.
//      Each node's pre-order index, leaf slot and depth are sums along its
//      root path, computed by pointer jumping (7 rounds cover depth < 128;
//      radix-tree depth is at most 64, one level per key bit);
//   7. the engine node records and the triangle records of every slot.
// Every float operation is written in a fixed order without contraction, so
// the CPU restatement in the oracle reproduces the arrays bit for bit.
#include "lbvh.h"

#include <algorithm>
#include <cstring>

#pragma clang fp contract(off)

namespace atlas_rt {

namespace {

constexpr int kB = 256;
constexpr uint32_t kIndexBits = 24;
constexpr uint32_t kMaxMeshes = 1024;
constexpr int kJumpRounds = 7;

inline uint32_t blocks(size_t n) { return uint32_t((n + kB - 1) / kB); }

// Largest m < k with off[m] <= g (off has k + 1 entries, off[0] = 0).
__device__ __forceinline__ uint32_t find_mesh(const uint32_t* __restrict__ off, uint32_t k, uint32_t g) {
    uint32_t lo = 0, hi = k;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (off[mid] <= g)
            lo = mid;
        else
            hi = mid;
    }
    return lo;
}

// Order-preserving float <-> int map for atomic min / max.
__device__ __forceinline__ int f2o(float f) {
    const int i = __float_as_int(f);
    return i >= 0 ? i : i ^ 0x7FFFFFFF;
}
__device__ __forceinline__ float o2f(int i) { return __int_as_float(i >= 0 ? i : i ^ 0x7FFFFFFF); }

__device__ __forceinline__ uint32_t expand_bits(uint32_t x) {
    x = (x * 0x00010001u) & 0xFF0000FFu;
    x = (x * 0x00000101u) & 0x0F00F00Fu;
    x = (x * 0x00000011u) & 0xC30C30C3u;
    x = (x * 0x00000005u) & 0x49249249u;
    return x;
}

__device__ __forceinline__ float box_area(float4 lo, float4 hi) {
    const float dx = hi.x - lo.x, dy = hi.y - lo.y, dz = hi.z - lo.z;
    return 2.0f * ((dx * dy + dy * dz) + dz * dx);
}

struct Batch {
    const float* verts;        // 9 floats per triangle, meshes concatenated
    const uint32_t* leaf_off;  // k + 1: first triangle of each mesh
    const uint32_t* inner_off; // k + 1: first inner node of each mesh (leaf_off[m] - m)
    uint32_t k;
    uint32_t total;            // triangles
    uint32_t inner;            // inner nodes
};

__global__ void k_fill_i32(int* p, uint32_t n, int v) {
    const uint32_t i = blockIdx.x * kB + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ __launch_bounds__(kB) void k_centroids(Batch B, float4* __restrict__ cen, int* __restrict__ bounds) {
    // lanes past the end take the last triangle's values, which leaves every
    // bound unchanged, so whole waves stay together for the reduction
    const uint32_t g0 = blockIdx.x * kB + threadIdx.x;
    const uint32_t g = g0 < B.total ? g0 : B.total - 1;
    const uint32_t m = find_mesh(B.leaf_off, B.k, g);
    const float* v = B.verts + 9 * size_t(g);
    const float cx = ((v[0] + v[3]) + v[6]) / 3.0f;
    const float cy = ((v[1] + v[4]) + v[7]) / 3.0f;
    const float cz = ((v[2] + v[5]) + v[8]) / 3.0f;
    if (g0 < B.total) cen[g] = make_float4(cx, cy, cz, 0.0f);
    int v6[6] = {f2o(cx), f2o(cy), f2o(cz), f2o(cx), f2o(cy), f2o(cz)};
    int* b = bounds + 6 * m;
    // A wave whose lanes all belong to one mesh reduces first: one atomic
    // per bound and wave instead of one per triangle.
    const uint32_t m0 = __shfl(m, 0);
    if (__all(m == m0)) {
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1)
#pragma unroll
            for (int a = 0; a < 6; a++) {
                const int o = __shfl_xor(v6[a], off);
                v6[a] = a < 3 ? min(v6[a], o) : max(v6[a], o);
            }
        if ((threadIdx.x & 63u) != 0) return;
    }
    for (int a = 0; a < 3; a++) atomicMin(&b[a], v6[a]);
    for (int a = 3; a < 6; a++) atomicMax(&b[a], v6[a]);
}

__device__ __forceinline__ uint32_t quantize(float c, float lo, float hi) {
    const float e = hi - lo;
    const float s = e > 0.0f ? 1024.0f / e : 0.0f;
    const float q = (c - lo) * s;
    const uint32_t u = q > 0.0f ? uint32_t(q) : 0u;
    return u < 1023u ? u : 1023u;
}

__global__ __launch_bounds__(kB) void k_keys(Batch B, const float4* __restrict__ cen, const int* __restrict__ bounds,
                                             uint64_t* __restrict__ keys) {
    const uint32_t g = blockIdx.x * kB + threadIdx.x;
    if (g >= B.total) return;
    const uint32_t m = find_mesh(B.leaf_off, B.k, g);
    const int* b = bounds + 6 * m;
    const float4 c = cen[g];
    const uint32_t qx = quantize(c.x, o2f(b[0]), o2f(b[3]));
    const uint32_t qy = quantize(c.y, o2f(b[1]), o2f(b[4]));
    const uint32_t qz = quantize(c.z, o2f(b[2]), o2f(b[5]));
    const uint32_t morton = (expand_bits(qx) << 2) | (expand_bits(qy) << 1) | expand_bits(qz);
    keys[g] = (uint64_t(m) << 54) | (uint64_t(morton) << kIndexBits) | uint64_t(g - B.leaf_off[m]);
}

// Common-prefix length of sorted keys a and b of one mesh (-1 outside it).
__device__ __forceinline__ int delta(const uint64_t* __restrict__ K, int n, int a, int b) {
    if (b < 0 || b >= n) return -1;
    return __clzll(K[a] ^ K[b]);
}

// Children: >= 0 inner node (global), < 0 ~leaf (global sorted position).
__global__ __launch_bounds__(kB) void k_karras(Batch B, const uint64_t* __restrict__ sorted, int2* __restrict__ child,
                                               int* __restrict__ parent_inner, int* __restrict__ parent_leaf,
                                               uint32_t* __restrict__ cnt) {
    const uint32_t g = blockIdx.x * kB + threadIdx.x;
    if (g >= B.inner) return;
    const uint32_t m = find_mesh(B.inner_off, B.k, g);
    const int i = int(g - B.inner_off[m]);
    const int n = int(B.leaf_off[m + 1] - B.leaf_off[m]);
    const uint64_t* K = sorted + B.leaf_off[m];
    const int d = delta(K, n, i, i + 1) - delta(K, n, i, i - 1) >= 0 ? 1 : -1;
    const int dmin = delta(K, n, i, i - d);
    int lmax = 2;
    while (delta(K, n, i, i + lmax * d) > dmin) lmax <<= 1;
    int l = 0;
    for (int t = lmax >> 1; t >= 1; t >>= 1)
        if (delta(K, n, i, i + (l + t) * d) > dmin) l += t;
    const int j = i + l * d;
    const int dnode = delta(K, n, i, j);
    int s = 0;
    int t = l;
    do {
        t = (t + 1) >> 1;
        if (delta(K, n, i, i + (s + t) * d) > dnode) s += t;
    } while (t > 1);
    const int gamma = i + s * d + (d < 0 ? -1 : 0);
    const int lo = i < j ? i : j, hi = i < j ? j : i;
    const int ib = int(B.inner_off[m]), lb = int(B.leaf_off[m]);
    const int c0 = lo == gamma ? ~(lb + gamma) : ib + gamma;
    const int c1 = hi == gamma + 1 ? ~(lb + gamma + 1) : ib + gamma + 1;
    child[g] = make_int2(c0, c1);
    cnt[g] = uint32_t(hi - lo + 1);  // leaves under g
    if (c0 >= 0)
        parent_inner[c0] = int(g);
    else
        parent_leaf[~c0] = int(g);
    if (c1 >= 0)
        parent_inner[c1] = int(g);
    else
        parent_leaf[~c1] = int(g);
}

__device__ __forceinline__ void child_box(int c, const float4* leaf_lo, const float4* leaf_hi,
                                          const float4* inner_lo, const float4* inner_hi, const uint32_t* cnt,
                                          float4& lo, float4& hi, uint32_t& n) {
    if (c < 0) {
        lo = leaf_lo[~c];
        hi = leaf_hi[~c];
        n = 1;
    } else {
        lo = inner_lo[c];
        hi = inner_hi[c];
        n = cnt[c];
    }
}

// Leaf boxes in sorted order.
__global__ __launch_bounds__(kB) void k_leaf_boxes(Batch B, const uint64_t* __restrict__ sorted,
                                                   float4* __restrict__ leaf_lo, float4* __restrict__ leaf_hi) {
    const uint32_t x = blockIdx.x * kB + threadIdx.x;
    if (x >= B.total) return;
    const uint32_t m = find_mesh(B.leaf_off, B.k, x);
    const uint32_t tri = B.leaf_off[m] + uint32_t(sorted[x] & ((1ull << kIndexBits) - 1));
    const float* v = B.verts + 9 * size_t(tri);
    leaf_lo[x] = make_float4(fminf(fminf(v[0], v[3]), v[6]), fminf(fminf(v[1], v[4]), v[7]),
                             fminf(fminf(v[2], v[5]), v[8]), 0.0f);
    leaf_hi[x] = make_float4(fmaxf(fmaxf(v[0], v[3]), v[6]), fmaxf(fmaxf(v[1], v[4]), v[7]),
                             fmaxf(fmaxf(v[2], v[5]), v[8]), 0.0f);
}

// Inner-node depth by pointer jumping over the parents: (depth so far, ancestor).
__global__ __launch_bounds__(kB) void k_depth_init(uint32_t inner, const int* __restrict__ parent_inner,
                                                   int2* __restrict__ d) {
    const uint32_t g = blockIdx.x * kB + threadIdx.x;
    if (g >= inner) return;
    const int p = parent_inner[g];
    d[g] = make_int2(p >= 0 ? 1 : 0, p);
}

__global__ __launch_bounds__(kB) void k_depth_jump(uint32_t inner, const int2* __restrict__ in, int2* __restrict__ out,
                                                   uint32_t* __restrict__ max_depth) {
    const uint32_t g0 = blockIdx.x * kB + threadIdx.x;
    const uint32_t g = g0 < inner ? g0 : inner - 1;
    const int2 s = in[g];
    const int2 r = s.y < 0 ? s : make_int2(s.x + in[s.y].x, in[s.y].y);
    if (g0 < inner) out[g] = r;
    if (max_depth) {
        uint32_t v = uint32_t(r.x);
        for (int off = 32; off >= 1; off >>= 1) v = max(v, uint32_t(__shfl_xor(int(v), off)));
        if ((threadIdx.x & 63u) == 0) atomicMax(max_depth, v);
    }
}

// Boxes of the inner nodes at one depth (their children are final).
__global__ __launch_bounds__(kB) void k_level_boxes(uint32_t inner, int level, const int2* __restrict__ depth,
                                                    const int2* __restrict__ child, const float4* leaf_lo,
                                                    const float4* leaf_hi, float4* inner_lo, float4* inner_hi) {
    const uint32_t g = blockIdx.x * kB + threadIdx.x;
    if (g >= inner || depth[g].x != level) return;
    const int2 c = child[g];
    const float4 lo0 = c.x < 0 ? leaf_lo[~c.x] : inner_lo[c.x], hi0 = c.x < 0 ? leaf_hi[~c.x] : inner_hi[c.x];
    const float4 lo1 = c.y < 0 ? leaf_lo[~c.y] : inner_lo[c.y], hi1 = c.y < 0 ? leaf_hi[~c.y] : inner_hi[c.y];
    inner_lo[g] = make_float4(fminf(lo0.x, lo1.x), fminf(lo0.y, lo1.y), fminf(lo0.z, lo1.z), 0.0f);
    inner_hi[g] = make_float4(fmaxf(hi0.x, hi1.x), fmaxf(hi0.y, hi1.y), fmaxf(hi0.z, hi1.z), 0.0f);
}

// The engine's child order of inner node p: the larger surface area first,
// the radix tree's order on a tie.
__device__ __forceinline__ void ordered_children(int2 c, const float4* leaf_lo, const float4* leaf_hi,
                                                 const float4* inner_lo, const float4* inner_hi, const uint32_t* cnt,
                                                 int& first, int& second, float4 (&box)[4], uint32_t& n_first) {
    uint32_t n0, n1;
    float4 lo0, hi0, lo1, hi1;
    child_box(c.x, leaf_lo, leaf_hi, inner_lo, inner_hi, cnt, lo0, hi0, n0);
    child_box(c.y, leaf_lo, leaf_hi, inner_lo, inner_hi, cnt, lo1, hi1, n1);
    const bool swap = box_area(lo0, hi0) < box_area(lo1, hi1);
    first = swap ? c.y : c.x;
    second = swap ? c.x : c.y;
    box[0] = swap ? lo1 : lo0;
    box[1] = swap ? hi1 : hi0;
    box[2] = swap ? lo0 : lo1;
    box[3] = swap ? hi0 : hi1;
    n_first = swap ? n1 : n0;
}

// Pointer jumping state per node (inner nodes first, then leaves):
// x = pre-order offset, y = leaf-slot offset, z = depth, w = ancestor (-1 at a root).
__global__ __launch_bounds__(kB) void k_jump_init(Batch B, const int2* __restrict__ child,
                                                  const int* __restrict__ parent_inner,
                                                  const int* __restrict__ parent_leaf, const float4* leaf_lo,
                                                  const float4* leaf_hi, const float4* inner_lo, const float4* inner_hi,
                                                  const uint32_t* cnt, int4* __restrict__ st) {
    const uint32_t x = blockIdx.x * kB + threadIdx.x;
    if (x >= B.inner + B.total) return;
    const bool is_inner = x < B.inner;
    const int p = is_inner ? parent_inner[x] : parent_leaf[x - B.inner];
    if (p < 0) {
        st[x] = make_int4(0, 0, 0, -1);
        return;
    }
    int first, second;
    float4 box[4];
    uint32_t n_first;
    ordered_children(child[p], leaf_lo, leaf_hi, inner_lo, inner_hi, cnt, first, second, box, n_first);
    const int me = is_inner ? int(x) : ~int(x - B.inner);
    if (me == first)
        st[x] = make_int4(1, 0, 1, p);
    else
        st[x] = make_int4(int(n_first), int(n_first), 1, p);  // 1 + (n_first - 1) inner nodes precede it
}

__global__ __launch_bounds__(kB) void k_jump(uint32_t n, const int4* __restrict__ in, int4* __restrict__ out) {
    const uint32_t x = blockIdx.x * kB + threadIdx.x;
    if (x >= n) return;
    const int4 s = in[x];
    if (s.w < 0) {
        out[x] = s;
        return;
    }
    const int4 a = in[s.w];
    out[x] = make_int4(s.x + a.x, s.y + a.y, s.z + a.z, a.w);
}

__global__ __launch_bounds__(kB) void k_write_nodes(Batch B, const LbvhMeshOut* __restrict__ outs,
                                                    const int2* __restrict__ child, const float4* leaf_lo,
                                                    const float4* leaf_hi, const float4* inner_lo,
                                                    const float4* inner_hi, const uint32_t* cnt,
                                                    const int4* __restrict__ st) {
    const uint32_t g = blockIdx.x * kB + threadIdx.x;
    if (g >= B.inner) return;
    const uint32_t m = find_mesh(B.inner_off, B.k, g);
    int first, second;
    float4 box[4];
    uint32_t n_first;
    ordered_children(child[g], leaf_lo, leaf_hi, inner_lo, inner_hi, cnt, first, second, box, n_first);
    const int lp = first >= 0 ? st[first].x : ~st[B.inner + uint32_t(~first)].y;
    const int rp = second >= 0 ? st[second].x : ~st[B.inner + uint32_t(~second)].y;
    float4* o = outs[m].nodes + 4 * size_t(st[g].x);
    o[0] = make_float4(box[0].x, box[0].y, box[0].z, box[1].x);
    o[1] = make_float4(box[1].y, box[1].z, box[2].x, box[2].y);
    o[2] = make_float4(box[2].z, box[3].x, box[3].y, box[3].z);
    o[3] = make_float4(__int_as_float(lp), __int_as_float(rp), 0.0f, 0.0f);
}

__global__ __launch_bounds__(kB) void k_write_tris(Batch B, const LbvhMeshOut* __restrict__ outs,
                                                   const uint64_t* __restrict__ sorted, const int4* __restrict__ st,
                                                   uint32_t* __restrict__ depth) {
    const uint32_t x0 = blockIdx.x * kB + threadIdx.x;
    const uint32_t x = x0 < B.total ? x0 : B.total - 1;  // whole waves stay for the depth reduction
    const uint32_t m = find_mesh(B.leaf_off, B.k, x);
    const uint32_t local = uint32_t(sorted[x] & ((1ull << kIndexBits) - 1));
    const float* v = B.verts + 9 * (size_t(B.leaf_off[m]) + local);
    const int4 s = st[B.inner + x];
    uint32_t dmax = uint32_t(s.z > 1 ? s.z : 1);
    const uint32_t m0 = __shfl(m, 0);
    if (__all(m == m0)) {
        for (int off = 32; off >= 1; off >>= 1) dmax = max(dmax, uint32_t(__shfl_xor(int(dmax), off)));
        if ((threadIdx.x & 63u) == 0) atomicMax(&depth[m], dmax);
    } else {
        atomicMax(&depth[m], dmax);
    }
    if (x0 >= B.total) return;
    const uint32_t slot = uint32_t(s.y);
    const LbvhMeshOut O = outs[m];
    // traversal layout: (v0, caller index), (v1 - v0, end of leaf), (v2 - v0, 0)
    O.tris[3 * size_t(slot)] = make_float4(v[0], v[1], v[2], __uint_as_float(local));
    O.tris[3 * size_t(slot) + 1] = make_float4(v[3] - v[0], v[4] - v[1], v[5] - v[2], 1.0f);
    O.tris[3 * size_t(slot) + 2] = make_float4(v[6] - v[0], v[7] - v[1], v[8] - v[2], 0.0f);
    // GPUBVHTriangle (MeshData.cpp:242-247): end flag, material 0, opacity 1
    O.vtx[3 * size_t(slot)] = make_float4(v[0], v[1], v[2], 1.0f);
    O.vtx[3 * size_t(slot) + 1] = make_float4(v[3], v[4], v[5], 0.0f);
    O.vtx[3 * size_t(slot) + 2] = make_float4(v[6], v[7], v[8], 1.0f);
    O.orig[slot] = local;
}

// ---------------------------------------------------------------------------
// Exclusive scan of the sort's digit counts: block sums, one small scan,
// then each block scans its chunk (16 consecutive values per thread).
// ---------------------------------------------------------------------------
constexpr uint32_t kScanChunk = kB * 16;

__global__ __launch_bounds__(kB) void k_scan_reduce(const uint32_t* __restrict__ v, uint32_t n,
                                                    uint32_t* __restrict__ part) {
    __shared__ uint32_t red[kB];
    const size_t base = size_t(blockIdx.x) * kScanChunk + size_t(threadIdx.x) * 16;
    uint32_t s = 0;
    for (int i = 0; i < 16; i++)
        if (base + i < n) s += v[base + i];
    red[threadIdx.x] = s;
    __syncthreads();
    for (uint32_t off = kB / 2; off >= 1; off >>= 1) {
        if (threadIdx.x < off) red[threadIdx.x] += red[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// Exclusive scan of the `n` block sums by one block (n <= 1024 * 16).
__global__ __launch_bounds__(1024) void k_scan_parts(uint32_t* __restrict__ part, uint32_t n) {
    __shared__ uint32_t sh[1024];
    const uint32_t lo = min(n, threadIdx.x * 16), hi = min(n, lo + 16);
    uint32_t s = 0;
    for (uint32_t i = lo; i < hi; i++) s += part[i];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (uint32_t off = 1; off < 1024; off <<= 1) {
        const uint32_t add = threadIdx.x >= off ? sh[threadIdx.x - off] : 0u;
        __syncthreads();
        sh[threadIdx.x] += add;
        __syncthreads();
    }
    uint32_t run = sh[threadIdx.x] - s;
    for (uint32_t i = lo; i < hi; i++) {
        const uint32_t x = part[i];
        part[i] = run;
        run += x;
    }
}

__global__ __launch_bounds__(kB) void k_scan_down(uint32_t* __restrict__ v, uint32_t n,
                                                  const uint32_t* __restrict__ part) {
    __shared__ uint32_t sh[kB];
    const size_t base = size_t(blockIdx.x) * kScanChunk + size_t(threadIdx.x) * 16;
    uint32_t x[16];
    uint32_t s = 0;
    for (int i = 0; i < 16; i++) {
        x[i] = base + i < n ? v[base + i] : 0u;
        s += x[i];
    }
    sh[threadIdx.x] = s;
    __syncthreads();
    for (uint32_t off = 1; off < kB; off <<= 1) {
        const uint32_t add = threadIdx.x >= off ? sh[threadIdx.x - off] : 0u;
        __syncthreads();
        sh[threadIdx.x] += add;
        __syncthreads();
    }
    uint32_t run = part[blockIdx.x] + sh[threadIdx.x] - s;
    for (int i = 0; i < 16; i++) {
        if (base + i < n) v[base + i] = run;
        run += x[i];
    }
}

// ---------------------------------------------------------------------------
// LSD radix sort of uint64 keys, 8 bits per pass.  A block owns a tile of
// kSortTile consecutive keys; keys keep their tile order within a digit, so
// each pass is stable.
// ---------------------------------------------------------------------------
constexpr int kSortItems = 16;
constexpr uint32_t kSortTile = kB * kSortItems;

// hist[digit * tiles + tile] = count of `digit` in the tile.
__global__ __launch_bounds__(kB) void k_sort_hist(const uint64_t* __restrict__ keys, uint32_t n, uint32_t shift,
                                                  uint32_t tiles, uint32_t* __restrict__ hist) {
    __shared__ uint32_t h[256];
    h[threadIdx.x] = 0;
    __syncthreads();
    const size_t base = size_t(blockIdx.x) * kSortTile;
    for (int r = 0; r < kSortItems; r++) {
        const size_t i = base + size_t(r) * kB + threadIdx.x;
        if (i < n) atomicAdd(&h[uint32_t(keys[i] >> shift) & 255u], 1u);
    }
    __syncthreads();
    hist[size_t(threadIdx.x) * tiles + blockIdx.x] = h[threadIdx.x];
}

__global__ __launch_bounds__(kB) void k_sort_scatter(const uint64_t* __restrict__ in, uint64_t* __restrict__ out,
                                                     uint32_t n, uint32_t shift, uint32_t tiles,
                                                     const uint32_t* __restrict__ offsets) {
    constexpr int kWaves = kB / 64;
    __shared__ uint32_t running[256];      // keys of each digit already placed from this tile
    __shared__ uint32_t wave_cnt[kWaves][256];
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    running[threadIdx.x] = offsets[size_t(threadIdx.x) * tiles + blockIdx.x];
    const size_t base = size_t(blockIdx.x) * kSortTile;
    for (int r = 0; r < kSortItems; r++) {
        const size_t i = base + size_t(r) * kB + threadIdx.x;
        const bool valid = i < n;
        const uint64_t key = valid ? in[i] : 0ull;
        const uint32_t digit = valid ? uint32_t(key >> shift) & 255u : 256u;
        // lanes of this wave holding the same digit
        uint64_t same = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; b++) {
            const uint64_t ones = __ballot(valid && ((digit >> b) & 1u));
            same &= ((digit >> b) & 1u) ? ones : ~ones;
        }
        for (int d = threadIdx.x; d < kWaves * 256; d += kB) (&wave_cnt[0][0])[d] = 0;
        __syncthreads();
        const uint32_t rank = __popcll(same & ((1ull << lane) - 1ull));
        if (valid && rank == 0) wave_cnt[wave][digit] = __popcll(same);
        __syncthreads();
        if (valid) {
            uint32_t before = running[digit];
            for (uint32_t w = 0; w < wave; w++) before += wave_cnt[w][digit];
            out[before + rank] = key;
        }
        __syncthreads();
        // advance the running offsets by this pass's counts
        uint32_t add = 0;
        for (int w = 0; w < kWaves; w++) add += wave_cnt[w][threadIdx.x];
        running[threadIdx.x] += add;
        __syncthreads();
    }
}

struct EventPair {
    hipEvent_t t0 = nullptr, t1 = nullptr;
    ~EventPair() {
        if (t0) (void)hipEventDestroy(t0);
        if (t1) (void)hipEventDestroy(t1);
    }
};

template <class T>
struct DevBuf {
    T* p = nullptr;
    ~DevBuf() {
        if (p) (void)hipFree(p);
    }
    hipError_t alloc(size_t n) { return hipMalloc(reinterpret_cast<void**>(&p), std::max<size_t>(n, 1) * sizeof(T)); }
};

}  // namespace

atlas_rt_status lbvh_build(const float* verts, const std::vector<uint32_t>& counts,
                           const std::vector<LbvhMeshOut>& outs, std::vector<uint32_t>& depth_out,
                           hipStream_t stream, float* device_ms) {
    const uint32_t k = uint32_t(counts.size());
    if (k == 0 || k > kMaxMeshes || !verts || outs.size() != k)
        return fail(ATLAS_RT_E_INVALID_ARG, "lbvh_build: 1..1024 meshes per call");
    std::vector<uint32_t> leaf_off(k + 1, 0), inner_off(k + 1, 0);
    for (uint32_t m = 0; m < k; m++) {
        if (counts[m] == 0 || counts[m] > (1u << kIndexBits))
            return fail(ATLAS_RT_E_INVALID_ARG, "lbvh_build: 1..2^24 triangles per mesh");
        leaf_off[m + 1] = leaf_off[m] + counts[m];
        inner_off[m + 1] = inner_off[m] + counts[m] - 1;
        if (leaf_off[m + 1] < leaf_off[m]) return fail(ATLAS_RT_E_INVALID_ARG, "lbvh_build: more than 2^32 triangles");
    }
    const uint32_t total = leaf_off[k], inner = inner_off[k], nodes = total + inner;
    DevBuf<uint32_t> d_off;
    DevBuf<LbvhMeshOut> d_outs;
    DevBuf<float4> cen, leaf_lo, leaf_hi, inner_lo, inner_hi;
    DevBuf<int> bounds, parent_inner, parent_leaf;
    DevBuf<uint64_t> keys, sorted;
    DevBuf<int2> child;
    DevBuf<uint32_t> cnt, depth, parts, maxd;
    DevBuf<int2> dep0, dep1;
    DevBuf<int4> st0, st1;
    hipError_t e = d_off.alloc(2 * (k + 1));
    if (e == hipSuccess) e = d_outs.alloc(k);
    if (e == hipSuccess) e = cen.alloc(total);
    if (e == hipSuccess) e = leaf_lo.alloc(total);
    if (e == hipSuccess) e = leaf_hi.alloc(total);
    if (e == hipSuccess) e = inner_lo.alloc(inner);
    if (e == hipSuccess) e = inner_hi.alloc(inner);
    if (e == hipSuccess) e = bounds.alloc(6 * k);
    if (e == hipSuccess) e = parent_inner.alloc(inner);
    if (e == hipSuccess) e = parent_leaf.alloc(total);
    if (e == hipSuccess) e = keys.alloc(total);
    if (e == hipSuccess) e = sorted.alloc(total);
    if (e == hipSuccess) e = child.alloc(inner);
    if (e == hipSuccess) e = cnt.alloc(inner);
    if (e == hipSuccess) e = dep0.alloc(inner);
    if (e == hipSuccess) e = dep1.alloc(inner);
    if (e == hipSuccess) e = maxd.alloc(1);
    if (e == hipSuccess) e = depth.alloc(k);
    if (e == hipSuccess) e = st0.alloc(nodes);
    if (e == hipSuccess) e = st1.alloc(nodes);
    if (e != hipSuccess) return hip_fail(e, "lbvh_build (scratch)");
    std::vector<uint32_t> offs(leaf_off);
    offs.insert(offs.end(), inner_off.begin(), inner_off.end());
    if (e == hipSuccess)
        e = hipMemcpyAsync(d_off.p, offs.data(), offs.size() * sizeof(uint32_t), hipMemcpyHostToDevice, stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(d_outs.p, outs.data(), k * sizeof(LbvhMeshOut), hipMemcpyHostToDevice, stream);
    if (e != hipSuccess) return hip_fail(e, "lbvh_build (upload)");
    // the radix sort's histograms (256 digits x tiles)
    const uint32_t tiles = uint32_t((size_t(total) + kSortTile - 1) / kSortTile);
    const uint32_t hist_n = 256 * tiles, scan_blocks = uint32_t((size_t(hist_n) + kScanChunk - 1) / kScanChunk);
    if (scan_blocks > 1024 * 16) return fail(ATLAS_RT_E_INVALID_ARG, "lbvh_build: too many triangles in one batch");
    DevBuf<uint32_t> hist;
    e = hist.alloc(hist_n);
    if (e == hipSuccess) e = parts.alloc(scan_blocks);
    if (e != hipSuccess) return hip_fail(e, "lbvh_build (sort storage)");

    EventPair ev;
    if (device_ms) {
        ATLAS_HIP_TRY(hipEventCreate(&ev.t0));
        ATLAS_HIP_TRY(hipEventCreate(&ev.t1));
        ATLAS_HIP_TRY(hipEventRecord(ev.t0, stream));
    }
    const Batch Bt{verts, d_off.p, d_off.p + (k + 1), k, total, inner};
    // running minimum starts at +FLT_MAX, running maximum at -FLT_MAX (ordered-int form)
    auto ordered = [](float f) {
        int i;
        std::memcpy(&i, &f, 4);
        return i >= 0 ? i : i ^ 0x7FFFFFFF;
    };
    std::vector<int> binit(6 * size_t(k));
    for (uint32_t m = 0; m < k; m++)
        for (int a = 0; a < 3; a++) {
            binit[6 * m + a] = ordered(3.4028234663852886e+38f);
            binit[6 * m + 3 + a] = ordered(-3.4028234663852886e+38f);
        }
    ATLAS_HIP_TRY(hipMemcpyAsync(bounds.p, binit.data(), binit.size() * sizeof(int), hipMemcpyHostToDevice, stream));
    if (inner) {
        hipLaunchKernelGGL(k_fill_i32, dim3(blocks(inner)), dim3(kB), 0, stream, parent_inner.p, inner, -1);
    }
    hipLaunchKernelGGL(k_fill_i32, dim3(blocks(total)), dim3(kB), 0, stream, parent_leaf.p, total, -1);
    hipLaunchKernelGGL(k_fill_i32, dim3(blocks(k)), dim3(kB), 0, stream, reinterpret_cast<int*>(depth.p), k, 0);
    hipLaunchKernelGGL(k_centroids, dim3(blocks(total)), dim3(kB), 0, stream, Bt, cen.p, bounds.p);
    hipLaunchKernelGGL(k_keys, dim3(blocks(total)), dim3(kB), 0, stream, Bt, cen.p, bounds.p, keys.p);
    {
        uint64_t* src = keys.p;
        uint64_t* dst = sorted.p;
        for (uint32_t shift = 0; shift < 64; shift += 8) {  // 8 passes: the result ends in `sorted`
            hipLaunchKernelGGL(k_sort_hist, dim3(tiles), dim3(kB), 0, stream, src, total, shift, tiles, hist.p);
            hipLaunchKernelGGL(k_scan_reduce, dim3(scan_blocks), dim3(kB), 0, stream, hist.p, hist_n, parts.p);
            hipLaunchKernelGGL(k_scan_parts, dim3(1), dim3(1024), 0, stream, parts.p, scan_blocks);
            hipLaunchKernelGGL(k_scan_down, dim3(scan_blocks), dim3(kB), 0, stream, hist.p, hist_n, parts.p);
            hipLaunchKernelGGL(k_sort_scatter, dim3(tiles), dim3(kB), 0, stream, src, dst, total, shift, tiles,
                               hist.p);
            std::swap(src, dst);
        }
        // an even number of passes leaves the keys in `keys`
        std::swap(keys.p, sorted.p);
    }
    hipLaunchKernelGGL(k_leaf_boxes, dim3(blocks(total)), dim3(kB), 0, stream, Bt, sorted.p, leaf_lo.p, leaf_hi.p);
    if (inner) {
        hipLaunchKernelGGL(k_karras, dim3(blocks(inner)), dim3(kB), 0, stream, Bt, sorted.p, child.p, parent_inner.p,
                           parent_leaf.p, cnt.p);
        // inner-node depths, then boxes level by level from the deepest
        hipLaunchKernelGGL(k_depth_init, dim3(blocks(inner)), dim3(kB), 0, stream, inner, parent_inner.p, dep0.p);
        hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(kB), 0, stream, reinterpret_cast<int*>(maxd.p), 1u, 0);
        int2* da = dep0.p;
        int2* db = dep1.p;
        for (int r = 0; r < kJumpRounds; r++) {
            hipLaunchKernelGGL(k_depth_jump, dim3(blocks(inner)), dim3(kB), 0, stream, inner, da, db,
                               r + 1 == kJumpRounds ? maxd.p : nullptr);
            std::swap(da, db);
        }
        uint32_t max_level = 0;
        ATLAS_HIP_TRY(hipMemcpyAsync(&max_level, maxd.p, sizeof(uint32_t), hipMemcpyDeviceToHost, stream));
        ATLAS_HIP_TRY(hipStreamSynchronize(stream));
        // 7 rounds resolve depths below 127; a radix tree over one mesh's
        // 54 key bits is at most 54 deep
        if (max_level >= (1u << kJumpRounds) - 1)
            return fail(ATLAS_RT_E_STACK_OVERFLOW, "lbvh_build: radix tree deeper than 126 levels");
        for (int level = int(max_level); level >= 0; level--)
            hipLaunchKernelGGL(k_level_boxes, dim3(blocks(inner)), dim3(kB), 0, stream, inner, level, da, child.p,
                               leaf_lo.p, leaf_hi.p, inner_lo.p, inner_hi.p);
    }
    hipLaunchKernelGGL(k_jump_init, dim3(blocks(nodes)), dim3(kB), 0, stream, Bt, child.p, parent_inner.p,
                       parent_leaf.p, leaf_lo.p, leaf_hi.p, inner_lo.p, inner_hi.p, cnt.p, st0.p);
    int4* a = st0.p;
    int4* b = st1.p;
    for (int r = 0; r < kJumpRounds; r++) {
        hipLaunchKernelGGL(k_jump, dim3(blocks(nodes)), dim3(kB), 0, stream, nodes, a, b);
        std::swap(a, b);
    }
    if (inner)
        hipLaunchKernelGGL(k_write_nodes, dim3(blocks(inner)), dim3(kB), 0, stream, Bt, d_outs.p, child.p, leaf_lo.p,
                           leaf_hi.p, inner_lo.p, inner_hi.p, cnt.p, a);
    hipLaunchKernelGGL(k_write_tris, dim3(blocks(total)), dim3(kB), 0, stream, Bt, d_outs.p, sorted.p, a, depth.p);
    ATLAS_HIP_TRY(hipGetLastError());
    if (device_ms) ATLAS_HIP_TRY(hipEventRecord(ev.t1, stream));
    depth_out.assign(k, 0);
    ATLAS_HIP_TRY(hipMemcpyAsync(depth_out.data(), depth.p, k * sizeof(uint32_t), hipMemcpyDeviceToHost, stream));
    ATLAS_HIP_TRY(hipStreamSynchronize(stream));
    if (device_ms) ATLAS_HIP_TRY(hipEventElapsedTime(device_ms, ev.t0, ev.t1));
    return ATLAS_RT_OK;
}

}  // namespace atlas_rt
