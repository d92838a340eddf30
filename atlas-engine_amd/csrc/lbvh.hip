// GPU linear-BVH builder, gfx950.
//
// Per mesh (all meshes of a batch in the same launches):
//   1. centroids ((v0 + v1) + v2) / 3 and their bounds (ordered-integer atomics);
//   2. 30-bit Morton codes on a 1024^3 grid over the centroid bounds, keys
//      (mesh << 54) | (morton << 24) | triangle index, so every key is unique
//      and the sort is deterministic;
//   3. rocPRIM radix sort of the 64-bit keys;
//   4. the binary radix tree over each mesh's sorted keys (Karras, "Maximizing
//      parallelism in the construction of BVHs, octrees, and k-d trees", HPG
//      2012): one thread per inner node finds its key range and split;
//   5. boxes and leaf counts bottom-up (the second child to arrive at a node
//      merges both, acquire/release atomics at device scope);
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

#include <rocprim/device/device_radix_sort.hpp>

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
    const uint32_t g = blockIdx.x * kB + threadIdx.x;
    if (g >= B.total) return;
    const uint32_t m = find_mesh(B.leaf_off, B.k, g);
    const float* v = B.verts + 9 * size_t(g);
    const float cx = ((v[0] + v[3]) + v[6]) / 3.0f;
    const float cy = ((v[1] + v[4]) + v[7]) / 3.0f;
    const float cz = ((v[2] + v[5]) + v[8]) / 3.0f;
    cen[g] = make_float4(cx, cy, cz, 0.0f);
    int* b = bounds + 6 * m;
    atomicMin(&b[0], f2o(cx));
    atomicMin(&b[1], f2o(cy));
    atomicMin(&b[2], f2o(cz));
    atomicMax(&b[3], f2o(cx));
    atomicMax(&b[4], f2o(cy));
    atomicMax(&b[5], f2o(cz));
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
                                               int* __restrict__ parent_inner, int* __restrict__ parent_leaf) {
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
    if (c0 >= 0)
        parent_inner[c0] = int(g);
    else
        parent_leaf[~c0] = int(g);
    if (c1 >= 0)
        parent_inner[c1] = int(g);
    else
        parent_leaf[~c1] = int(g);
}

// (Plain pointers: during the bottom-up pass other threads write these.)
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

// Leaf boxes, then boxes and leaf counts up to each mesh's root.
__global__ __launch_bounds__(kB) void k_bottom_up(Batch B, const uint64_t* __restrict__ sorted,
                                                  const int2* __restrict__ child, const int* __restrict__ parent_inner,
                                                  const int* __restrict__ parent_leaf, float4* leaf_lo,
                                                  float4* leaf_hi, float4* inner_lo, float4* inner_hi,
                                                  uint32_t* cnt, uint32_t* __restrict__ arrivals) {
    const uint32_t x = blockIdx.x * kB + threadIdx.x;
    if (x >= B.total) return;
    const uint32_t m = find_mesh(B.leaf_off, B.k, x);
    const uint32_t tri = B.leaf_off[m] + uint32_t(sorted[x] & ((1ull << kIndexBits) - 1));
    const float* v = B.verts + 9 * size_t(tri);
    leaf_lo[x] = make_float4(fminf(fminf(v[0], v[3]), v[6]), fminf(fminf(v[1], v[4]), v[7]),
                             fminf(fminf(v[2], v[5]), v[8]), 0.0f);
    leaf_hi[x] = make_float4(fmaxf(fmaxf(v[0], v[3]), v[6]), fmaxf(fmaxf(v[1], v[4]), v[7]),
                             fmaxf(fmaxf(v[2], v[5]), v[8]), 0.0f);
    int p = parent_leaf[x];
    while (p >= 0) {
        // the first child to arrive leaves; the second sees both boxes
        const uint32_t seen = __hip_atomic_fetch_add(&arrivals[p], 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
        if (seen == 0) return;
        const int2 c = child[p];
        float4 lo0, hi0, lo1, hi1;
        uint32_t n0, n1;
        child_box(c.x, leaf_lo, leaf_hi, inner_lo, inner_hi, cnt, lo0, hi0, n0);
        child_box(c.y, leaf_lo, leaf_hi, inner_lo, inner_hi, cnt, lo1, hi1, n1);
        inner_lo[p] = make_float4(fminf(lo0.x, lo1.x), fminf(lo0.y, lo1.y), fminf(lo0.z, lo1.z), 0.0f);
        inner_hi[p] = make_float4(fmaxf(hi0.x, hi1.x), fmaxf(hi0.y, hi1.y), fmaxf(hi0.z, hi1.z), 0.0f);
        cnt[p] = n0 + n1;
        p = parent_inner[p];
    }
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
    const uint32_t x = blockIdx.x * kB + threadIdx.x;
    if (x >= B.total) return;
    const uint32_t m = find_mesh(B.leaf_off, B.k, x);
    const uint32_t local = uint32_t(sorted[x] & ((1ull << kIndexBits) - 1));
    const float* v = B.verts + 9 * (size_t(B.leaf_off[m]) + local);
    const int4 s = st[B.inner + x];
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
    atomicMax(&depth[m], uint32_t(s.z > 1 ? s.z : 1));
}

template <class T>
struct DevBuf {
    T* p = nullptr;
    ~DevBuf() {
        if (p) (void)hipFree(p);
    }
    hipError_t alloc(size_t n) { return hipMalloc(reinterpret_cast<void**>(&p), std::max<size_t>(n, 1) * sizeof(T)); }
};

}  // namespace

atlas_rt_status lbvh_build(const std::vector<const float*>& verts, const std::vector<uint32_t>& counts,
                           const std::vector<LbvhMeshOut>& outs, std::vector<uint32_t>& depth_out,
                           hipStream_t stream, float* device_ms) {
    const uint32_t k = uint32_t(counts.size());
    if (k == 0 || k > kMaxMeshes || verts.size() != k || outs.size() != k)
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
    DevBuf<float> d_verts;
    DevBuf<uint32_t> d_off;
    DevBuf<LbvhMeshOut> d_outs;
    DevBuf<float4> cen, leaf_lo, leaf_hi, inner_lo, inner_hi;
    DevBuf<int> bounds, parent_inner, parent_leaf;
    DevBuf<uint64_t> keys, sorted;
    DevBuf<int2> child;
    DevBuf<uint32_t> cnt, arrivals, depth;
    DevBuf<int4> st0, st1;
    hipError_t e = d_verts.alloc(9 * size_t(total));
    if (e == hipSuccess) e = d_off.alloc(2 * (k + 1));
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
    if (e == hipSuccess) e = arrivals.alloc(inner);
    if (e == hipSuccess) e = depth.alloc(k);
    if (e == hipSuccess) e = st0.alloc(nodes);
    if (e == hipSuccess) e = st1.alloc(nodes);
    if (e != hipSuccess) return hip_fail(e, "lbvh_build (scratch)");
    for (uint32_t m = 0; m < k && e == hipSuccess; m++)
        e = hipMemcpyAsync(d_verts.p + 9 * size_t(leaf_off[m]), verts[m], 9 * size_t(counts[m]) * sizeof(float),
                           hipMemcpyHostToDevice, stream);
    std::vector<uint32_t> offs(leaf_off);
    offs.insert(offs.end(), inner_off.begin(), inner_off.end());
    if (e == hipSuccess)
        e = hipMemcpyAsync(d_off.p, offs.data(), offs.size() * sizeof(uint32_t), hipMemcpyHostToDevice, stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(d_outs.p, outs.data(), k * sizeof(LbvhMeshOut), hipMemcpyHostToDevice, stream);
    if (e != hipSuccess) return hip_fail(e, "lbvh_build (upload)");
    // the radix sort's temporary storage
    size_t temp_bytes = 0;
    e = rocprim::radix_sort_keys(nullptr, temp_bytes, keys.p, sorted.p, total, 0, 64, stream);
    DevBuf<char> temp;
    if (e == hipSuccess) e = temp.alloc(temp_bytes);
    if (e != hipSuccess) return hip_fail(e, "lbvh_build (sort storage)");

    hipEvent_t t0 = nullptr, t1 = nullptr;
    if (device_ms) {
        ATLAS_HIP_TRY(hipEventCreate(&t0));
        ATLAS_HIP_TRY(hipEventCreate(&t1));
        ATLAS_HIP_TRY(hipEventRecord(t0, stream));
    }
    const Batch Bt{d_verts.p, d_off.p, d_off.p + (k + 1), k, total, inner};
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
        hipLaunchKernelGGL(k_fill_i32, dim3(blocks(inner)), dim3(kB), 0, stream, reinterpret_cast<int*>(arrivals.p),
                           inner, 0);
    }
    hipLaunchKernelGGL(k_fill_i32, dim3(blocks(total)), dim3(kB), 0, stream, parent_leaf.p, total, -1);
    hipLaunchKernelGGL(k_fill_i32, dim3(blocks(k)), dim3(kB), 0, stream, reinterpret_cast<int*>(depth.p), k, 0);
    hipLaunchKernelGGL(k_centroids, dim3(blocks(total)), dim3(kB), 0, stream, Bt, cen.p, bounds.p);
    hipLaunchKernelGGL(k_keys, dim3(blocks(total)), dim3(kB), 0, stream, Bt, cen.p, bounds.p, keys.p);
    e = rocprim::radix_sort_keys(temp.p, temp_bytes, keys.p, sorted.p, total, 0, 64, stream);
    if (e != hipSuccess) return hip_fail(e, "lbvh_build (sort)");
    if (inner)
        hipLaunchKernelGGL(k_karras, dim3(blocks(inner)), dim3(kB), 0, stream, Bt, sorted.p, child.p, parent_inner.p,
                           parent_leaf.p);
    hipLaunchKernelGGL(k_bottom_up, dim3(blocks(total)), dim3(kB), 0, stream, Bt, sorted.p, child.p, parent_inner.p,
                       parent_leaf.p, leaf_lo.p, leaf_hi.p, inner_lo.p, inner_hi.p, cnt.p, arrivals.p);
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
    if (device_ms) ATLAS_HIP_TRY(hipEventRecord(t1, stream));
    depth_out.assign(k, 0);
    ATLAS_HIP_TRY(hipMemcpyAsync(depth_out.data(), depth.p, k * sizeof(uint32_t), hipMemcpyDeviceToHost, stream));
    ATLAS_HIP_TRY(hipStreamSynchronize(stream));
    if (device_ms) {
        ATLAS_HIP_TRY(hipEventElapsedTime(device_ms, t0, t1));
        (void)hipEventDestroy(t0);
        (void)hipEventDestroy(t1);
    }
    return ATLAS_RT_OK;
}

}  // namespace atlas_rt
