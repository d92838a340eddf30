// TEST INFRASTRUCTURE - the oracle.  Only tests/, __graft_entry__.smoke() and
// bench.py's cpu_baseline leg may load it, and only as the checker.
//
// Serial CPU restatement of the GPU linear-BVH builder (the product's
// atlas_rt_blas_build_device): centroids ((v0 + v1) + v2) / 3, 30-bit Morton
// codes over the centroid bounds (1024 cells per axis), keys
// (morton << 24) | triangle index sorted ascending, the binary radix tree of
// the sorted keys (Karras, HPG 2012: each inner node's range and split from
// common-prefix lengths), boxes bottom-up, and the engine's flattening
// (src/engine/volume/BVH.cpp:408-441): larger surface area left (ties keep
// the radix-tree order), pre-order inner nodes, one triangle per leaf,
// leaf pointer ~slot.  No FMA contraction, so every float rounds as written.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <functional>
#include <vector>

#include "../include/atlas_rt.h"

namespace {

int ordered(float f) {
    int i;
    std::memcpy(&i, &f, 4);
    return i >= 0 ? i : i ^ 0x7FFFFFFF;
}
float from_ordered(int i) {
    const int b = i >= 0 ? i : i ^ 0x7FFFFFFF;
    float f;
    std::memcpy(&f, &b, 4);
    return f;
}

uint32_t expand_bits(uint32_t x) {
    x = (x * 0x00010001u) & 0xFF0000FFu;
    x = (x * 0x00000101u) & 0x0F00F00Fu;
    x = (x * 0x00000011u) & 0xC30C30C3u;
    x = (x * 0x00000005u) & 0x49249249u;
    return x;
}

uint32_t quantize(float c, float lo, float hi) {
    const float e = hi - lo;
    const float s = e > 0.0f ? 1024.0f / e : 0.0f;
    const float q = (c - lo) * s;
    const uint32_t u = q > 0.0f ? uint32_t(q) : 0u;
    return u < 1023u ? u : 1023u;
}

struct Box {
    float lo[3], hi[3];
};

float area(const Box& b) {
    const float dx = b.hi[0] - b.lo[0], dy = b.hi[1] - b.lo[1], dz = b.hi[2] - b.lo[2];
    return 2.0f * ((dx * dy + dy * dz) + dz * dx);
}

int delta(const std::vector<uint64_t>& K, int a, int b) {
    if (b < 0 || b >= int(K.size())) return -1;
    return __builtin_clzll(K[size_t(a)] ^ K[size_t(b)]);
}

// Children of inner node i: >= 0 inner node, < 0 ~leaf (sorted position).
void karras(const std::vector<uint64_t>& K, int i, int& c0, int& c1) {
    const int d = delta(K, i, i + 1) - delta(K, i, i - 1) >= 0 ? 1 : -1;
    const int dmin = delta(K, i, i - d);
    int lmax = 2;
    while (delta(K, i, i + lmax * d) > dmin) lmax <<= 1;
    int l = 0;
    for (int t = lmax >> 1; t >= 1; t >>= 1)
        if (delta(K, i, i + (l + t) * d) > dmin) l += t;
    const int j = i + l * d;
    const int dnode = delta(K, i, j);
    int s = 0, t = l;
    do {
        t = (t + 1) >> 1;
        if (delta(K, i, i + (s + t) * d) > dnode) s += t;
    } while (t > 1);
    const int gamma = i + s * d + (d < 0 ? -1 : 0);
    const int lo = std::min(i, j), hi = std::max(i, j);
    c0 = lo == gamma ? ~gamma : gamma;
    c1 = hi == gamma + 1 ? ~(gamma + 1) : gamma + 1;
}

}  // namespace

extern "C" {

// The radix tree of n sorted, distinct 64-bit keys: children[2 i], [2 i + 1]
// of inner node i (>= 0 inner, < 0 ~leaf).
void oracle_radix_tree(const uint64_t* keys, uint32_t n, int32_t* children) {
    std::vector<uint64_t> K(keys, keys + n);
    for (uint32_t i = 0; i + 1 < n; i++) karras(K, int(i), children[2 * i], children[2 * i + 1]);
}

// verts: 9 floats per triangle.  Outputs: nodes[n - 1], tris[n] (engine
// GPUBVHTriangle: end flag 1, material 0, opacity 1), orig[n] (slot ->
// triangle index), *max_depth (inner-node ancestors of the deepest leaf, at
// least 1).
void oracle_lbvh_build(const float* verts, uint32_t n, atlas_rt_gpu_bvh_node* nodes, atlas_rt_gpu_bvh_triangle* tris,
                       uint32_t* orig, uint32_t* max_depth) {
    std::vector<float> cen(3 * size_t(n));
    int lo[3], hi[3];
    for (int a = 0; a < 3; a++) {
        lo[a] = ordered(3.4028234663852886e+38f);
        hi[a] = ordered(-3.4028234663852886e+38f);
    }
    for (uint32_t g = 0; g < n; g++) {
        const float* v = verts + 9 * size_t(g);
        for (int a = 0; a < 3; a++) {
            const float c = ((v[a] + v[3 + a]) + v[6 + a]) / 3.0f;
            cen[3 * size_t(g) + a] = c;
            lo[a] = std::min(lo[a], ordered(c));
            hi[a] = std::max(hi[a], ordered(c));
        }
    }
    std::vector<uint64_t> K(n);
    for (uint32_t g = 0; g < n; g++) {
        uint32_t q[3];
        for (int a = 0; a < 3; a++) q[a] = quantize(cen[3 * size_t(g) + a], from_ordered(lo[a]), from_ordered(hi[a]));
        const uint32_t morton = (expand_bits(q[0]) << 2) | (expand_bits(q[1]) << 1) | expand_bits(q[2]);
        K[g] = (uint64_t(morton) << 24) | g;
    }
    std::sort(K.begin(), K.end());
    const int inner = int(n) - 1;
    std::vector<int> c0(std::max(inner, 0)), c1(std::max(inner, 0));
    for (int i = 0; i < inner; i++) karras(K, i, c0[size_t(i)], c1[size_t(i)]);
    auto tri_of = [&](int leaf) { return uint32_t(K[size_t(leaf)] & ((1ull << 24) - 1)); };
    std::vector<Box> leaf_box(n), inner_box(std::max(inner, 0));
    for (uint32_t x = 0; x < n; x++) {
        const float* v = verts + 9 * size_t(tri_of(int(x)));
        for (int a = 0; a < 3; a++) {
            leaf_box[x].lo[a] = std::fmin(std::fmin(v[a], v[3 + a]), v[6 + a]);
            leaf_box[x].hi[a] = std::fmax(std::fmax(v[a], v[3 + a]), v[6 + a]);
        }
    }
    std::vector<uint8_t> done(std::max(inner, 0), 0);
    std::function<const Box&(int)> box_of = [&](int c) -> const Box& {
        if (c < 0) return leaf_box[size_t(~c)];
        Box& b = inner_box[size_t(c)];
        if (done[size_t(c)]) return b;
        done[size_t(c)] = 1;
        const Box& l = box_of(c0[size_t(c)]);
        const Box& r = box_of(c1[size_t(c)]);
        for (int a = 0; a < 3; a++) {
            b.lo[a] = std::fmin(l.lo[a], r.lo[a]);
            b.hi[a] = std::fmax(l.hi[a], r.hi[a]);
        }
        return b;
    };
    uint32_t depth_max = 1;
    uint32_t next_node = 0, next_slot = 0;
    auto write_leaf = [&](int leaf, uint32_t depth) -> int32_t {
        const uint32_t slot = next_slot++;
        const uint32_t t = tri_of(leaf);
        const float* v = verts + 9 * size_t(t);
        atlas_rt_gpu_bvh_triangle& o = tris[slot];
        std::memcpy(o.v0, v, 12);
        std::memcpy(o.v1, v + 3, 12);
        std::memcpy(o.v2, v + 6, 12);
        o.v0[3] = 1.0f;
        o.v1[3] = 0.0f;
        o.v2[3] = 1.0f;
        orig[slot] = t;
        depth_max = std::max(depth_max, depth);
        return ~int32_t(slot);
    };
    if (inner == 0) {
        write_leaf(0, 0);
    } else {
        box_of(0);
        std::function<void(int, uint32_t)> flatten = [&](int node, uint32_t depth) {
            const uint32_t me = next_node++;
            int first = c0[size_t(node)], second = c1[size_t(node)];
            if (area(box_of(first)) < area(box_of(second))) std::swap(first, second);
            atlas_rt_gpu_bvh_node& o = nodes[me];
            std::memset(&o, 0, sizeof(o));
            const Box& bl = box_of(first);
            const Box& br = box_of(second);
            std::memcpy(o.left_min, bl.lo, 12);
            std::memcpy(o.left_max, bl.hi, 12);
            std::memcpy(o.right_min, br.lo, 12);
            std::memcpy(o.right_max, br.hi, 12);
            if (first < 0) {
                o.left_ptr = write_leaf(~first, depth + 1);
            } else {
                o.left_ptr = int32_t(next_node);
                flatten(first, depth + 1);
            }
            atlas_rt_gpu_bvh_node& o2 = nodes[me];
            if (second < 0) {
                o2.right_ptr = write_leaf(~second, depth + 1);
            } else {
                o2.right_ptr = int32_t(next_node);
                flatten(second, depth + 1);
            }
        };
        flatten(0, 0);
    }
    *max_depth = depth_max;
}

}  // extern "C"
