// Collapse of the engine's binary BVH into 4- or 8-wide nodes.
#include "wide.h"

#include <algorithm>
#include <cstring>

namespace atlas_rt {

namespace {

struct Child {
    float bmin[3], bmax[3];
    int32_t ptr;
    bool leaf = false;
};

float area(const Child& c) {
    const float dx = c.bmax[0] - c.bmin[0], dy = c.bmax[1] - c.bmin[1], dz = c.bmax[2] - c.bmin[2];
    return 2.0f * (dx * dy + dy * dz + dz * dx);
}

void children_of(const atlas_rt_gpu_bvh_node& n, Child& l, Child& r) {
    std::memcpy(l.bmin, n.left_min, 12);
    std::memcpy(l.bmax, n.left_max, 12);
    std::memcpy(r.bmin, n.right_min, 12);
    std::memcpy(r.bmax, n.right_max, 12);
    l.ptr = n.left_ptr;
    r.ptr = n.right_ptr;
    l.leaf = r.leaf = false;
}

}  // namespace

WideTree collapse_wide(const std::vector<atlas_rt_gpu_bvh_node>& nodes, const std::vector<uint8_t>& end_of_leaf,
                       int width, int max_leaf) {
    WideTree out;
    if (nodes.empty()) return out;
    const int G = width / 4;
    const int F = wide_node_float4s(width);
    max_leaf = std::max(1, std::min(max_leaf, 7));
    // slot range (first, count) under every binary child pointer; children
    // come after their parent in pre-order, so one reverse sweep suffices
    auto leaf_range = [&](int32_t ptr, uint32_t& first, uint32_t& count) {
        first = uint32_t(~ptr);
        uint32_t e = first;
        while (e < end_of_leaf.size() && !end_of_leaf[e]) e++;
        count = e - first + 1;
    };
    std::vector<uint32_t> sub_first(nodes.size()), sub_count(nodes.size());
    std::vector<uint8_t> contiguous(nodes.size(), 1);
    for (size_t n = nodes.size(); n-- > 0;) {
        uint32_t f[2], c[2];
        uint8_t ok = 1;
        const int32_t ptr[2] = {nodes[n].left_ptr, nodes[n].right_ptr};
        for (int k = 0; k < 2; k++) {
            if (ptr[k] < 0) {
                leaf_range(ptr[k], f[k], c[k]);
            } else {
                f[k] = sub_first[size_t(ptr[k])];
                c[k] = sub_count[size_t(ptr[k])];
                ok &= contiguous[size_t(ptr[k])];
            }
        }
        sub_first[n] = f[0];
        sub_count[n] = c[0] + c[1];
        contiguous[n] = ok && f[1] == f[0] + c[0];
    }
    // a child becomes a wide leaf when its whole slot range fits
    auto as_leaf = [&](Child& ch) {
        uint32_t first, count;
        if (ch.ptr < 0) {
            leaf_range(ch.ptr, first, count);
            if (count > uint32_t(max_leaf)) {
                // a binary leaf larger than a wide leaf: runs to its end flag
                ch.ptr = ~int32_t((first << 3) | 7u);
                ch.leaf = true;
                return true;
            }
        } else {
            if (!contiguous[size_t(ch.ptr)]) return false;
            first = sub_first[size_t(ch.ptr)];
            count = sub_count[size_t(ch.ptr)];
        }
        if (count > uint32_t(max_leaf)) return false;
        ch.ptr = ~int32_t((first << 3) | (count - 1));
        ch.leaf = true;
        return true;
    };
    // wide nodes in breadth-first order: node w is written when it is taken
    // from the queue, and its inner children get the next free numbers
    const size_t per = size_t(F) * 4;
    std::vector<int32_t> queue{0};  // binary node of every wide node
    queue.reserve(nodes.size() / size_t(width - 1) + 1);
    out.data.reserve(queue.capacity() * per);
    Child c[8];
    for (size_t w = 0; w < queue.size(); w++) {
        int n = 2;
        children_of(nodes[size_t(queue[w])], c[0], c[1]);
        as_leaf(c[0]);
        as_leaf(c[1]);
        while (n < width) {
            int best = -1;
            float best_sa = -1.0f;
            for (int i = 0; i < n; i++)
                if (!c[i].leaf && c[i].ptr >= 0 && area(c[i]) > best_sa) {
                    best_sa = area(c[i]);
                    best = i;
                }
            if (best < 0) break;
            Child l, r;
            children_of(nodes[size_t(c[best].ptr)], l, r);
            as_leaf(l);
            as_leaf(r);
            c[best] = l;
            c[n++] = r;
        }
        out.data.resize(out.data.size() + per, 0.0f);
        float* nd = out.data.data() + w * per;
        for (int i = 0; i < n; i++) {
            if (!c[i].leaf && c[i].ptr >= 0) {
                const int32_t id = int32_t(queue.size());
                queue.push_back(c[i].ptr);
                c[i].ptr = id;
            }
            const size_t q = size_t(i) / 4, lane = size_t(i) % 4;
            for (int ax = 0; ax < 3; ax++) {
                nd[(size_t(ax) * G + q) * 4 + lane] = c[i].bmin[ax];
                nd[(size_t(3 + ax) * G + q) * 4 + lane] = c[i].bmax[ax];
            }
            std::memcpy(&nd[(6 * size_t(G) + q) * 4 + lane], &c[i].ptr, 4);
        }
        std::memcpy(&nd[7 * size_t(G) * 4], &n, 4);
    }
    out.node_count = uint32_t(queue.size());
    // stack bound: along any root-to-leaf path, a node pushes at most
    // (children - 1) entries; children are numbered after their parents
    std::vector<uint32_t> depth(queue.size(), 0);
    for (size_t w = queue.size(); w-- > 0;) {
        const float* nd = out.data.data() + w * per;
        int32_t n;
        std::memcpy(&n, &nd[7 * size_t(G) * 4], 4);
        uint32_t below = 0;
        for (int i = 0; i < n; i++) {
            int32_t ptr;
            std::memcpy(&ptr, &nd[(6 * size_t(G) + size_t(i) / 4) * 4 + size_t(i) % 4], 4);
            if (ptr >= 0) below = std::max(below, depth[size_t(ptr)]);
        }
        depth[w] = uint32_t(n) - 1 + below;
    }
    out.max_stack = depth[0];
    return out;
}

}  // namespace atlas_rt
