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
    struct Item {
        int32_t bin;     // binary node
        uint32_t wide;   // wide node it becomes
    };
    std::vector<Item> work{{0, 0}};
    std::vector<std::vector<Child>> kids(1);
    std::vector<std::vector<uint32_t>> inner_children(1);
    for (size_t w = 0; w < work.size(); w++) {
        std::vector<Child> c(2);
        children_of(nodes[size_t(work[w].bin)], c[0], c[1]);
        for (Child& ch : c) as_leaf(ch);
        while (int(c.size()) < width) {
            int best = -1;
            float best_sa = -1.0f;
            for (int i = 0; i < int(c.size()); i++)
                if (!c[size_t(i)].leaf && c[size_t(i)].ptr >= 0 && area(c[size_t(i)]) > best_sa) {
                    best_sa = area(c[size_t(i)]);
                    best = i;
                }
            if (best < 0) break;
            Child l, r;
            children_of(nodes[size_t(c[size_t(best)].ptr)], l, r);
            as_leaf(l);
            as_leaf(r);
            c[size_t(best)] = l;
            c.push_back(r);
        }
        for (Child& ch : c)
            if (!ch.leaf && ch.ptr >= 0) {
                const uint32_t id = uint32_t(work.size());
                work.push_back({ch.ptr, id});
                kids.emplace_back();
                inner_children.emplace_back();
                inner_children[w].push_back(id);
                ch.ptr = int32_t(id);
            }
        kids[w] = std::move(c);
    }
    out.node_count = uint32_t(work.size());
    out.data.assign(size_t(out.node_count) * F * 4, 0.0f);
    for (size_t w = 0; w < work.size(); w++) {
        float* nd = out.data.data() + w * size_t(F) * 4;
        const auto& c = kids[w];
        for (size_t i = 0; i < c.size(); i++) {
            const size_t q = i / 4, lane = i % 4;
            for (int a = 0; a < 3; a++) {
                nd[(size_t(a) * G + q) * 4 + lane] = c[i].bmin[a];
                nd[(size_t(3 + a) * G + q) * 4 + lane] = c[i].bmax[a];
            }
            std::memcpy(&nd[(6 * size_t(G) + q) * 4 + lane], &c[i].ptr, 4);
        }
        const int32_t n = int32_t(c.size());
        std::memcpy(&nd[7 * size_t(G) * 4], &n, 4);
    }
    // stack bound: along any root-to-leaf path, a node pushes at most
    // (children - 1) entries; children are numbered after their parents
    std::vector<uint32_t> depth(work.size(), 0);
    for (size_t w = work.size(); w-- > 0;) {
        uint32_t below = 0;
        for (uint32_t ch : inner_children[w]) below = std::max(below, depth[ch]);
        depth[w] = uint32_t(kids[w].size()) - 1 + below;
    }
    out.max_stack = depth[0];
    return out;
}

}  // namespace atlas_rt
