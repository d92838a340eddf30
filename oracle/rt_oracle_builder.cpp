// TEST INFRASTRUCTURE - the oracle.  Only tests/, __graft_entry__.smoke() and
// bench.py's cpu_baseline leg may load it, and only as the checker.
//
// Serial restatement of the engine's BVH builder, src/engine/volume/BVH.cpp:14-869:
//   - BVH::BVH(aabbs, data) (BVH.cpp:14-56): root node runs the 4-argument Build
//     (object split + SBVH spatial split, BVH.cpp:249-341); every child runs
//     the 3-argument AABB-only Build (BVH.cpp:343-406): leaf iff one reference,
//     otherwise binned SAH object split or the median split fallback.
//   - BVH::BVH(aabbs) (BVH.cpp:58-101): the TLAS builder, 3-argument Build from
//     the root with 64 bins, plus the single-instance special node (BVH.cpp:80-88).
//   - Flatten (BVH.cpp:408-441): pre-order, larger surface area child left.
// The parallel job split at depth <= 6 (BVH.cpp:310-328) does not change the
// result, so the restatement recurses serially.
#include <algorithm>
#include <vector>
#include <memory>

#include "rt_oracle.h"

namespace oracle {

namespace {

struct Ref {
    uint32_t idx = 0;
    uint32_t nodeIdx = 0;
    bool endOfNode = false;
    Box aabb = initial_box();
};

struct Tri {
    V3 v0, v1, v2;
};

struct Split {
    float cost = FLT_MAX;
    int32_t axis = -1;
    uint32_t binIdx = 0;
    float pos = 0.0f;
    Box leftAABB = initial_box();
    Box rightAABB = initial_box();
};

struct Bin {
    Box aabb = initial_box();
    uint32_t primitiveCount = 0;
    uint32_t enter = 0;
    uint32_t exit = 0;
};

struct Node {
    uint32_t depth;
    uint32_t binCount;
    float minOverlap;
    float nodeCost;
    Box aabb;
    std::vector<Ref> refs;  // non-empty => leaf
    std::unique_ptr<Node> left, right;

    Node(Box b, uint32_t d, size_t refCount, float overlap, uint32_t bins)
        : depth(d), binCount(bins), minOverlap(overlap), aabb(b) {
        nodeCost = float(refCount) * surface_area(aabb);  // BVH.cpp:230,238
    }

    uint32_t bins_at_depth() const { return std::max(binCount / (depth + 1), 16u); }  // BVH.cpp:447

    // BVHBuilder::FindObjectSplit (BVH.cpp:444-528).
    Split find_object_split(const std::vector<Ref>& refs) const {
        Split split;
        const uint32_t B = bins_at_depth();
        std::vector<Bin> bins(B);
        std::vector<Box> rightBoxes(B);
        for (int32_t i = 0; i < 3; i++) {
            for (auto& bin : bins) {
                bin.aabb = initial_box();
                bin.primitiveCount = 0;
            }
            for (auto& b : rightBoxes) b = initial_box();
            float start = aabb.min[i];
            float stop = aabb.max[i];
            if (std::fabs(stop - start) < 1e-3f) continue;
            float binSize = (stop - start) / float(B);
            float invBinSize = 1.0f / binSize;
            for (const auto& ref : refs) {
                const float value = 0.5f * (ref.aabb.min[i] + ref.aabb.max[i]);
                uint32_t binIdx = uint32_t(gclamp((value - start) * invBinSize, 0.0f, float(B) - 1.0f));
                bins[binIdx].primitiveCount++;
                grow(bins[binIdx].aabb, ref.aabb);
            }
            Box rb = initial_box();
            for (size_t j = B - 1; j > 0; j--) {
                grow(rb, bins[j].aabb);
                rightBoxes[j - 1] = rb;
            }
            Box lb = initial_box();
            uint32_t primitivesLeft = 0;
            for (size_t j = 1; j < B; j++) {
                grow(lb, bins[j - 1].aabb);
                primitivesLeft += bins[j - 1].primitiveCount;
                const Box rbox = rightBoxes[j - 1];
                const uint32_t primitivesRight = uint32_t(refs.size()) - primitivesLeft;
                if (!primitivesLeft || !primitivesRight) continue;
                const float cost = surface_area(lb) * float(primitivesLeft) +
                                   surface_area(rbox) * float(primitivesRight);
                if (cost < split.cost) {
                    split.cost = cost;
                    split.axis = i;
                    split.binIdx = uint32_t(j);
                    split.pos = start + float(j) * binSize;
                    split.leftAABB = lb;
                    split.rightAABB = rbox;
                }
            }
        }
        return split;
    }

    // BVHBuilder::PerformObjectSplit (BVH.cpp:530-556).
    void perform_object_split(const std::vector<Ref>& refs, std::vector<Ref>& rightRefs,
                              std::vector<Ref>& leftRefs, const Split& split) const {
        const uint32_t B = bins_at_depth();
        float start = aabb.min[split.axis];
        float stop = aabb.max[split.axis];
        float binSize = (stop - start) / float(B);
        float invBinSize = 1.0f / binSize;
        for (const auto& ref : refs) {
            const float value = 0.5f * (ref.aabb.min[split.axis] + ref.aabb.max[split.axis]);
            uint32_t binIdx = uint32_t(gclamp((value - start) * invBinSize, 0.0f, float(B) - 1.0f));
            if (binIdx < split.binIdx)
                leftRefs.push_back(ref);
            else
                rightRefs.push_back(ref);
        }
    }

    // BVHBuilder::SplitReference (BVH.cpp:760-798).
    static void split_reference(const Tri& tri, const Ref& current, Ref& leftRef, Ref& rightRef,
                                float planePos, int32_t axis) {
        for (int i = 0; i < 3; i++) {
            V3 a, b;
            if (i == 0) { a = tri.v0; b = tri.v1; }
            else if (i == 1) { a = tri.v1; b = tri.v2; }
            else { a = tri.v2; b = tri.v0; }
            if ((a[axis] < planePos && b[axis] > planePos) || (a[axis] > planePos && b[axis] < planePos)) {
                float off = gclamp((planePos - a[axis]) / (b[axis] - a[axis]), 0.0f, 1.0f);
                V3 p = gmix(a, b, off);
                grow(leftRef.aabb, p);
                grow(rightRef.aabb, p);
            }
            if (a[axis] <= planePos) grow(leftRef.aabb, a);
            if (a[axis] >= planePos) grow(rightRef.aabb, a);
        }
        leftRef.aabb.max[axis] = planePos;
        rightRef.aabb.min[axis] = planePos;
        intersect(leftRef.aabb, current.aabb);
        intersect(rightRef.aabb, current.aabb);
    }

    // BVHBuilder::FindSpatialSplit (BVH.cpp:558-667).
    Split find_spatial_split(const std::vector<Ref>& refs, const std::vector<Tri>& data) const {
        Split split;
        const uint32_t B = bins_at_depth();
        std::vector<Bin> bins(B);
        std::vector<Box> rightBoxes(B);
        for (int32_t i = 0; i < 3; i++) {
            for (auto& bin : bins) {
                bin.aabb = initial_box();
                bin.enter = 0;
                bin.exit = 0;
            }
            for (auto& b : rightBoxes) b = initial_box();
            float start = aabb.min[i];
            float stop = aabb.max[i];
            if (std::fabs(stop - start) < 1e-3f) continue;
            float binSize = (stop - start) / float(B);
            float invBinSize = 1.0f / binSize;
            for (const auto& ref : refs) {
                const uint32_t s = uint32_t(gclamp((ref.aabb.min[i] - start) * invBinSize, 0.0f, float(B) - 1.0f));
                const uint32_t e = uint32_t(gclamp((ref.aabb.max[i] - start) * invBinSize, 0.0f, float(B) - 1.0f));
                if (s == e) {
                    bins[s].enter++;
                    bins[s].exit++;
                    grow(bins[s].aabb, ref.aabb);
                    continue;
                }
                Ref current = ref;
                for (uint32_t j = s; j < e; j++) {
                    Ref l, r;
                    const float planePos = start + float(j + 1) * binSize;
                    split_reference(data[ref.idx], current, l, r, planePos, i);
                    grow(bins[j].aabb, l.aabb);
                    current = r;
                }
                grow(bins[e].aabb, current.aabb);
                bins[s].enter++;
                bins[e].exit++;
            }
            Box rb = initial_box();
            for (size_t j = B - 1; j > 0; j--) {
                grow(rb, bins[j].aabb);
                rightBoxes[j - 1] = rb;
            }
            Box lb = initial_box();
            uint32_t primitivesRight = uint32_t(refs.size());
            uint32_t primitivesLeft = 0;
            for (size_t j = 1; j < B; j++) {
                grow(lb, bins[j - 1].aabb);
                primitivesLeft += bins[j - 1].enter;
                primitivesRight -= bins[j - 1].exit;
                const Box rbox = rightBoxes[j - 1];
                if (!primitivesLeft || !primitivesRight) continue;
                const float cost = surface_area(lb) * float(primitivesLeft) +
                                   surface_area(rbox) * float(primitivesRight);
                if (cost < split.cost) {
                    split.cost = cost;
                    split.axis = i;
                    split.binIdx = uint32_t(j);
                    split.pos = start + float(j) * binSize;
                    split.leftAABB = lb;
                    split.rightAABB = rbox;
                }
            }
        }
        return split;
    }

    // BVHBuilder::PerformSpatialSplit (BVH.cpp:669-758).  `split` is updated in place.
    void perform_spatial_split(const std::vector<Ref>& refs, const std::vector<Tri>& data,
                               std::vector<Ref>& rightRefs, std::vector<Ref>& leftRefs, Split& split) const {
        const uint32_t B = bins_at_depth();
        float start = aabb.min[split.axis];
        float stop = aabb.max[split.axis];
        float binSize = (stop - start) / float(B);
        float invBinSize = 1.0f / binSize;
        auto bin_of = [&](float x) {
            return uint32_t(gclamp((x - start) * invBinSize, 0.0f, float(B) - 1.0f));
        };
        for (const auto& ref : refs) {
            const uint32_t s = bin_of(ref.aabb.min[split.axis]);
            const uint32_t e = bin_of(ref.aabb.max[split.axis]);
            if (e < split.binIdx) {
                leftRefs.push_back(ref);
                grow(split.leftAABB, ref.aabb);
            } else if (s >= split.binIdx) {
                rightRefs.push_back(ref);
                grow(split.rightAABB, ref.aabb);
            }
        }
        for (const auto& ref : refs) {
            const uint32_t s = bin_of(ref.aabb.min[split.axis]);
            const uint32_t e = bin_of(ref.aabb.max[split.axis]);
            if (e >= split.binIdx && s < split.binIdx) {
                Ref l, r;
                split_reference(data[ref.idx], ref, l, r, split.pos, split.axis);
                Box unsplitLeft = split.leftAABB, unsplitRight = split.rightAABB;
                Box dupLeft = split.leftAABB, dupRight = split.rightAABB;
                grow(unsplitLeft, ref.aabb);
                grow(unsplitRight, ref.aabb);
                grow(dupLeft, l.aabb);
                grow(dupRight, r.aabb);
                float leftRefCost = float(leftRefs.size());
                float rightRefCost = float(rightRefs.size());
                float leftModRefCost = float(leftRefs.size() + 1);
                float rightModRefCost = float(rightRefs.size() + 1);
                float unsplitLeftSAH = surface_area(unsplitLeft) * leftModRefCost +
                                       surface_area(split.rightAABB) * rightRefCost;
                float unsplitRightSAH = surface_area(split.leftAABB) * leftRefCost +
                                        surface_area(unsplitRight) * rightModRefCost;
                float duplicateSAH = surface_area(dupLeft) * leftModRefCost +
                                     surface_area(dupRight) * rightModRefCost;
                float minSAH = gmin(duplicateSAH, gmin(unsplitRightSAH, unsplitLeftSAH));
                if (minSAH == unsplitLeftSAH) {
                    split.leftAABB = unsplitLeft;
                    leftRefs.push_back(ref);
                } else if (minSAH == unsplitRightSAH) {
                    split.rightAABB = unsplitRight;
                    rightRefs.push_back(ref);
                } else {
                    l.idx = ref.idx;
                    r.idx = ref.idx;
                    split.leftAABB = dupLeft;
                    split.rightAABB = dupRight;
                    leftRefs.push_back(l);
                    rightRefs.push_back(r);
                }
            }
        }
    }

    // BVHBuilder::PerformMedianSplit (BVH.cpp:800-855).
    Split perform_median_split(std::vector<Ref>& refs, std::vector<Ref>& rightRefs,
                               std::vector<Ref>& leftRefs) const {
        Split split;
        V3 dim = aabb.max - aabb.min;
        int axis = 0;
        axis = dim.y > dim[axis] ? 1 : axis;
        axis = dim.z > dim[axis] ? 2 : axis;
        float cutoff = aabb.min[axis] + dim[axis] / 2.0f;
        for (auto& ref : refs) {
            float center = (ref.aabb.max[axis] - ref.aabb.min[axis]) * 0.5f + ref.aabb.min[axis];
            if (center < cutoff) {
                grow(split.leftAABB, ref.aabb);
                leftRefs.push_back(ref);
            } else {
                grow(split.rightAABB, ref.aabb);
                rightRefs.push_back(ref);
            }
        }
        if (!leftRefs.size() || !rightRefs.size()) {
            split = Split();
            leftRefs.clear();
            rightRefs.clear();
            // Sorted by extent, not centre, exactly as BVH.cpp:832-836 does.
            std::sort(refs.begin(), refs.end(), [&](const Ref& a, const Ref& b) {
                float ea = a.aabb.max[axis] - a.aabb.min[axis];
                float eb = b.aabb.max[axis] - b.aabb.min[axis];
                return ea < eb;
            });
            uint32_t splitIdx = uint32_t(refs.size() / 2);
            for (uint32_t i = 0; i < splitIdx; i++) {
                grow(split.leftAABB, refs[i].aabb);
                leftRefs.push_back(refs[i]);
            }
            for (uint32_t i = splitIdx; i < uint32_t(refs.size()); i++) {
                grow(split.rightAABB, refs[i].aabb);
                rightRefs.push_back(refs[i]);
            }
        }
        return split;
    }

    void make_children(const Split& split, std::vector<Ref>& leftRefs, std::vector<Ref>& rightRefs) {
        if (leftRefs.size()) {
            left.reset(new Node(split.leftAABB, depth + 1, leftRefs.size(), 0.0f, binCount));
            left->build3(leftRefs);
        }
        if (rightRefs.size()) {
            right.reset(new Node(split.rightAABB, depth + 1, rightRefs.size(), 0.0f, binCount));
            right->build3(rightRefs);
        }
    }

    // 4-argument Build (BVH.cpp:249-341); only ever runs at the root.
    void build4(std::vector<Ref>& refs, const std::vector<Tri>& data) {
        if ((refs.size() <= 2 || depth >= 32) && depth > 0) {
            this->refs = refs;
            return;
        }
        Split objectSplit = find_object_split(refs);
        Split spatialSplit;
        if (depth <= 16) {
            Box overlap = objectSplit.leftAABB;
            intersect(overlap, objectSplit.rightAABB);
            if (surface_area(overlap) >= minOverlap) spatialSplit = find_spatial_split(refs, data);
        }
        std::vector<Ref> leftRefs, rightRefs;
        Split split;
        if ((objectSplit.axis < 0 || objectSplit.cost >= nodeCost) &&
            (spatialSplit.axis < 0 || spatialSplit.cost >= nodeCost)) {
            if (depth == 0) {
                split = perform_median_split(refs, rightRefs, leftRefs);
            } else {
                this->refs = refs;
                return;
            }
        } else {
            if (spatialSplit.cost < objectSplit.cost) {
                perform_spatial_split(refs, data, rightRefs, leftRefs, spatialSplit);
                split = spatialSplit;
            }
            if (objectSplit.cost <= spatialSplit.cost) {
                leftRefs.clear();
                rightRefs.clear();
                perform_object_split(refs, rightRefs, leftRefs, objectSplit);
                split = objectSplit;
            }
        }
        // "Last resort" (BVH.cpp:303-305): the node becomes a leaf, but the
        // children are still built (and later ignored by Flatten).
        if (leftRefs.empty() || rightRefs.empty()) this->refs = refs;
        refs.clear();
        make_children(split, leftRefs, rightRefs);
    }

    // 3-argument Build (BVH.cpp:343-406).
    void build3(std::vector<Ref>& refs) {
        if (refs.size() == 1) {
            this->refs = refs;
            return;
        }
        Split objectSplit = find_object_split(refs);
        std::vector<Ref> leftRefs, rightRefs;
        Split split;
        if (objectSplit.axis < 0 || objectSplit.cost >= nodeCost) {
            split = perform_median_split(refs, rightRefs, leftRefs);
        } else {
            perform_object_split(refs, rightRefs, leftRefs, objectSplit);
            split = objectSplit;
        }
        refs.clear();
        make_children(split, leftRefs, rightRefs);
    }

    // BVHBuilder::Flatten (BVH.cpp:408-441).
    void flatten(std::vector<BVHNodeOut>& nodes, std::vector<Ref>& out) {
        if (this->refs.size()) {
            for (auto& r : this->refs) r.nodeIdx = uint32_t(nodes.size() - 1);
            out.insert(out.end(), this->refs.begin(), this->refs.end());
            out[out.size() - 1].endOfNode = true;
            return;
        }
        const size_t nodeIdx = nodes.size();
        nodes.push_back(BVHNodeOut());
        if (surface_area(left->aabb) < surface_area(right->aabb)) std::swap(left, right);
        if (left) {
            bool leaf = left->refs.size() > 0;
            nodes[nodeIdx].leftMin = left->aabb.min;
            nodes[nodeIdx].leftMax = left->aabb.max;
            nodes[nodeIdx].leftPtr = leaf ? ~int32_t(out.size()) : int32_t(nodes.size());
            left->flatten(nodes, out);
        }
        if (right) {
            bool leaf = right->refs.size() > 0;
            nodes[nodeIdx].rightMin = right->aabb.min;
            nodes[nodeIdx].rightMax = right->aabb.max;
            nodes[nodeIdx].rightPtr = leaf ? ~int32_t(out.size()) : int32_t(nodes.size());
            right->flatten(nodes, out);
        }
    }
};

}  // namespace

// BVH::BVH(aabbs, data, parallelBuild) (BVH.cpp:14-56).
bool build_triangle_bvh(const std::vector<Box>& aabbs, const float* verts, uint32_t n, BuildResult& res) {
    res = BuildResult();
    if (aabbs.size() != n || n == 0) return false;
    std::vector<Tri> data(n);
    for (uint32_t i = 0; i < n; i++) {
        const float* v = verts + 9 * size_t(i);
        data[i] = Tri{V3{v[0], v[1], v[2]}, V3{v[3], v[4], v[5]}, V3{v[6], v[7], v[8]}};
    }
    std::vector<Ref> refs(n);
    for (uint32_t i = 0; i < n; i++) {
        refs[i].idx = i;
        refs[i].aabb = aabbs[i];
    }
    Box root{V3{FLT_MAX, FLT_MAX, FLT_MAX}, V3{-FLT_MAX, -FLT_MAX, -FLT_MAX}};
    for (auto& r : refs) grow(root, aabbs[r.idx]);
    float minOverlap = surface_area(root) * 10e-6f;
    Node builder(root, 0, refs.size(), minOverlap, 256);
    builder.build4(refs, data);
    // Flatten appends to the caller's (now empty) reference vector (BVH.cpp:40-42).
    std::vector<Ref>& out = refs;
    builder.flatten(res.nodes, out);
    res.refIdx.resize(out.size());
    res.endOfNode.resize(out.size());
    for (size_t i = 0; i < out.size(); i++) {
        res.refIdx[i] = out[i].idx;
        res.endOfNode[i] = out[i].endOfNode ? 1 : 0;
    }
    return true;
}

// BVH::BVH(aabbs, parallelBuild) (BVH.cpp:58-101), the TLAS builder.
bool build_aabb_bvh(const std::vector<Box>& aabbs, BuildResult& res) {
    res = BuildResult();
    uint32_t n = uint32_t(aabbs.size());
    if (n == 0) return false;
    std::vector<Ref> refs(n);
    for (uint32_t i = 0; i < n; i++) {
        refs[i].idx = i;
        refs[i].aabb = aabbs[i];
    }
    Box root{V3{FLT_MAX, FLT_MAX, FLT_MAX}, V3{-FLT_MAX, -FLT_MAX, -FLT_MAX}};
    for (auto& r : refs) grow(root, aabbs[r.idx]);
    Node builder(root, 0, refs.size(), 0.0f, 64);
    builder.build3(refs);
    if (res.nodes.empty() && n == 1) {
        BVHNodeOut node;
        node.leftPtr = ~0;
        node.rightPtr = ~0;
        node.leftMin = root.min;
        node.leftMax = root.max;
        node.rightMin = V3{0.0f, 0.0f, 0.0f};
        node.rightMax = V3{0.0f, 0.0f, 0.0f};
        res.nodes.push_back(node);
    }
    // The 3-argument Build only clears the reference vector on the inner-node
    // path, so for a single instance the original reference stays in front of
    // the flattened one (BVH.cpp:60-63,75,90 operate on the member `refs`).
    std::vector<Ref>& out = refs;
    builder.flatten(res.nodes, out);
    res.refIdx.resize(out.size());
    res.endOfNode.resize(out.size());
    for (size_t i = 0; i < out.size(); i++) {
        res.refIdx[i] = out[i].idx;
        res.endOfNode[i] = out[i].endOfNode ? 1 : 0;
    }
    return true;
}

}  // namespace oracle
