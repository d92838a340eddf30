// Host BVH builder (see builder.h for the reference rules it reproduces).
//
// Structure: a tree of BuildNode objects filled by tasks.  Subtrees at depth
// <= kParallelDepth are handed to a small std::thread pool (the engine spawns
// jobs to the same depth, BVH.cpp:310-328); deeper subtrees recurse inline.
// Every node keeps its references in the order the engine would hold them, so
// the order-sensitive parts (min/max of signed zeros, std::sort in the median
// fallback) resolve identically.
#include "builder.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <functional>
#include <memory>
#include <mutex>
#include <thread>

#include "rt_hostmath.h"

namespace atlas_rt {

namespace {

constexpr uint32_t kParallelDepth = 6;

struct Ref {
    uint32_t idx;
    HBox box;
};

struct Cand {
    float cost = FLT_MAX;
    int32_t axis = -1;
    uint32_t bin = 0;
    float pos = 0.0f;
    HBox left = empty_box();
    HBox right = empty_box();
};

struct BuildNode {
    HBox box;
    uint32_t depth = 0;
    float cost = 0.0f;             // refs * area (BVH.cpp:230)
    std::vector<Ref> leaf;         // non-empty => leaf
    std::unique_ptr<BuildNode> side[2];  // [0] = split's left side, [1] = right side
};

// Minimal task pool; the calling thread helps until the pool drains.
class Pool {
  public:
    explicit Pool(int threads) {
        for (int i = 0; i + 1 < threads; i++) workers_.emplace_back([this] { loop(); });
    }
    ~Pool() {
        {
            std::lock_guard<std::mutex> g(m_);
            stop_ = true;
        }
        cv_.notify_all();
        for (auto& t : workers_) t.join();
    }
    bool parallel() const { return !workers_.empty(); }
    void submit(std::function<void()> f) {
        pending_.fetch_add(1);
        {
            std::lock_guard<std::mutex> g(m_);
            q_.push_back(std::move(f));
        }
        cv_.notify_one();
    }
    void drain() {
        while (pending_.load() != 0) {
            std::function<void()> f;
            {
                std::unique_lock<std::mutex> g(m_);
                if (q_.empty()) {
                    done_cv_.wait_for(g, std::chrono::milliseconds(1));
                    continue;
                }
                f = std::move(q_.front());
                q_.pop_front();
            }
            run(f);
        }
    }

  private:
    void run(std::function<void()>& f) {
        f();
        if (pending_.fetch_sub(1) == 1) done_cv_.notify_all();
    }
    void loop() {
        for (;;) {
            std::function<void()> f;
            {
                std::unique_lock<std::mutex> g(m_);
                cv_.wait(g, [this] { return stop_ || !q_.empty(); });
                if (stop_ && q_.empty()) return;
                f = std::move(q_.front());
                q_.pop_front();
            }
            run(f);
        }
    }
    std::vector<std::thread> workers_;
    std::deque<std::function<void()>> q_;
    std::mutex m_;
    std::condition_variable cv_, done_cv_;
    std::atomic<int> pending_{0};
    bool stop_ = false;
};

struct Builder {
    const float* verts = nullptr;  // 9 floats per triangle; null for the AABB builder
    uint32_t bin_base = 256;       // 256 for BLAS (BVH.cpp:34), 64 for TLAS (BVH.cpp:72)
    float min_overlap = 0.0f;
    Pool* pool = nullptr;

    uint32_t bins_for(uint32_t depth) const { return std::max(bin_base / (depth + 1), 16u); }

    static uint32_t bin_index(float x, float start, float inv, uint32_t B) {
        return uint32_t(gl_clamp((x - start) * inv, 0.0f, float(B) - 1.0f));
    }

    // Prefix/suffix sweep over the bins (BVH.cpp:484-522 / 621-661).  `enter`
    // and `leave` give the left count increment and right count decrement of
    // each bin (equal to the bin count for object splits).
    static void sweep(const std::vector<HBox>& bb, const std::vector<uint32_t>& enter,
                      const std::vector<uint32_t>& leave, uint32_t total, int axis, float start, float binSize,
                      Cand& best) {
        const uint32_t B = uint32_t(bb.size());
        std::vector<HBox> suffix(B);
        HBox acc = empty_box();
        for (uint32_t j = B - 1; j > 0; j--) {
            box_grow(acc, bb[j]);
            suffix[j - 1] = acc;
        }
        HBox pre = empty_box();
        uint32_t nl = 0, nr = total;
        for (uint32_t j = 1; j < B; j++) {
            box_grow(pre, bb[j - 1]);
            nl += enter[j - 1];
            nr -= leave[j - 1];
            if (!nl || !nr) continue;
            const float c = box_area(pre) * float(nl) + box_area(suffix[j - 1]) * float(nr);
            if (c < best.cost) {
                best.cost = c;
                best.axis = axis;
                best.bin = j;
                best.pos = start + float(j) * binSize;
                best.left = pre;
                best.right = suffix[j - 1];
            }
        }
    }

    // Centroid-binned SAH (FindObjectSplit, BVH.cpp:444-528).
    Cand object_split(const BuildNode& n, const std::vector<Ref>& refs) const {
        Cand best;
        const uint32_t B = bins_for(n.depth);
        std::vector<HBox> bb(B);
        std::vector<uint32_t> cnt(B);
        for (int a = 0; a < 3; a++) {
            const float start = n.box.lo[a], stop = n.box.hi[a];
            if (std::fabs(stop - start) < 1e-3f) continue;
            const float binSize = (stop - start) / float(B);
            const float inv = 1.0f / binSize;
            std::fill(bb.begin(), bb.end(), empty_box());
            std::fill(cnt.begin(), cnt.end(), 0u);
            for (const Ref& r : refs) {
                const float c = 0.5f * (r.box.lo[a] + r.box.hi[a]);
                const uint32_t k = bin_index(c, start, inv, B);
                cnt[k]++;
                box_grow(bb[k], r.box);
            }
            sweep(bb, cnt, cnt, uint32_t(refs.size()), a, start, binSize, best);
        }
        return best;
    }

    // PerformObjectSplit (BVH.cpp:530-556).
    void apply_object_split(const BuildNode& n, const std::vector<Ref>& refs, const Cand& s,
                            std::vector<Ref>& L, std::vector<Ref>& R) const {
        const uint32_t B = bins_for(n.depth);
        const float start = n.box.lo[s.axis], stop = n.box.hi[s.axis];
        const float binSize = (stop - start) / float(B);
        const float inv = 1.0f / binSize;
        for (const Ref& r : refs) {
            const float c = 0.5f * (r.box.lo[s.axis] + r.box.hi[s.axis]);
            (bin_index(c, start, inv, B) < s.bin ? L : R).push_back(r);
        }
    }

    // Clip one triangle reference at an axis-aligned plane (SplitReference, BVH.cpp:760-798).
    void clip_ref(uint32_t tri, const HBox& cur, float plane, int a, HBox& lbox, HBox& rbox) const {
        const float* v = verts + 9 * size_t(tri);
        lbox = empty_box();
        rbox = empty_box();
        for (int e = 0; e < 3; e++) {
            const float* p0 = v + 3 * e;
            const float* p1 = v + 3 * ((e + 1) % 3);
            if ((p0[a] < plane && p1[a] > plane) || (p0[a] > plane && p1[a] < plane)) {
                const float t = gl_clamp((plane - p0[a]) / (p1[a] - p0[a]), 0.0f, 1.0f);
                float q[3];
                for (int k = 0; k < 3; k++) q[k] = p0[k] + t * (p1[k] - p0[k]);
                box_grow_point(lbox, q);
                box_grow_point(rbox, q);
            }
            if (p0[a] <= plane) box_grow_point(lbox, p0);
            if (p0[a] >= plane) box_grow_point(rbox, p0);
        }
        lbox.hi[a] = plane;
        rbox.lo[a] = plane;
        box_clip(lbox, cur);
        box_clip(rbox, cur);
    }

    // SBVH spatial binning (FindSpatialSplit, BVH.cpp:558-667).
    Cand spatial_split(const BuildNode& n, const std::vector<Ref>& refs) const {
        Cand best;
        const uint32_t B = bins_for(n.depth);
        std::vector<HBox> bb(B);
        std::vector<uint32_t> enter(B), leave(B);
        for (int a = 0; a < 3; a++) {
            const float start = n.box.lo[a], stop = n.box.hi[a];
            if (std::fabs(stop - start) < 1e-3f) continue;
            const float binSize = (stop - start) / float(B);
            const float inv = 1.0f / binSize;
            std::fill(bb.begin(), bb.end(), empty_box());
            std::fill(enter.begin(), enter.end(), 0u);
            std::fill(leave.begin(), leave.end(), 0u);
            for (const Ref& r : refs) {
                const uint32_t b0 = bin_index(r.box.lo[a], start, inv, B);
                const uint32_t b1 = bin_index(r.box.hi[a], start, inv, B);
                if (b0 == b1) {
                    enter[b0]++;
                    leave[b0]++;
                    box_grow(bb[b0], r.box);
                    continue;
                }
                HBox cur = r.box;
                for (uint32_t j = b0; j < b1; j++) {
                    HBox lb, rb;
                    clip_ref(r.idx, cur, start + float(j + 1) * binSize, a, lb, rb);
                    box_grow(bb[j], lb);
                    cur = rb;
                }
                box_grow(bb[b1], cur);
                enter[b0]++;
                leave[b1]++;
            }
            sweep(bb, enter, leave, uint32_t(refs.size()), a, start, binSize, best);
        }
        return best;
    }

    // PerformSpatialSplit with reference unsplitting (BVH.cpp:669-758); s is updated.
    void apply_spatial_split(const BuildNode& n, const std::vector<Ref>& refs, Cand& s, std::vector<Ref>& L,
                             std::vector<Ref>& R) const {
        const uint32_t B = bins_for(n.depth);
        const int a = s.axis;
        const float start = n.box.lo[a], stop = n.box.hi[a];
        const float binSize = (stop - start) / float(B);
        const float inv = 1.0f / binSize;
        std::vector<uint8_t> straddles(refs.size(), 0);
        for (size_t i = 0; i < refs.size(); i++) {
            const Ref& r = refs[i];
            const uint32_t b0 = bin_index(r.box.lo[a], start, inv, B);
            const uint32_t b1 = bin_index(r.box.hi[a], start, inv, B);
            if (b1 < s.bin) {
                L.push_back(r);
                box_grow(s.left, r.box);
            } else if (b0 >= s.bin) {
                R.push_back(r);
                box_grow(s.right, r.box);
            } else {
                straddles[i] = 1;
            }
        }
        for (size_t i = 0; i < refs.size(); i++) {
            if (!straddles[i]) continue;
            const Ref& r = refs[i];
            HBox lb, rb;
            clip_ref(r.idx, r.box, s.pos, a, lb, rb);
            HBox keepL = s.left, keepR = s.right, dupL = s.left, dupR = s.right;
            box_grow(keepL, r.box);
            box_grow(keepR, r.box);
            box_grow(dupL, lb);
            box_grow(dupR, rb);
            const float nL = float(L.size()), nR = float(R.size());
            const float nL1 = float(L.size() + 1), nR1 = float(R.size() + 1);
            const float costKeepL = box_area(keepL) * nL1 + box_area(s.right) * nR;
            const float costKeepR = box_area(s.left) * nL + box_area(keepR) * nR1;
            const float costDup = box_area(dupL) * nL1 + box_area(dupR) * nR1;
            const float m = gl_min(costDup, gl_min(costKeepR, costKeepL));
            if (m == costKeepL) {
                s.left = keepL;
                L.push_back(r);
            } else if (m == costKeepR) {
                s.right = keepR;
                R.push_back(r);
            } else {
                s.left = dupL;
                s.right = dupR;
                L.push_back(Ref{r.idx, lb});
                R.push_back(Ref{r.idx, rb});
            }
        }
    }

    // PerformMedianSplit (BVH.cpp:800-855): spatial median on the longest axis,
    // falling back to sorting by extent and halving.
    Cand median_split(const BuildNode& n, std::vector<Ref>& refs, std::vector<Ref>& L, std::vector<Ref>& R) const {
        Cand s;
        const float dim[3] = {n.box.hi[0] - n.box.lo[0], n.box.hi[1] - n.box.lo[1], n.box.hi[2] - n.box.lo[2]};
        int a = 0;
        a = dim[1] > dim[a] ? 1 : a;
        a = dim[2] > dim[a] ? 2 : a;
        const float cut = n.box.lo[a] + dim[a] / 2.0f;
        for (const Ref& r : refs) {
            const float c = (r.box.hi[a] - r.box.lo[a]) * 0.5f + r.box.lo[a];
            if (c < cut) {
                box_grow(s.left, r.box);
                L.push_back(r);
            } else {
                box_grow(s.right, r.box);
                R.push_back(r);
            }
        }
        if (L.empty() || R.empty()) {
            s = Cand();
            L.clear();
            R.clear();
            std::sort(refs.begin(), refs.end(), [a](const Ref& x, const Ref& y) {
                return (x.box.hi[a] - x.box.lo[a]) < (y.box.hi[a] - y.box.lo[a]);
            });
            const uint32_t half = uint32_t(refs.size() / 2);
            for (uint32_t i = 0; i < uint32_t(refs.size()); i++) {
                if (i < half) {
                    box_grow(s.left, refs[i].box);
                    L.push_back(refs[i]);
                } else {
                    box_grow(s.right, refs[i].box);
                    R.push_back(refs[i]);
                }
            }
        }
        return s;
    }

    void spawn_children(BuildNode& n, const Cand& s, std::vector<Ref>&& L, std::vector<Ref>&& R) {
        std::vector<Ref>* sides[2] = {&L, &R};
        const HBox boxes[2] = {s.left, s.right};
        const bool par = pool && pool->parallel() && n.depth <= kParallelDepth;
        for (int k = 0; k < 2; k++) {
            if (sides[k]->empty()) continue;
            n.side[k].reset(new BuildNode());
            BuildNode* c = n.side[k].get();
            c->box = boxes[k];
            c->depth = n.depth + 1;
            c->cost = float(sides[k]->size()) * box_area(c->box);
            if (par) {
                auto refs = std::make_shared<std::vector<Ref>>(std::move(*sides[k]));
                pool->submit([this, c, refs] { build_inner(*c, *refs); });
            } else {
                build_inner(*c, *sides[k]);
            }
        }
    }

    // Root of a triangle BVH (the 4-argument Build, BVH.cpp:249-341).
    void build_root(BuildNode& n, std::vector<Ref>& refs) {
        Cand obj = object_split(n, refs);
        Cand spa;
        HBox ov = obj.left;
        box_clip(ov, obj.right);
        if (box_area(ov) >= min_overlap) spa = spatial_split(n, refs);
        std::vector<Ref> L, R;
        L.reserve(refs.size());
        R.reserve(refs.size());
        Cand s;
        const bool objBad = obj.axis < 0 || obj.cost >= n.cost;
        const bool spaBad = spa.axis < 0 || spa.cost >= n.cost;
        if (objBad && spaBad) {
            s = median_split(n, refs, L, R);
        } else if (spa.cost < obj.cost) {
            apply_spatial_split(n, refs, spa, L, R);
            s = spa;
        } else {
            apply_object_split(n, refs, obj, L, R);
            s = obj;
        }
        // A split with an empty side leaves the node a leaf (BVH.cpp:303-305);
        // its children are still built but never reached by Flatten.
        if (L.empty() || R.empty()) n.leaf = refs;
        refs.clear();
        spawn_children(n, s, std::move(L), std::move(R));
    }

    // Every non-root node (3-argument Build, BVH.cpp:343-406).
    void build_inner(BuildNode& n, std::vector<Ref>& refs) {
        if (refs.size() == 1) {
            n.leaf = refs;
            return;
        }
        Cand obj = object_split(n, refs);
        std::vector<Ref> L, R;
        L.reserve(refs.size() / 2);
        R.reserve(refs.size() / 2);
        Cand s;
        if (obj.axis < 0 || obj.cost >= n.cost) {
            s = median_split(n, refs, L, R);
        } else {
            apply_object_split(n, refs, obj, L, R);
            s = obj;
        }
        refs.clear();
        refs.shrink_to_fit();
        spawn_children(n, s, std::move(L), std::move(R));
    }
};

void emit_box(const HBox& b, float* lo, float* hi) {
    for (int k = 0; k < 3; k++) {
        lo[k] = b.lo[k];
        hi[k] = b.hi[k];
    }
}

// Pre-order flattening (BVHBuilder::Flatten, BVH.cpp:408-441).  `prefix` holds
// references that precede the flattened ones (the single-instance TLAS case).
void flatten(BuildNode& root, std::vector<Ref>& prefix, BuildOutput& out) {
    out.ref_idx.clear();
    out.end_of_node.clear();
    for (const Ref& r : prefix) {
        out.ref_idx.push_back(r.idx);
        out.end_of_node.push_back(0);
    }
    // Recursive order: node, then left subtree, then right subtree.  Inner
    // node pointers are the index the child will get when it is emitted.
    std::function<void(BuildNode*, uint32_t)> visit = [&](BuildNode* n, uint32_t depth) {
        if (!n->leaf.empty()) {
            for (const Ref& r : n->leaf) {
                out.ref_idx.push_back(r.idx);
                out.end_of_node.push_back(0);
            }
            out.end_of_node.back() = 1;
            if (depth > out.max_depth) out.max_depth = depth;
            return;
        }
        const size_t me = out.nodes.size();
        out.nodes.push_back(atlas_rt_gpu_bvh_node{});
        std::memset(&out.nodes[me], 0, sizeof(atlas_rt_gpu_bvh_node));
        if (box_area(n->side[0]->box) < box_area(n->side[1]->box)) std::swap(n->side[0], n->side[1]);
        for (int k = 0; k < 2; k++) {
            BuildNode* c = n->side[k].get();
            if (!c) continue;
            const bool leaf = !c->leaf.empty();
            const int32_t ptr = leaf ? ~int32_t(out.ref_idx.size()) : int32_t(out.nodes.size());
            atlas_rt_gpu_bvh_node& g = out.nodes[me];
            if (k == 0) {
                emit_box(c->box, g.left_min, g.left_max);
                g.left_ptr = ptr;
            } else {
                emit_box(c->box, g.right_min, g.right_max);
                g.right_ptr = ptr;
            }
            visit(c, depth + 1);
        }
    };
    visit(&root, 0);
}

}  // namespace

bool build_triangle_bvh(const float* boxes, const float* verts, uint32_t count, int threads, BuildOutput& out) {
    out = BuildOutput();
    if (!count || !verts || !boxes) return false;
    std::vector<Ref> refs(count);
    HBox root = empty_box();
    for (uint32_t i = 0; i < count; i++) {
        refs[i].idx = i;
        std::memcpy(refs[i].box.lo, boxes + 6 * size_t(i), 3 * sizeof(float));
        std::memcpy(refs[i].box.hi, boxes + 6 * size_t(i) + 3, 3 * sizeof(float));
        box_grow(root, refs[i].box);
    }
    Pool pool(threads);
    Builder b;
    b.verts = verts;
    b.bin_base = 256;
    b.min_overlap = box_area(root) * 10e-6f;  // BVH.cpp:33
    b.pool = &pool;
    BuildNode n;
    n.box = root;
    n.depth = 0;
    n.cost = float(count) * box_area(root);
    b.build_root(n, refs);
    pool.drain();
    std::vector<Ref> none;
    flatten(n, none, out);
    return true;
}

bool build_aabb_bvh(const float* boxes, uint32_t count, int threads, BuildOutput& out) {
    out = BuildOutput();
    if (!count || !boxes) return false;
    std::vector<Ref> refs(count);
    HBox root = empty_box();
    for (uint32_t i = 0; i < count; i++) {
        refs[i].idx = i;
        std::memcpy(refs[i].box.lo, boxes + 6 * size_t(i), 3 * sizeof(float));
        std::memcpy(refs[i].box.hi, boxes + 6 * size_t(i) + 3, 3 * sizeof(float));
        box_grow(root, refs[i].box);
    }
    Pool pool(threads);
    Builder b;
    b.verts = nullptr;
    b.bin_base = 64;
    b.pool = &pool;
    BuildNode n;
    n.box = root;
    n.depth = 0;
    n.cost = float(count) * box_area(root);
    // The engine's 3-argument Build leaves its input vector untouched when the
    // root is a leaf, so a single instance appears twice in the reference list,
    // and BVH.cpp:80-88 adds a node whose both children point at it.
    std::vector<Ref> keep = (count == 1) ? refs : std::vector<Ref>();
    b.build_inner(n, refs);
    pool.drain();
    if (count == 1) {
        atlas_rt_gpu_bvh_node g;
        std::memset(&g, 0, sizeof(g));
        emit_box(root, g.left_min, g.left_max);
        g.left_ptr = ~0;
        g.right_ptr = ~0;
        out.nodes.push_back(g);
    }
    flatten(n, keep, out);
    if (count == 1) out.max_depth = 1;
    return true;
}

bool validate_tree(const atlas_rt_gpu_bvh_node* nodes, uint32_t node_count, uint32_t leaf_items,
                   const uint8_t* end_flags, uint32_t* max_depth) {
    std::vector<uint32_t> depth(node_count, 0);
    uint32_t md = node_count ? 0 : 1;
    for (uint32_t i = 0; i < node_count; i++) {
        const int32_t ptr[2] = {nodes[i].left_ptr, nodes[i].right_ptr};
        for (int k = 0; k < 2; k++) {
            if (ptr[k] >= 0) {
                if (uint32_t(ptr[k]) >= node_count || uint32_t(ptr[k]) <= i) return false;
                depth[ptr[k]] = depth[i] + 1;
            } else {
                const uint32_t first = uint32_t(~ptr[k]);
                if (first >= leaf_items) return false;
                if (end_flags) {
                    uint32_t j = first;
                    while (j < leaf_items && !end_flags[j]) j++;
                    if (j >= leaf_items) return false;
                }
                md = std::max(md, depth[i] + 1);
            }
        }
    }
    if (max_depth) *max_depth = md;
    return true;
}

}  // namespace atlas_rt
