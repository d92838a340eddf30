// TEST INFRASTRUCTURE - the oracle.  Only tests/, __graft_entry__.smoke() and
// bench.py's cpu_baseline leg may load it, and only as the checker.
//
// Serial CPU restatement of the device TLAS refit (the product's
// atlas_rt_world_refit): the last build's tree and instance order are kept;
// each instance takes its new matrices and a new world box (AABB::Transform,
// src/engine/volume/AABB.cpp:35-60, over the BLAS's object box: the union of
// its root's two child boxes, or of its vertices when it is a single leaf);
// every node's child box becomes the union of the boxes below it, walking the
// tree recursively and each leaf's instances along their nextInstance chain
// (RayTracingWorld.cpp:287-295).  glm 0.9.8 min/max, no FMA contraction.
#include <cstdint>
#include <cstring>
#include <vector>

#include "../include/atlas_rt.h"

namespace {

float gmin(float a, float b) { return (b < a) ? b : a; }
float gmax(float a, float b) { return (a < b) ? b : a; }

struct B3 {
    float lo[3], hi[3];
};

B3 empty3() {
    B3 b;
    for (int k = 0; k < 3; k++) {
        b.lo[k] = 3.4028234663852886e+38f;
        b.hi[k] = -3.4028234663852886e+38f;
    }
    return b;
}

void grow(B3& b, const B3& o) {
    for (int k = 0; k < 3; k++) {
        b.hi[k] = gmax(b.hi[k], o.hi[k]);
        b.lo[k] = gmin(b.lo[k], o.lo[k]);
    }
}

struct Refit {
    atlas_rt_gpu_bvh_node* tlas;
    const atlas_rt_gpu_bvh_instance* inst;
    std::vector<B3> ibox;

    B3 child(int32_t ptr) {
        B3 b = empty3();
        if (ptr >= 0) {
            atlas_rt_gpu_bvh_node& n = tlas[ptr];
            const B3 l = child(n.left_ptr);
            B3 r;
            if (n.right_ptr < 0 && n.right_ptr == n.left_ptr) {
                // single instance (BVH.cpp:80-88): both children are the one
                // leaf and the second box stays as the builder left it (zero)
                std::memcpy(r.lo, n.right_min, 12);
                std::memcpy(r.hi, n.right_max, 12);
            } else {
                r = child(n.right_ptr);
            }
            std::memcpy(n.left_min, l.lo, 12);
            std::memcpy(n.left_max, l.hi, 12);
            std::memcpy(n.right_min, r.lo, 12);
            std::memcpy(n.right_max, r.hi, 12);
            grow(b, l);
            grow(b, r);
        } else {
            for (int32_t s = ~ptr;; s++) {
                grow(b, ibox[size_t(s)]);
                if (inst[s].next_instance < 0) break;
            }
        }
        return b;
    }
};

}  // namespace

extern "C" {

// tlas (node_count nodes), inst and xform (12 floats per slot) are updated in
// place; order[slot] = caller index into desc; blas_nodes / blas_tris with
// their counts: every BLAS of the world's table; boxes_out (may be NULL): the
// instances' new world boxes in slot order.
void oracle_world_refit(atlas_rt_gpu_bvh_node* tlas, uint32_t node_count, atlas_rt_gpu_bvh_instance* inst,
                        float* xform, const uint32_t* order, uint32_t count, const atlas_rt_instance_desc* desc,
                        const atlas_rt_gpu_bvh_node* const* blas_nodes, const uint32_t* blas_node_counts,
                        const atlas_rt_gpu_bvh_triangle* const* blas_tris, const uint32_t* blas_tri_counts,
                        float* boxes_out) {
    Refit R{tlas, inst, std::vector<B3>(count)};
    for (uint32_t s = 0; s < count; s++) {
        const atlas_rt_instance_desc& d = desc[order[s]];
        std::memcpy(inst[s].inverse_matrix, d.world_to_object, 48);
        std::memcpy(&xform[12 * size_t(s)], d.object_to_world, 48);
        const int32_t k = inst[s].mesh_offset;
        B3 ob = empty3();
        if (blas_node_counts[k]) {
            const atlas_rt_gpu_bvh_node& r = blas_nodes[k][0];
            B3 l, q;
            std::memcpy(l.lo, r.left_min, 12);
            std::memcpy(l.hi, r.left_max, 12);
            std::memcpy(q.lo, r.right_min, 12);
            std::memcpy(q.hi, r.right_max, 12);
            grow(ob, l);
            grow(ob, q);
        } else {
            for (uint32_t t = 0; t < blas_tri_counts[k]; t++) {
                const float* v[3] = {blas_tris[k][t].v0, blas_tris[k][t].v1, blas_tris[k][t].v2};
                for (int c = 0; c < 3; c++)
                    for (int a = 0; a < 3; a++) {
                        ob.hi[a] = gmax(v[c][a], ob.hi[a]);
                        ob.lo[a] = gmin(v[c][a], ob.lo[a]);
                    }
            }
        }
        B3& w = R.ibox[s];
        for (int c = 0; c < 8; c++) {
            const float p[3] = {(c & 2) ? ob.hi[0] : ob.lo[0], (c & 4) ? ob.hi[1] : ob.lo[1],
                                (c & 1) ? ob.hi[2] : ob.lo[2]};
            for (int r = 0; r < 3; r++) {
                const float* m = d.object_to_world[r];
                const float x = (m[0] * p[0] + m[1] * p[1]) + (m[2] * p[2] + m[3] * 1.0f);
                w.lo[r] = c == 0 ? x : gmin(w.lo[r], x);
                w.hi[r] = c == 0 ? x : gmax(w.hi[r], x);
            }
        }
    }
    if (node_count) R.child(0);
    if (boxes_out)  // per slot: min xyz, max xyz
        for (uint32_t s = 0; s < count; s++) {
            std::memcpy(&boxes_out[6 * size_t(s)], R.ibox[s].lo, 12);
            std::memcpy(&boxes_out[6 * size_t(s) + 3], R.ibox[s].hi, 12);
        }
}

}  // extern "C"
