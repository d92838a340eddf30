// World objects: the software-ray-tracing half of RayTracingWorld
// (src/engine/raytracing/RayTracingWorld.cpp:32-162, 267-307) and the C ABI of
// the two-level traces.
#include "world.h"

#include <algorithm>
#include <cstring>

#include "builder.h"
#include "rt_hostmath.h"

namespace atlas_rt {
int default_threads();
}

using namespace atlas_rt;

void atlas_rt_world::free_device() {
    void* ptrs[] = {d_tlas, d_instances, d_table, d_order, d_xform, d_blas_box, d_inst_box, d_refit_leaves,
                    d_refit_nodes};
    for (void* p : ptrs)
        if (p) hipFree(p);
    d_tlas = d_instances = d_table = d_order = d_xform = nullptr;
    d_blas_box = d_inst_box = d_refit_leaves = d_refit_nodes = nullptr;
    refit_level_off.clear();
    host_stale = false;
}

atlas_rt_status atlas_rt_world::sync_host() {
    if (!host_stale) return ATLAS_RT_OK;
    ATLAS_HIP_TRY(hipDeviceSynchronize());
    ATLAS_HIP_TRY(hipMemcpy(tlas.data(), d_tlas, tlas.size() * sizeof(tlas[0]), hipMemcpyDeviceToHost));
    ATLAS_HIP_TRY(
        hipMemcpy(instances.data(), d_instances, instances.size() * sizeof(instances[0]), hipMemcpyDeviceToHost));
    ATLAS_HIP_TRY(hipMemcpy(xform.data(), d_xform, xform.size() * sizeof(float), hipMemcpyDeviceToHost));
    host_stale = false;
    return ATLAS_RT_OK;
}

// The object-space box AABB::Transform starts from: the union of the root's
// two child boxes, or of the vertices of a BLAS that is a single leaf.
static HBox blas_object_box(const atlas_rt_gpu_bvh_node* root, uint32_t root_count,
                            const atlas_rt_gpu_bvh_triangle* tris, uint32_t tri_count) {
    HBox b = empty_box();
    if (root_count) {
        HBox l, r;
        for (int k = 0; k < 3; k++) {
            l.lo[k] = root->left_min[k];
            l.hi[k] = root->left_max[k];
            r.lo[k] = root->right_min[k];
            r.hi[k] = root->right_max[k];
        }
        box_grow(b, l);
        box_grow(b, r);
    } else {
        for (uint32_t i = 0; i < tri_count; i++) {
            box_grow_point(b, tris[i].v0);
            box_grow_point(b, tris[i].v1);
            box_grow_point(b, tris[i].v2);
        }
    }
    return b;
}

// AABB::Transform (AABB.cpp:35-60) with glm's mat4 * vec4 evaluation order
// ((m0*x + m1*y) + (m2*z + m3*w)).
static void transform_box(const float m[3][4], const atlas_rt_gpu_bvh_node* root, uint32_t root_count,
                          const atlas_rt_gpu_bvh_triangle* tris, uint32_t tri_count, float out[6]) {
    const HBox b = blas_object_box(root, root_count, tris, tri_count);
    const float cx[8] = {b.lo[0], b.lo[0], b.hi[0], b.hi[0], b.lo[0], b.lo[0], b.hi[0], b.hi[0]};
    const float cy[8] = {b.lo[1], b.lo[1], b.lo[1], b.lo[1], b.hi[1], b.hi[1], b.hi[1], b.hi[1]};
    const float cz[8] = {b.lo[2], b.hi[2], b.lo[2], b.hi[2], b.lo[2], b.hi[2], b.lo[2], b.hi[2]};
    float lo[3] = {0.0f, 0.0f, 0.0f}, hi[3] = {0.0f, 0.0f, 0.0f};  // set at i == 0
    for (int i = 0; i < 8; i++) {
        float p[3];
        for (int r = 0; r < 3; r++)
            p[r] = (m[r][0] * cx[i] + m[r][1] * cy[i]) + (m[r][2] * cz[i] + m[r][3] * 1.0f);
        for (int r = 0; r < 3; r++) {
            if (i == 0) {
                lo[r] = hi[r] = p[r];
            } else {
                lo[r] = gl_min(lo[r], p[r]);
                hi[r] = gl_max(hi[r], p[r]);
            }
        }
    }
    for (int r = 0; r < 3; r++) {
        out[r] = lo[r];
        out[3 + r] = hi[r];
    }
}

static atlas_rt_status check_world(const atlas_rt_world* w, const char* who) {
    if (!w) return fail(ATLAS_RT_E_INVALID_ARG, std::string(who) + ": NULL world");
    if (!w->valid) return fail(ATLAS_RT_E_INVALID_ARG, std::string(who) + ": world has no instances (call update)");
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev != w->device)
        return fail(ATLAS_RT_E_INVALID_ARG, std::string(who) + ": world lives on another device");
    return ATLAS_RT_OK;
}

template <class T>
static atlas_rt_status upload_vec(void*& dst, const std::vector<T>& v) {
    if (dst) hipFree(dst);
    dst = nullptr;
    if (v.empty()) return ATLAS_RT_OK;
    ATLAS_HIP_TRY(hipMalloc(&dst, v.size() * sizeof(T)));
    ATLAS_HIP_TRY(hipMemcpy(dst, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
    return ATLAS_RT_OK;
}

extern "C" {

atlas_rt_status atlas_rt_aabb_tree_build(const atlas_rt_aabb* aabbs, uint32_t count, uint32_t flags,
                                         atlas_rt_gpu_bvh_node* nodes, uint32_t* node_count, uint32_t* ref_idx,
                                         uint8_t* end_of_node, uint32_t* ref_count) {
    if (!aabbs || !node_count || !ref_count) return fail(ATLAS_RT_E_INVALID_ARG, "atlas_rt_aabb_tree_build: NULL argument");
    if (count == 0) return fail(ATLAS_RT_E_EMPTY, "atlas_rt_aabb_tree_build: no items");
    std::vector<float> boxes(6 * size_t(count));
    for (uint32_t i = 0; i < count; i++) {
        std::memcpy(&boxes[6 * size_t(i)], aabbs[i].min, 12);
        std::memcpy(&boxes[6 * size_t(i) + 3], aabbs[i].max, 12);
    }
    BuildOutput bo;
    const int threads = (flags & ATLAS_RT_BUILD_PARALLEL) ? default_threads() : 1;
    if (!build_aabb_bvh(boxes.data(), count, threads, bo)) return fail(ATLAS_RT_E_EMPTY, "atlas_rt_aabb_tree_build");
    const uint32_t nn = uint32_t(bo.nodes.size()), nr = uint32_t(bo.ref_idx.size());
    const bool fits = *node_count >= nn && *ref_count >= nr;
    *node_count = nn;
    *ref_count = nr;
    if (!fits) return fail(ATLAS_RT_E_INVALID_ARG, "atlas_rt_aabb_tree_build: output capacity too small");
    if (nodes) std::memcpy(nodes, bo.nodes.data(), nn * sizeof(atlas_rt_gpu_bvh_node));
    if (ref_idx) std::memcpy(ref_idx, bo.ref_idx.data(), nr * sizeof(uint32_t));
    if (end_of_node) std::memcpy(end_of_node, bo.end_of_node.data(), nr);
    return ATLAS_RT_OK;
}

atlas_rt_status atlas_rt_world_create(atlas_rt_world** out) {
    if (!out) return fail(ATLAS_RT_E_INVALID_ARG, "atlas_rt_world_create: out is NULL");
    *out = new atlas_rt_world();
    return ATLAS_RT_OK;
}

atlas_rt_status atlas_rt_world_destroy(atlas_rt_world* w) {
    if (!w) return ATLAS_RT_OK;
    w->free_device();
    w->spill.release();
    w->flags.release();
    w->alpha_tab.release();
    w->timing.release();
    delete w;
    return ATLAS_RT_OK;
}

atlas_rt_status atlas_rt_world_update(atlas_rt_world* w, const atlas_rt_instance_desc* inst, uint32_t n,
                                      const atlas_rt_aabb* world_aabbs, const atlas_rt_blas* const* blases,
                                      uint32_t blas_count) {
    if (!w) return fail(ATLAS_RT_E_INVALID_ARG, "atlas_rt_world_update: NULL world");
    if (n == 0) return fail(ATLAS_RT_E_EMPTY, "atlas_rt_world_update: no instances");
    if (!inst || !blases || blas_count == 0) return fail(ATLAS_RT_E_INVALID_ARG, "atlas_rt_world_update: NULL array");
    int dev = 0;
    ATLAS_HIP_TRY(hipGetDevice(&dev));
    std::vector<BlasDev> table(blas_count);
    uint32_t blas_depth = 0;
    for (uint32_t k = 0; k < blas_count; k++) {
        const atlas_rt_blas* b = blases[k];
        if (!b || !b->on_device || b->device != dev)
            return fail(ATLAS_RT_E_INVALID_ARG, "atlas_rt_world_update: BLAS " + std::to_string(k) +
                                                    " is NULL, host-only or on another device");
        table[k].nodes = b->d_nodes;
        table[k].tris = b->d_tris;
        table[k].vtx = b->d_vtx;
        table[k].nodes4 = b->d_nodes4;
        table[k].nodes8 = b->d_nodes8;
        table[k].node_count = uint32_t(b->nodes.size());
        table[k].tri_count = uint32_t(b->tris.size());
        blas_depth = std::max(blas_depth, b->max_depth);
    }
    std::vector<float> boxes(6 * size_t(n));
    for (uint32_t i = 0; i < n; i++) {
        if (inst[i].blas >= blas_count)
            return fail(ATLAS_RT_E_INVALID_ARG, "atlas_rt_world_update: instance " + std::to_string(i) +
                                                    " names BLAS " + std::to_string(inst[i].blas));
        if (world_aabbs) {
            std::memcpy(&boxes[6 * size_t(i)], world_aabbs[i].min, 12);
            std::memcpy(&boxes[6 * size_t(i) + 3], world_aabbs[i].max, 12);
        } else {
            const atlas_rt_blas* b = blases[inst[i].blas];
            transform_box(inst[i].object_to_world, b->nodes.empty() ? nullptr : &b->nodes[0],
                          uint32_t(b->nodes.size()), b->tris.data(), uint32_t(b->tris.size()), &boxes[6 * size_t(i)]);
        }
    }
    BuildOutput bo;
    if (!build_aabb_bvh(boxes.data(), n, n > 4096 ? default_threads() : 1, bo))
        return fail(ATLAS_RT_E_EMPTY, "atlas_rt_world_update: TLAS build failed");
    // Instances in TLAS leaf order, nextInstance chaining within a leaf
    // (RayTracingWorld.cpp:287-295).
    const size_t m = bo.ref_idx.size();
    std::vector<atlas_rt_gpu_bvh_instance> packed(m);
    std::vector<uint32_t> order(m);
    std::vector<float> xform(12 * m);
    for (size_t i = 0; i < m; i++) {
        const atlas_rt_instance_desc& d = inst[bo.ref_idx[i]];
        std::memcpy(&xform[12 * i], d.object_to_world, 12 * sizeof(float));
        atlas_rt_gpu_bvh_instance& g = packed[i];
        std::memcpy(g.inverse_matrix, d.world_to_object, sizeof(g.inverse_matrix));
        g.mesh_offset = int32_t(d.blas);
        g.material_offset = d.material_offset;
        g.next_instance = bo.end_of_node[i] ? -1 : int32_t(i) + 1;
        g.mask = d.mask;
        order[i] = bo.ref_idx[i];
    }
    // traces and frames in flight on any stream still read the arrays freed here
    ATLAS_HIP_TRY(hipDeviceSynchronize());
    w->free_device();
    w->valid = false;
    w->generation++;
    atlas_rt_status st = upload_vec(w->d_tlas, bo.nodes);
    if (st == ATLAS_RT_OK) st = upload_vec(w->d_instances, packed);
    if (st == ATLAS_RT_OK) st = upload_vec(w->d_table, table);
    if (st == ATLAS_RT_OK) st = upload_vec(w->d_order, order);
    if (st == ATLAS_RT_OK) st = upload_vec(w->d_xform, xform);
    // refit schedule: BLAS object boxes, per-node leaf ranges, nodes by depth
    std::vector<float> blas_box(8 * size_t(blas_count), 0.0f);
    for (uint32_t k = 0; k < blas_count; k++) {
        const atlas_rt_blas* b = blases[k];
        const HBox ob = blas_object_box(b->nodes.empty() ? nullptr : &b->nodes[0], uint32_t(b->nodes.size()),
                                        b->tris.data(), uint32_t(b->tris.size()));
        for (int a = 0; a < 3; a++) {
            blas_box[8 * size_t(k) + a] = ob.lo[a];
            blas_box[8 * size_t(k) + 4 + a] = ob.hi[a];
        }
    }
    const size_t nn = bo.nodes.size();
    std::vector<int32_t> leaves(4 * nn, 0);
    std::vector<uint32_t> depth(nn, 0);
    uint32_t deepest = 0;
    for (size_t g = 0; g < nn; g++) {
        const int32_t ptr[2] = {bo.nodes[g].left_ptr, bo.nodes[g].right_ptr};
        for (int side = 0; side < 2; side++) {
            if (ptr[side] >= 0) {
                depth[size_t(ptr[side])] = depth[g] + 1;
                deepest = std::max(deepest, depth[g] + 1);
            } else if (side == 1 && ptr[1] == ptr[0]) {
                // a single instance: the builder's root node points both
                // children at it and leaves the second box zero (count -1: kept)
                leaves[4 * g + 3] = -1;
            } else {
                uint32_t first = uint32_t(~ptr[side]), last = first;
                while (!bo.end_of_node[last]) last++;
                leaves[4 * g + 2 * side] = int32_t(first);
                leaves[4 * g + 2 * side + 1] = int32_t(last - first + 1);
            }
        }
    }
    std::vector<uint32_t> level_off(deepest + 2, 0), by_level(nn);
    for (size_t g = 0; g < nn; g++) level_off[deepest - depth[g] + 1]++;
    for (size_t l = 1; l < level_off.size(); l++) level_off[l] += level_off[l - 1];
    {
        std::vector<uint32_t> next(level_off.begin(), level_off.end() - 1);
        for (size_t g = 0; g < nn; g++) by_level[next[deepest - depth[g]]++] = uint32_t(g);
    }
    if (st == ATLAS_RT_OK) st = upload_vec(w->d_blas_box, blas_box);
    if (st == ATLAS_RT_OK) st = upload_vec(w->d_refit_leaves, leaves);
    if (st == ATLAS_RT_OK) st = upload_vec(w->d_refit_nodes, by_level);
    if (st == ATLAS_RT_OK) st = upload_vec(w->d_inst_box, std::vector<float>(8 * m, 0.0f));
    if (st != ATLAS_RT_OK) {
        w->free_device();
        return st;
    }
    w->refit_level_off = std::move(level_off);
    w->tlas = std::move(bo.nodes);
    w->instances = std::move(packed);
    w->order = std::move(order);
    w->table = std::move(table);
    w->blas_handles.assign(blases, blases + blas_count);
    w->xform = std::move(xform);
    w->tlas_depth = bo.max_depth;
    w->max_depth = bo.max_depth + blas_depth;
    w->device = dev;
    w->valid = true;
    return ATLAS_RT_OK;
}

atlas_rt_status atlas_rt_world_refit(atlas_rt_world* w, const atlas_rt_instance_desc* d_instances, void* stream) {
    atlas_rt_status st = check_world(w, "atlas_rt_world_refit");
    if (st != ATLAS_RT_OK) return st;
    if (!d_instances) return fail(ATLAS_RT_E_INVALID_ARG, "atlas_rt_world_refit: NULL instances");
    return launch_world_refit(*w, d_instances, reinterpret_cast<hipStream_t>(stream));
}

atlas_rt_status atlas_rt_world_set_stack_limit(atlas_rt_world* w, uint32_t entries) {
    if (!w) return fail(ATLAS_RT_E_INVALID_ARG, "atlas_rt_world_set_stack_limit: NULL world");
    w->stack_limit = entries;
    return ATLAS_RT_OK;
}

atlas_rt_status atlas_rt_world_set_history(atlas_rt_world* w, const float* last_object_to_world, uint32_t n) {
    if (!w) return fail(ATLAS_RT_E_INVALID_ARG, "atlas_rt_world_set_history: NULL world");
    if (!w->valid) return fail(ATLAS_RT_E_INVALID_ARG, "atlas_rt_world_set_history: the world has no instances");
    if (!last_object_to_world) {
        w->history.clear();
        w->history_gen = ~uint64_t(0);
        w->history_serial++;
        return ATLAS_RT_OK;
    }
    if (n != w->order.size())
        return fail(ATLAS_RT_E_SIZE_MISMATCH, "atlas_rt_world_set_history: " + std::to_string(n) +
                                                  " matrices for " + std::to_string(w->order.size()) + " instances");
    w->history.assign(last_object_to_world, last_object_to_world + 12 * size_t(n));
    w->history_gen = w->generation;
    w->history_serial++;
    return ATLAS_RT_OK;
}

atlas_rt_status atlas_rt_world_trace_status(atlas_rt_world* w, void* stream) {
    if (!w) return fail(ATLAS_RT_E_INVALID_ARG, "atlas_rt_world_trace_status: NULL world");
    return take_overflow_flag(w->flags.ptr, stream, "atlas_rt_world_trace_status");
}

atlas_rt_status atlas_rt_world_info(const atlas_rt_world* w, uint32_t* tlas_node_count, uint32_t* instance_count,
                                    uint32_t* max_depth) {
    if (!w) return fail(ATLAS_RT_E_INVALID_ARG, "atlas_rt_world_info: NULL world");
    if (tlas_node_count) *tlas_node_count = uint32_t(w->tlas.size());
    if (instance_count) *instance_count = uint32_t(w->instances.size());
    if (max_depth) *max_depth = w->max_depth;
    return ATLAS_RT_OK;
}

atlas_rt_status atlas_rt_world_export(const atlas_rt_world* w, atlas_rt_gpu_bvh_node* tlas_nodes,
                                      atlas_rt_gpu_bvh_instance* instances, uint32_t* instance_order) {
    if (!w) return fail(ATLAS_RT_E_INVALID_ARG, "atlas_rt_world_export: NULL world");
    atlas_rt_status st = const_cast<atlas_rt_world*>(w)->sync_host();
    if (st != ATLAS_RT_OK) return st;
    if (tlas_nodes && !w->tlas.empty()) std::memcpy(tlas_nodes, w->tlas.data(), w->tlas.size() * sizeof(*tlas_nodes));
    if (instances && !w->instances.empty())
        std::memcpy(instances, w->instances.data(), w->instances.size() * sizeof(*instances));
    if (instance_order && !w->order.empty())
        std::memcpy(instance_order, w->order.data(), w->order.size() * sizeof(uint32_t));
    return ATLAS_RT_OK;
}

atlas_rt_status atlas_rt_trace_closest(const atlas_rt_world* w, const atlas_rt_packed_ray* d_in,
                                       atlas_rt_packed_ray* d_out, const uint32_t* d_count, uint32_t max_count,
                                       float tmin, float tmax, uint32_t mask, void* stream) {
    atlas_rt_status st = check_world(w, "atlas_rt_trace_closest");
    if (st != ATLAS_RT_OK) return st;
    if (max_count && (!d_in || !d_out)) return fail(ATLAS_RT_E_INVALID_ARG, "atlas_rt_trace_closest: NULL buffer");
    return launch_world_closest_packed(*const_cast<atlas_rt_world*>(w), d_in, d_out, d_count, max_count, tmin, tmax,
                                       mask, reinterpret_cast<hipStream_t>(stream));
}

atlas_rt_status atlas_rt_trace_closest_transparency(const atlas_rt_world* w, const atlas_rt_packed_ray* d_in,
                                                    atlas_rt_packed_ray* d_out, const uint32_t* d_count,
                                                    uint32_t max_count, float tmin, float tmax, uint32_t mask,
                                                    void* stream) {
    atlas_rt_status st = check_world(w, "atlas_rt_trace_closest_transparency");
    if (st != ATLAS_RT_OK) return st;
    if (max_count && (!d_in || !d_out))
        return fail(ATLAS_RT_E_INVALID_ARG, "atlas_rt_trace_closest_transparency: NULL buffer");
    return launch_world_alpha_packed(*const_cast<atlas_rt_world*>(w), d_in, d_out, nullptr, d_count, max_count, tmin,
                                     tmax, mask, reinterpret_cast<hipStream_t>(stream));
}

atlas_rt_status atlas_rt_trace_transparency(const atlas_rt_world* w, const atlas_rt_packed_ray* d_in,
                                            float* d_transparency, const uint32_t* d_count, uint32_t max_count,
                                            float tmin, float tmax, uint32_t mask, void* stream) {
    atlas_rt_status st = check_world(w, "atlas_rt_trace_transparency");
    if (st != ATLAS_RT_OK) return st;
    if (max_count && (!d_in || !d_transparency))
        return fail(ATLAS_RT_E_INVALID_ARG, "atlas_rt_trace_transparency: NULL buffer");
    return launch_world_alpha_packed(*const_cast<atlas_rt_world*>(w), d_in, nullptr, d_transparency, d_count,
                                     max_count, tmin, tmax, mask, reinterpret_cast<hipStream_t>(stream));
}

atlas_rt_status atlas_rt_trace_any(const atlas_rt_world* w, const atlas_rt_packed_ray* d_in, uint32_t* d_occluded,
                                   const uint32_t* d_count, uint32_t max_count, float tmin, float tmax, uint32_t mask,
                                   uint32_t flags, void* stream) {
    atlas_rt_status st = check_world(w, "atlas_rt_trace_any");
    if (st != ATLAS_RT_OK) return st;
    if (max_count && (!d_in || !d_occluded)) return fail(ATLAS_RT_E_INVALID_ARG, "atlas_rt_trace_any: NULL buffer");
    if (flags & ~ATLAS_RT_ANY_RESTORE_ALWAYS) return fail(ATLAS_RT_E_INVALID_ARG, "atlas_rt_trace_any: unknown flags");
    return launch_world_any_packed(*const_cast<atlas_rt_world*>(w), d_in, d_occluded, d_count, max_count, tmin, tmax,
                                   mask, flags, reinterpret_cast<hipStream_t>(stream));
}

atlas_rt_status atlas_rt_world_trace_closest_rays(const atlas_rt_world* w, const atlas_rt_ray* d_rays,
                                                  atlas_rt_hit* d_hits, int32_t* d_instance, uint32_t count,
                                                  uint32_t mask, void* stream) {
    atlas_rt_status st = check_world(w, "atlas_rt_world_trace_closest_rays");
    if (st != ATLAS_RT_OK) return st;
    if (count && (!d_rays || !d_hits))
        return fail(ATLAS_RT_E_INVALID_ARG, "atlas_rt_world_trace_closest_rays: NULL buffer");
    return launch_world_closest_rays(*const_cast<atlas_rt_world*>(w), d_rays, d_hits, d_instance, count, mask,
                                     reinterpret_cast<hipStream_t>(stream));
}

atlas_rt_status atlas_rt_world_trace_any_rays(const atlas_rt_world* w, const atlas_rt_ray* d_rays, uint32_t* d_hit,
                                              uint32_t count, uint32_t mask, uint32_t flags, void* stream) {
    atlas_rt_status st = check_world(w, "atlas_rt_world_trace_any_rays");
    if (st != ATLAS_RT_OK) return st;
    if (count && (!d_rays || !d_hit)) return fail(ATLAS_RT_E_INVALID_ARG, "atlas_rt_world_trace_any_rays: NULL buffer");
    if (flags & ~ATLAS_RT_ANY_RESTORE_ALWAYS)
        return fail(ATLAS_RT_E_INVALID_ARG, "atlas_rt_world_trace_any_rays: unknown flags");
    return launch_world_any_rays(*const_cast<atlas_rt_world*>(w), d_rays, d_hit, count, mask, flags,
                                 reinterpret_cast<hipStream_t>(stream));
}

}  // extern "C"
