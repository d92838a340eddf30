// Two-level world: TLAS nodes, packed instances and the BLAS table.
#pragma once

#include <vector>

#include "rt_internal.h"

namespace atlas_rt {

// Device view of a world.
struct WorldDev {
    const float4* tlas;       // engine GPUBVHNode, 4 float4 per node
    const float4* instances;  // GPUBVHInstance, 4 float4 per instance (3 matrix columns + meta)
    const BlasDev* blases;    // BLAS table indexed by GPUBVHInstance.meshOffset
    const int32_t* inst_order;  // packed instance slot -> caller instance index
};

atlas_rt_status launch_world_closest_rays(atlas_rt_world& w, const void* rays, void* hits, int32_t* inst,
                                          uint32_t count, uint32_t mask, hipStream_t stream);
atlas_rt_status launch_world_any_rays(atlas_rt_world& w, const void* rays, uint32_t* out, uint32_t count,
                                      uint32_t mask, uint32_t flags, hipStream_t stream);
atlas_rt_status launch_world_closest_packed(atlas_rt_world& w, const void* in, void* out, const uint32_t* d_count,
                                            uint32_t max_count, float tmin, float tmax, uint32_t mask,
                                            hipStream_t stream);
atlas_rt_status launch_world_alpha_packed(atlas_rt_world& w, const void* in, void* out_rays, float* out_trans,
                                          const uint32_t* d_count, uint32_t max_count, float tmin, float tmax,
                                          uint32_t mask, hipStream_t stream);
// Device TLAS refit with new instance matrices (refit.hip).
atlas_rt_status launch_world_refit(atlas_rt_world& w, const atlas_rt_instance_desc* d_desc, hipStream_t stream);

atlas_rt_status launch_world_any_packed(atlas_rt_world& w, const void* in, uint32_t* out, const uint32_t* d_count,
                                        uint32_t max_count, float tmin, float tmax, uint32_t mask, uint32_t flags,
                                        hipStream_t stream);

}  // namespace atlas_rt

struct atlas_rt_world {
    // Host copies (engine layout, TLAS leaf order).
    std::vector<atlas_rt_gpu_bvh_node> tlas;
    std::vector<atlas_rt_gpu_bvh_instance> instances;
    std::vector<uint32_t> order;  // packed slot -> caller instance index
    std::vector<atlas_rt::BlasDev> table;
    std::vector<const atlas_rt_blas*> blas_handles;  // borrowed; the caller keeps them alive
    std::vector<float> xform;                        // object-to-world rows, 12 floats per packed instance
    uint32_t tlas_depth = 0;
    uint32_t max_depth = 0;  // tlas depth + deepest BLAS
    bool valid = false;
    // Bumped by every atlas_rt_world_update, which replaces the device arrays
    // (a refit rewrites them in place): objects that keep device pointers of
    // the world rebind when it changes.
    uint64_t generation = 0;
    // Most traversal stack entries per ray (atlas_rt_world_set_stack_limit);
    // 0: sized from the TLAS and BLAS depths, which never overflows.
    uint32_t stack_limit = 0;
    uint32_t stack_cap() const { return atlas_rt::limit_cap(max_depth + 1, stack_limit); }
    // Last frame's object-to-world rows per instance in the caller's order of
    // the last update (atlas_rt_world_set_history); valid while
    // history_gen == generation.
    std::vector<float> history;
    uint64_t history_gen = ~uint64_t(0);
    uint64_t history_serial = 0;  // bumped by every atlas_rt_world_set_history

    int device = -1;
    void* d_tlas = nullptr;
    void* d_instances = nullptr;
    void* d_table = nullptr;
    void* d_order = nullptr;
    void* d_xform = nullptr;  // object-to-world rows, 3 float4 per packed instance
    // refit schedule, fixed by the last update: per BLAS its object box (2
    // float4), per slot a world box (2 float4), per TLAS node (int4) the
    // instance range of a leaf child (left first, count, right first,
    // count; count 0 = inner child), and the nodes grouped by depth,
    // deepest first
    void* d_blas_box = nullptr;
    void* d_inst_box = nullptr;
    void* d_refit_leaves = nullptr;
    void* d_refit_nodes = nullptr;
    std::vector<uint32_t> refit_level_off;
    bool host_stale = false;  // a refit changed the device copies after the host ones
    atlas_rt::Scratch spill;
    atlas_rt::Scratch flags;
    atlas_rt::LaunchTiming timing;
    // per-BLAS shading / uv pointers for the opacity-aware traces
    atlas_rt::Scratch alpha_tab;
    std::vector<const void*> alpha_tab_host;

    atlas_rt::WorldDev dev() const {
        atlas_rt::WorldDev d;
        d.tlas = reinterpret_cast<const float4*>(d_tlas);
        d.instances = reinterpret_cast<const float4*>(d_instances);
        d.blases = reinterpret_cast<const atlas_rt::BlasDev*>(d_table);
        d.inst_order = reinterpret_cast<const int32_t*>(d_order);
        return d;
    }
    void free_device();
    // Brings tlas, instances and xform up to date after a refit.
    atlas_rt_status sync_host();
};
