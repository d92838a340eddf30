// Batch sets: a BLAS table plus (BLAS, ray range) descriptors traced by one
// launch, with autotuned launch configuration and optional graph replay.
#pragma once

#include <map>
#include <string>
#include <vector>

#include "rt_internal.h"

namespace atlas_rt {

struct BatchDev {
    const BlasDev* blases;
    const uint4* batches;   // (blas, ray_offset, ray_count, 0)
    const uint32_t* prefix; // exclusive prefix of ray_count, batch_count + 1 entries
    uint32_t batch_count;
    uint32_t total;
};

struct BatchLaunch {
    const void* rays = nullptr;
    void* hits = nullptr;
    const uint32_t* d_count = nullptr;
    bool persistent = true;
    int node_w = 2;       // 2, 4 or 8 children per node
    bool reg = false;     // register-topped traversal stack
    bool fslab = false;   // widened float slab test (binary nodes)
    bool leaf_batch = false;  // leaf-batched resident waves (binary nodes)
    uint32_t block = 256;
    uint32_t refill = 40;
    hipStream_t stream = nullptr;
};

}  // namespace atlas_rt

struct atlas_rt_batch_set {
    std::vector<atlas_rt::BlasDev> table;
    std::vector<atlas_rt_batch> batches;
    std::vector<uint32_t> prefix;
    uint32_t total = 0;
    uint32_t max_depth = 0;
    uint32_t max_cap4 = 1, max_cap8 = 1;  // wide stack bounds over the set's BLASes
    int device = -1;
    void* d_table = nullptr;
    void* d_batches = nullptr;
    void* d_prefix = nullptr;
    atlas_rt::Scratch spill;
    atlas_rt::Scratch flags;
    atlas_rt::LaunchTiming timing;
    // Tuned launch (valid once `tuned`).
    bool tuned = false;
    atlas_rt_launch_config best{};
    std::vector<std::pair<atlas_rt_launch_config, float>> trials;
    // Captured graph.
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;

    atlas_rt::BatchDev dev() const {
        atlas_rt::BatchDev d;
        d.blases = reinterpret_cast<const atlas_rt::BlasDev*>(d_table);
        d.batches = reinterpret_cast<const uint4*>(d_batches);
        d.prefix = reinterpret_cast<const uint32_t*>(d_prefix);
        d.batch_count = uint32_t(batches.size());
        d.total = total;
        return d;
    }
};

namespace atlas_rt {
atlas_rt_status launch_batch(atlas_rt_batch_set& s, const BatchLaunch& L);
}
