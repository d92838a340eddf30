// Batch sets (C ABI), the launch autotuner and graph capture, plus the
// many-mesh BLAS builder.
#include "batch.h"

#include <algorithm>
#include <atomic>
#include <cstring>
#include <thread>
#include <tuple>

namespace atlas_rt {
int default_threads();
}

using namespace atlas_rt;

namespace {

template <class T>
atlas_rt_status upload(void*& dst, const std::vector<T>& v) {
    if (dst) (void)hipFree(dst);
    dst = nullptr;
    if (v.empty()) return ATLAS_RT_OK;
    ATLAS_HIP_TRY(hipMalloc(&dst, v.size() * sizeof(T)));
    ATLAS_HIP_TRY(hipMemcpy(dst, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
    return ATLAS_RT_OK;
}

BatchLaunch to_launch(const atlas_rt_launch_config& c) {
    BatchLaunch L;
    const int base = c.kernel & ATLAS_RT_KERNEL_VARIANT_MASK;
    L.persistent = base != ATLAS_RT_KERNEL_BVH2_EXACT;
    L.node_w = base == ATLAS_RT_KERNEL_BVH4 ? 4 : (base == ATLAS_RT_KERNEL_BVH8 ? 8 : 2);
    L.reg = (c.kernel & ATLAS_RT_KERNEL_STACK_REGISTER) != 0;
    L.fslab = (c.kernel & ATLAS_RT_KERNEL_SLAB_FLOAT) != 0;
    L.leaf_batch = (c.kernel & ATLAS_RT_KERNEL_LEAF_BATCH) != 0;
    L.block = c.block ? c.block : 256;
    L.refill = c.refill ? c.refill : 40;
    return L;
}

bool valid_kernel(int32_t k) {
    const int32_t base = k & ATLAS_RT_KERNEL_VARIANT_MASK;
    const bool fslab_ok = !(k & ATLAS_RT_KERNEL_SLAB_FLOAT) || base == ATLAS_RT_KERNEL_BVH2_EXACT ||
                          base == ATLAS_RT_KERNEL_BVH2_PERSISTENT;
    const bool batch_ok = !(k & ATLAS_RT_KERNEL_LEAF_BATCH) || base == ATLAS_RT_KERNEL_BVH2_PERSISTENT;
    return (k & ~int32_t(ATLAS_RT_KERNEL_VARIANT_MASK | ATLAS_RT_KERNEL_STACK_REGISTER | ATLAS_RT_KERNEL_SLAB_FLOAT |
                         ATLAS_RT_KERNEL_LEAF_BATCH)) == 0 &&
           fslab_ok && batch_ok && base >= ATLAS_RT_KERNEL_BVH2_EXACT && base <= ATLAS_RT_KERNEL_BVH8;
}

// Candidates: node width {2, 4, 8} x stack placement {LDS, register top} x
// work-group size {64, 128, 256}, for binary nodes both launch shapes (one
// ray per lane; resident waves at refill thresholds 16 / 32 / 48), both slab
// tests (engine-exact; widened float) and leaf-batched resident waves.
std::vector<atlas_rt_launch_config> candidates() {
    std::vector<atlas_rt_launch_config> out;
    const int32_t reg = ATLAS_RT_KERNEL_STACK_REGISTER;
    const int32_t fs = ATLAS_RT_KERNEL_SLAB_FLOAT;
    const int32_t lb = ATLAS_RT_KERNEL_LEAF_BATCH;
    for (uint32_t block : {64u, 128u, 256u}) {
        out.push_back({ATLAS_RT_KERNEL_BVH2_EXACT, block, 0, 0});
        out.push_back({ATLAS_RT_KERNEL_BVH2_EXACT | reg, block, 0, 0});
        out.push_back({ATLAS_RT_KERNEL_BVH2_EXACT | reg | fs, block, 0, 0});
        for (uint32_t refill : {16u, 32u, 48u}) out.push_back({ATLAS_RT_KERNEL_BVH2_PERSISTENT, block, refill, 0});
        out.push_back({ATLAS_RT_KERNEL_BVH2_PERSISTENT | reg, block, 32, 0});
        out.push_back({ATLAS_RT_KERNEL_BVH2_PERSISTENT | reg | fs, block, 32, 0});
        out.push_back({ATLAS_RT_KERNEL_BVH2_PERSISTENT | lb, block, 32, 0});
        out.push_back({ATLAS_RT_KERNEL_BVH2_PERSISTENT | reg | lb, block, 32, 0});
        out.push_back({ATLAS_RT_KERNEL_BVH2_PERSISTENT | fs | lb, block, 32, 0});
        for (int32_t wide : {ATLAS_RT_KERNEL_BVH4, ATLAS_RT_KERNEL_BVH8}) {
            out.push_back({wide, block, 32, 0});
            out.push_back({wide | reg, block, 32, 0});
        }
    }
    return out;
}

atlas_rt_status timed(atlas_rt_batch_set& s, BatchLaunch L, int reps, float& ms) {
    hipEvent_t a, b;
    ATLAS_HIP_TRY(hipEventCreate(&a));
    ATLAS_HIP_TRY(hipEventCreate(&b));
    atlas_rt_status st = launch_batch(s, L);  // warm-up
    if (st == ATLAS_RT_OK && hipEventRecord(a, L.stream) != hipSuccess) st = ATLAS_RT_E_HIP;
    for (int i = 0; st == ATLAS_RT_OK && i < reps; i++) st = launch_batch(s, L);
    if (st == ATLAS_RT_OK && hipEventRecord(b, L.stream) != hipSuccess) st = ATLAS_RT_E_HIP;
    if (st == ATLAS_RT_OK && hipEventSynchronize(b) != hipSuccess) st = ATLAS_RT_E_HIP;
    if (st == ATLAS_RT_OK) {
        (void)hipEventElapsedTime(&ms, a, b);
        ms /= float(reps);
    }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    return st;
}

// Tuned configurations shared by sets of the same shape on the same device.
struct ShapeKey {
    int device;
    uint32_t batches, log2_rays, depth;
    bool operator<(const ShapeKey& o) const {
        return std::tie(device, batches, log2_rays, depth) < std::tie(o.device, o.batches, o.log2_rays, o.depth);
    }
};
std::map<ShapeKey, atlas_rt_launch_config>& tuned_cache() {
    static std::map<ShapeKey, atlas_rt_launch_config> m;
    return m;
}

ShapeKey key_of(const atlas_rt_batch_set& s) {
    uint32_t lg = 0;
    while ((1u << lg) < s.total && lg < 31) lg++;
    return ShapeKey{s.device, uint32_t(s.batches.size()), lg, s.max_depth};
}

atlas_rt_status autotune(atlas_rt_batch_set& s, const BatchLaunch& base) {
    auto it = tuned_cache().find(key_of(s));
    if (it != tuned_cache().end()) {
        s.best = it->second;
        s.tuned = true;
        return ATLAS_RT_OK;
    }
    s.trials.clear();
    float best_ms = 1e30f;
    for (const auto& c : candidates()) {
        BatchLaunch L = to_launch(c);
        L.rays = base.rays;
        L.hits = base.hits;
        L.d_count = base.d_count;
        L.stream = base.stream;
        float ms = 0;
        atlas_rt_status st = timed(s, L, 3, ms);
        if (st != ATLAS_RT_OK) return st;
        s.trials.push_back({c, ms});
        if (ms < best_ms) {
            best_ms = ms;
            s.best = c;
        }
    }
    s.tuned = true;
    tuned_cache()[key_of(s)] = s.best;
    return ATLAS_RT_OK;
}

}  // namespace

extern "C" {

atlas_rt_status atlas_rt_batch_create(const atlas_rt_blas* const* blases, uint32_t blas_count,
                                      const atlas_rt_batch* batches, uint32_t batch_count, atlas_rt_batch_set** out) {
    if (!out) return fail(ATLAS_RT_E_INVALID_ARG, "atlas_rt_batch_create: out is NULL");
    *out = nullptr;
    if (!blases || !batches || !blas_count || !batch_count)
        return fail(ATLAS_RT_E_INVALID_ARG, "atlas_rt_batch_create: empty BLAS or batch table");
    int dev = 0;
    ATLAS_HIP_TRY(hipGetDevice(&dev));
    auto* s = new atlas_rt_batch_set();
    s->table.resize(blas_count);
    for (uint32_t k = 0; k < blas_count; k++) {
        const atlas_rt_blas* b = blases[k];
        if (!b || !b->on_device || b->device != dev) {
            delete s;
            return fail(ATLAS_RT_E_INVALID_ARG, "atlas_rt_batch_create: BLAS " + std::to_string(k) +
                                                    " is NULL, host-only or on another device");
        }
        s->table[k] = BlasDev{b->d_nodes,  b->d_tris, b->d_vtx, b->d_nodes4, b->d_nodes8, uint32_t(b->nodes.size()),
                              uint32_t(b->tris.size())};
    }
    s->batches.assign(batches, batches + batch_count);
    s->prefix.resize(batch_count + 1);
    uint64_t total = 0;
    std::vector<uint32_t> used(blas_count, 0);
    for (uint32_t i = 0; i < batch_count; i++) {
        if (batches[i].blas >= blas_count) {
            delete s;
            return fail(ATLAS_RT_E_INVALID_ARG, "atlas_rt_batch_create: batch names a BLAS out of range");
        }
        s->prefix[i] = uint32_t(total);
        total += batches[i].ray_count;
        used[batches[i].blas] = 1;
        s->max_depth = std::max(s->max_depth, blases[batches[i].blas]->max_depth);
        s->max_cap4 = std::max(s->max_cap4, blases[batches[i].blas]->cap4);
        s->max_cap8 = std::max(s->max_cap8, blases[batches[i].blas]->cap8);
    }
    if (total >= (1ull << 32)) {
        delete s;
        return fail(ATLAS_RT_E_INVALID_ARG, "atlas_rt_batch_create: more than 2^32 rays");
    }
    s->prefix[batch_count] = uint32_t(total);
    s->total = uint32_t(total);
    s->device = dev;
    atlas_rt_status st = upload(s->d_table, s->table);
    if (st == ATLAS_RT_OK) st = upload(s->d_batches, s->batches);
    if (st == ATLAS_RT_OK) st = upload(s->d_prefix, s->prefix);
    if (st == ATLAS_RT_OK) st = s->flags.ensure(kFlagBytes);
    if (st == ATLAS_RT_OK && hipMemset(s->flags.ptr, 0, kFlagBytes) != hipSuccess) st = fail(ATLAS_RT_E_HIP, "hipMemset");
    if (st != ATLAS_RT_OK) {
        atlas_rt_batch_destroy(s);
        return st;
    }
    *out = s;
    return ATLAS_RT_OK;
}

atlas_rt_status atlas_rt_batch_destroy(atlas_rt_batch_set* s) {
    if (!s) return ATLAS_RT_OK;
    if (s->graph_exec) (void)hipGraphExecDestroy(s->graph_exec);
    if (s->graph) (void)hipGraphDestroy(s->graph);
    for (void* p : {s->d_table, s->d_batches, s->d_prefix})
        if (p) (void)hipFree(p);
    s->spill.release();
    s->flags.release();
    s->timing.release();
    delete s;
    return ATLAS_RT_OK;
}

atlas_rt_status atlas_rt_batch_trace_closest(atlas_rt_batch_set* s, const atlas_rt_ray* d_rays, atlas_rt_hit* d_hits,
                                             const uint32_t* d_count, const atlas_rt_launch_config* config,
                                             void* stream) {
    if (!s) return fail(ATLAS_RT_E_INVALID_ARG, "atlas_rt_batch_trace_closest: NULL set");
    if (s->total && (!d_rays || !d_hits)) return fail(ATLAS_RT_E_INVALID_ARG, "atlas_rt_batch_trace_closest: NULL buffer");
    BatchLaunch base;
    base.rays = d_rays;
    base.hits = d_hits;
    base.d_count = d_count;
    base.stream = reinterpret_cast<hipStream_t>(stream);
    atlas_rt_launch_config c{};
    if (config && config->kernel != ATLAS_RT_KERNEL_AUTO) {
        if (!valid_kernel(config->kernel))
            return fail(ATLAS_RT_E_UNSUPPORTED, "atlas_rt_batch_trace_closest: kernel variant not available");
        c = *config;
    } else {
        if (!s->tuned) {
            atlas_rt_status st = autotune(*s, base);
            if (st != ATLAS_RT_OK) return st;
        }
        c = s->best;
    }
    BatchLaunch L = to_launch(c);
    L.rays = base.rays;
    L.hits = base.hits;
    L.d_count = base.d_count;
    L.stream = base.stream;
    atlas_rt_status st = s->timing.begin(L.stream);
    if (st != ATLAS_RT_OK) return st;
    st = launch_batch(*s, L);
    if (st != ATLAS_RT_OK) return st;
    return s->timing.end(L.stream);
}

atlas_rt_status atlas_rt_batch_tuning(const atlas_rt_batch_set* s, atlas_rt_launch_config* best, uint32_t* n_trials,
                                      atlas_rt_launch_config* trials, float* trial_ms) {
    if (!s) return fail(ATLAS_RT_E_INVALID_ARG, "atlas_rt_batch_tuning: NULL set");
    if (best) *best = s->best;
    if (n_trials) {
        const uint32_t cap = *n_trials;
        *n_trials = uint32_t(s->trials.size());
        for (uint32_t i = 0; i < std::min(cap, uint32_t(s->trials.size())); i++) {
            if (trials) trials[i] = s->trials[i].first;
            if (trial_ms) trial_ms[i] = s->trials[i].second;
        }
    }
    return ATLAS_RT_OK;
}

atlas_rt_status atlas_rt_batch_last_kernel_ms(atlas_rt_batch_set* s, float* ms) {
    if (!s || !ms) return fail(ATLAS_RT_E_INVALID_ARG, "atlas_rt_batch_last_kernel_ms: NULL argument");
    return s->timing.read(ms);
}

atlas_rt_status atlas_rt_batch_graph_capture(atlas_rt_batch_set* s, const atlas_rt_ray* d_rays, atlas_rt_hit* d_hits,
                                             const uint32_t* d_count, const atlas_rt_launch_config* config) {
    if (!s) return fail(ATLAS_RT_E_INVALID_ARG, "atlas_rt_batch_graph_capture: NULL set");
    hipStream_t cs;
    ATLAS_HIP_TRY(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
    atlas_rt_launch_config c{};
    if (config && config->kernel != ATLAS_RT_KERNEL_AUTO) {
        if (!valid_kernel(config->kernel)) {
            (void)hipStreamDestroy(cs);
            return fail(ATLAS_RT_E_UNSUPPORTED, "atlas_rt_batch_graph_capture: kernel variant not available");
        }
        c = *config;
    } else {
        if (!s->tuned) {
            BatchLaunch base;
            base.rays = d_rays;
            base.hits = d_hits;
            base.d_count = d_count;
            base.stream = cs;
            atlas_rt_status st = autotune(*s, base);
            if (st != ATLAS_RT_OK) {
                (void)hipStreamDestroy(cs);
                return st;
            }
        }
        c = s->best;
    }
    // Scratch must exist before capture (no allocation inside the graph).
    BatchLaunch L = to_launch(c);
    L.rays = d_rays;
    L.hits = d_hits;
    L.d_count = d_count;
    L.stream = cs;
    atlas_rt_status st = launch_batch(*s, L);
    if (st == ATLAS_RT_OK && hipStreamSynchronize(cs) != hipSuccess) st = fail(ATLAS_RT_E_HIP, "hipStreamSynchronize");
    if (st != ATLAS_RT_OK) {
        (void)hipStreamDestroy(cs);
        return st;
    }
    if (s->graph_exec) (void)hipGraphExecDestroy(s->graph_exec);
    if (s->graph) (void)hipGraphDestroy(s->graph);
    s->graph_exec = nullptr;
    s->graph = nullptr;
    hipError_t e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
    if (e == hipSuccess) {
        st = launch_batch(*s, L);
        hipError_t e2 = hipStreamEndCapture(cs, &s->graph);
        if (st == ATLAS_RT_OK && e2 != hipSuccess) st = hip_fail(e2, "hipStreamEndCapture");
    } else {
        st = hip_fail(e, "hipStreamBeginCapture");
    }
    if (st == ATLAS_RT_OK) {
        e = hipGraphInstantiate(&s->graph_exec, s->graph, nullptr, nullptr, 0);
        if (e != hipSuccess) st = hip_fail(e, "hipGraphInstantiate");
    }
    (void)hipStreamDestroy(cs);
    return st;
}

atlas_rt_status atlas_rt_batch_graph_launch(atlas_rt_batch_set* s, void* stream) {
    if (!s || !s->graph_exec) return fail(ATLAS_RT_E_INVALID_ARG, "atlas_rt_batch_graph_launch: no captured graph");
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    atlas_rt_status r = s->timing.begin(st);
    if (r != ATLAS_RT_OK) return r;
    ATLAS_HIP_TRY(hipGraphLaunch(s->graph_exec, st));
    return s->timing.end(st);
}

atlas_rt_status atlas_rt_blas_build_many(const atlas_rt_bvh_triangle* const* tris, const uint32_t* counts,
                                         uint32_t count, uint32_t flags, atlas_rt_blas** out) {
    if (!tris || !counts || !out) return fail(ATLAS_RT_E_INVALID_ARG, "atlas_rt_blas_build_many: NULL argument");
    for (uint32_t k = 0; k < count; k++) out[k] = nullptr;
    // Host builds in parallel (one mesh per thread, each serial); device
    // uploads afterwards on this thread.
    std::vector<atlas_rt_blas*> built(count, nullptr);
    std::vector<atlas_rt_status> status(count, ATLAS_RT_OK);
    std::atomic<uint32_t> next{0};
    const int threads = std::max(1, std::min<int>(default_threads(), int(count)));
    const uint32_t host_flags = (flags & ~ATLAS_RT_BUILD_PARALLEL) | ATLAS_RT_BUILD_HOST_ONLY;
    std::vector<std::thread> pool;
    std::vector<std::string> errors(count);
    for (int t = 0; t < threads; t++)
        pool.emplace_back([&] {
            for (uint32_t k; (k = next.fetch_add(1)) < count;) {
                status[k] = atlas_rt_blas_build(nullptr, tris[k], counts[k], host_flags, &built[k]);
                if (status[k] != ATLAS_RT_OK) errors[k] = atlas_rt_last_error();
            }
        });
    for (auto& th : pool) th.join();
    atlas_rt_status st = ATLAS_RT_OK;
    for (uint32_t k = 0; k < count && st == ATLAS_RT_OK; k++) {
        if (status[k] != ATLAS_RT_OK) {
            st = fail(status[k], "atlas_rt_blas_build_many: mesh " + std::to_string(k) + ": " + errors[k]);
            break;
        }
        if (!(flags & ATLAS_RT_BUILD_HOST_ONLY)) st = upload_blas(*built[k]);
    }
    if (st != ATLAS_RT_OK) {
        for (uint32_t k = 0; k < count; k++) atlas_rt_blas_destroy(built[k]);
        return st;
    }
    for (uint32_t k = 0; k < count; k++) out[k] = built[k];
    return ATLAS_RT_OK;
}

}  // extern "C"
