// BLAS objects built on the GPU (lbvh.hip): device arrays written by the
// builder, host mirrors copied back for export, world bounds, shading and
// the wide-node collapse.
//
// Memory: every batch of up to 1024 meshes shares two device allocations,
// one for the builder's outputs (nodes, triangle records, vertex records)
// and one for the wide nodes; each BLAS of the batch holds a reference and
// the last one destroyed frees them.  That keeps a build of many small
// meshes at two hipMalloc calls per batch.
#include <algorithm>
#include <atomic>
#include <cstring>
#include <thread>

#include "lbvh.h"
#include "rt_internal.h"
#include "wide.h"

using namespace atlas_rt;

namespace atlas_rt {
int default_threads();
}

namespace {

constexpr size_t kAlign = 256;

size_t align_up(size_t x) { return (x + kAlign - 1) / kAlign * kAlign; }

struct DeviceArena {
    void* build = nullptr;
    void* wide = nullptr;
    ~DeviceArena() {
        if (build) (void)hipFree(build);
        if (wide) (void)hipFree(wide);
    }
};

void free_partial(std::vector<atlas_rt_blas*>& made) {
    for (atlas_rt_blas* b : made) atlas_rt_blas_destroy(b);
    made.clear();
}

// Runs fn(m) for m in [0, n) on up to default_threads() host threads.
template <class F>
void parallel_for(uint32_t n, F&& fn) {
    const int threads = std::max(1, std::min<int>(default_threads(), int(n)));
    std::atomic<uint32_t> next{0};
    std::vector<std::thread> pool;
    for (int t = 1; t < threads; t++)
        pool.emplace_back([&] {
            for (uint32_t m; (m = next.fetch_add(1)) < n;) fn(m);
        });
    for (uint32_t m; (m = next.fetch_add(1)) < n;) fn(m);
    for (auto& th : pool) th.join();
}

// One batch (at most 1024 meshes) into made[first .. first + n).
atlas_rt_status build_batch(const float* const* verts, const uint32_t* counts, uint32_t n, int dev, hipStream_t s,
                            std::vector<atlas_rt_blas*>& made, float* ms) {
    std::vector<const float*> v(verts, verts + n);
    std::vector<uint32_t> c(counts, counts + n);
    // build arena: [nodes of every mesh][vertex records][triangle records]
    std::vector<size_t> off_n(n), off_v(n), off_t(n);
    size_t at = 0;
    for (uint32_t m = 0; m < n; m++) {
        off_n[m] = at;
        at = align_up(at + size_t(c[m] - 1) * sizeof(atlas_rt_gpu_bvh_node));
    }
    for (uint32_t m = 0; m < n; m++) {
        off_v[m] = at;
        at = align_up(at + size_t(c[m]) * sizeof(atlas_rt_gpu_bvh_triangle));
    }
    const size_t tris_begin = at;
    for (uint32_t m = 0; m < n; m++) {
        off_t[m] = at;
        at = align_up(at + size_t(c[m]) * 3 * sizeof(float4));
    }
    size_t total = 0;
    std::vector<size_t> off_o(n);
    for (uint32_t m = 0; m < n; m++) {
        off_o[m] = total;
        total += c[m];
    }
    auto arena = std::make_shared<DeviceArena>();
    hipError_t e = hipMalloc(&arena->build, std::max<size_t>(at, 1));
    uint32_t* orig = nullptr;
    if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void**>(&orig), total * sizeof(uint32_t));
    if (e != hipSuccess) {
        if (orig) (void)hipFree(orig);
        return hip_fail(e, "atlas_rt_blas_build_device (outputs)");
    }
    char* base = static_cast<char*>(arena->build);
    std::vector<LbvhMeshOut> o(n);
    const size_t first = made.size();
    for (uint32_t m = 0; m < n; m++) {
        auto* b = new atlas_rt_blas();
        made.push_back(b);
        b->device = dev;
        b->on_device = true;
        b->arena = arena;
        b->d_nodes = c[m] > 1 ? reinterpret_cast<float4*>(base + off_n[m]) : nullptr;
        b->d_vtx = reinterpret_cast<float4*>(base + off_v[m]);
        b->d_tris = reinterpret_cast<float4*>(base + off_t[m]);
        o[m] = LbvhMeshOut{b->d_nodes, b->d_tris, b->d_vtx, orig + off_o[m]};
    }
    std::vector<uint32_t> depth;
    atlas_rt_status st = lbvh_build(v, c, o, depth, s, ms);
    // host mirrors: node and vertex records in two copies, split on the host
    std::vector<char> host;
    std::vector<uint32_t> host_orig(total);
    if (st == ATLAS_RT_OK) {
        host.resize(tris_begin);
        e = hipMemcpy(host.data(), base, tris_begin, hipMemcpyDeviceToHost);
        if (e == hipSuccess) e = hipMemcpy(host_orig.data(), orig, total * sizeof(uint32_t), hipMemcpyDeviceToHost);
        if (e != hipSuccess) st = hip_fail(e, "atlas_rt_blas_build_device (host mirror)");
    }
    (void)hipFree(orig);
    if (st != ATLAS_RT_OK) return st;
    // mirrors and the wide collapse, one mesh per host thread
    std::vector<WideTree> w4(n), w8(n);
    parallel_for(n, [&](uint32_t m) {
        atlas_rt_blas* b = made[first + m];
        const size_t nt = c[m], nn = nt - 1;
        b->nodes.resize(nn);
        b->tris.resize(nt);
        if (nn) std::memcpy(b->nodes.data(), host.data() + off_n[m], nn * sizeof(atlas_rt_gpu_bvh_node));
        std::memcpy(b->tris.data(), host.data() + off_v[m], nt * sizeof(atlas_rt_gpu_bvh_triangle));
        b->orig_idx.assign(host_orig.begin() + off_o[m], host_orig.begin() + off_o[m] + nt);
        b->max_depth = depth[m];
        std::vector<uint8_t> end_flags(nt);
        for (size_t i = 0; i < nt; i++) end_flags[i] = b->tris[i].v0[3] > 0.0f ? 1 : 0;
        w4[m] = collapse_wide(b->nodes, end_flags, 4, 4);
        w8[m] = collapse_wide(b->nodes, end_flags, 8, 4);
    });
    // wide arena: [4-wide nodes of every mesh][8-wide nodes]
    std::vector<size_t> off4(n), off8(n);
    size_t wat = 0;
    for (uint32_t m = 0; m < n; m++) {
        off4[m] = wat;
        wat = align_up(wat + w4[m].data.size() * sizeof(float));
    }
    for (uint32_t m = 0; m < n; m++) {
        off8[m] = wat;
        wat = align_up(wat + w8[m].data.size() * sizeof(float));
    }
    std::vector<char> wide_host(wat);
    parallel_for(n, [&](uint32_t m) {
        std::memcpy(wide_host.data() + off4[m], w4[m].data.data(), w4[m].data.size() * sizeof(float));
        std::memcpy(wide_host.data() + off8[m], w8[m].data.data(), w8[m].data.size() * sizeof(float));
    });
    e = hipMalloc(&arena->wide, std::max<size_t>(wat, 1));
    if (e == hipSuccess) e = hipMemcpy(arena->wide, wide_host.data(), wat, hipMemcpyHostToDevice);
    if (e != hipSuccess) return hip_fail(e, "atlas_rt_blas_build_device (wide nodes)");
    char* wbase = static_cast<char*>(arena->wide);
    for (uint32_t m = 0; m < n; m++) {
        atlas_rt_blas* b = made[first + m];
        b->d_nodes4 = w4[m].node_count ? reinterpret_cast<float4*>(wbase + off4[m]) : nullptr;
        b->d_nodes8 = w8[m].node_count ? reinterpret_cast<float4*>(wbase + off8[m]) : nullptr;
        b->cap4 = w4[m].max_stack + 1;
        b->cap8 = w8[m].max_stack + 1;
    }
    return ATLAS_RT_OK;
}

}  // namespace

extern "C" {

atlas_rt_status atlas_rt_blas_build_device(const float* const* verts, const uint32_t* counts, uint32_t count,
                                           uint32_t flags, void* stream, atlas_rt_blas** out, float* device_ms) {
    if (!verts || !counts || !out) return fail(ATLAS_RT_E_INVALID_ARG, "atlas_rt_blas_build_device: NULL argument");
    if (count == 0) return fail(ATLAS_RT_E_EMPTY, "atlas_rt_blas_build_device: no meshes");
    if (flags) return fail(ATLAS_RT_E_INVALID_ARG, "atlas_rt_blas_build_device: unknown flag bits");
    for (uint32_t k = 0; k < count; k++) {
        out[k] = nullptr;
        if (!verts[k]) return fail(ATLAS_RT_E_INVALID_ARG, "atlas_rt_blas_build_device: NULL triangle array");
        if (counts[k] == 0) return fail(ATLAS_RT_E_EMPTY, "atlas_rt_blas_build_device: a mesh has no triangles");
        if (counts[k] > (1u << 24))
            return fail(ATLAS_RT_E_INVALID_ARG, "atlas_rt_blas_build_device: more than 2^24 triangles in a mesh");
    }
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    int dev = 0;
    ATLAS_HIP_TRY(hipGetDevice(&dev));
    std::vector<atlas_rt_blas*> made;
    float total_ms = 0.0f;
    for (uint32_t first = 0; first < count; first += 1024) {
        const uint32_t n = std::min<uint32_t>(1024, count - first);
        float ms = 0.0f;
        const atlas_rt_status st = build_batch(verts + first, counts + first, n, dev, s, made,
                                               device_ms ? &ms : nullptr);
        if (st != ATLAS_RT_OK) {
            free_partial(made);
            return st;
        }
        total_ms += ms;
    }
    for (uint32_t k = 0; k < count; k++) out[k] = made[k];
    if (device_ms) *device_ms = total_ms;
    return ATLAS_RT_OK;
}

}  // extern "C"
