// BLAS objects built on the GPU (lbvh.hip): device arrays written by the
// builder, host mirrors copied back for export, world bounds, shading and
// the wide-node collapse.  The mirrors come back, and the wide nodes go up,
// through two pinned chunks while a team of host threads spreads each chunk
// over the per-BLAS vectors: a pageable copy of the whole arena runs at a
// few GB/s through the runtime's staging on one host thread.
//
// Memory: every batch of up to 1024 meshes shares two device allocations,
// one for the builder's outputs (nodes, triangle records, vertex records)
// and one for the wide nodes; each BLAS of the batch holds a reference and
// the last one destroyed frees them.  That keeps a build of many small
// meshes at two hipMalloc calls per batch.
#include <algorithm>
#include <atomic>
#include <condition_variable>
#include <cstring>
#include <exception>
#include <functional>
#include <mutex>
#include <new>
#include <string>
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

// A temporary device allocation of the build, freed on every way out
// (hipFree waits for the work that uses it).
struct DevTemp {
    void* p = nullptr;
    ~DevTemp() { reset(); }
    void reset() {
        if (p) (void)hipFree(p);
        p = nullptr;
    }
};

void free_partial(std::vector<atlas_rt_blas*>& made) {
    for (atlas_rt_blas* b : made) atlas_rt_blas_destroy(b);
    made.clear();
}

// A fixed team of host threads that runs one parallel loop after another:
// run(n, fn) calls fn(i) for every i in [0, n), the calling thread taking
// part, and returns when all calls have returned, rethrowing the first
// exception a call threw.  The team starts its threads once, so the many
// short loops of a staged copy cost a wake-up each.
class Team {
  public:
    explicit Team(int threads) {
        try {
            for (int t = 1; t < threads; t++) pool_.emplace_back([this] { work(); });
        } catch (...) {
            stop();  // the threads already started
            throw;
        }
    }
    ~Team() { stop(); }
    template <class F>
    void run(uint32_t n, F&& fn) {
        if (n == 0) return;
        if (pool_.empty() || n == 1) {
            for (uint32_t i = 0; i < n; i++) fn(i);
            return;
        }
        std::function<void(uint32_t)> job(std::forward<F>(fn));
        {
            std::lock_guard<std::mutex> g(mu_);
            job_ = &job;
            n_ = n;
            next_.store(0);
            busy_ = uint32_t(pool_.size());
            gen_++;
        }
        wake_.notify_all();
        drain(job, n);
        std::unique_lock<std::mutex> g(mu_);
        idle_.wait(g, [this] { return busy_ == 0; });
        job_ = nullptr;
        if (err_) {
            std::exception_ptr e = err_;
            err_ = nullptr;
            g.unlock();
            std::rethrow_exception(e);
        }
    }

  private:
    void stop() {
        {
            std::lock_guard<std::mutex> g(mu_);
            quit_ = true;
        }
        wake_.notify_all();
        for (auto& th : pool_) th.join();
        pool_.clear();
    }
    void work() {
        uint64_t seen = 0;
        for (;;) {
            std::function<void(uint32_t)>* job;
            uint32_t n;
            {
                std::unique_lock<std::mutex> g(mu_);
                wake_.wait(g, [&] { return quit_ || gen_ != seen; });
                if (quit_) return;
                seen = gen_;
                job = job_;
                n = n_;
            }
            drain(*job, n);
            std::lock_guard<std::mutex> g(mu_);
            if (--busy_ == 0) idle_.notify_one();
        }
    }
    // Takes indices until none is left; after a throw, takes the rest
    // without calling the job.
    void drain(std::function<void(uint32_t)>& job, uint32_t n) {
        for (uint32_t i; (i = next_.fetch_add(1)) < n;) {
            try {
                job(i);
            } catch (...) {
                std::lock_guard<std::mutex> g(mu_);
                if (!err_) err_ = std::current_exception();
                next_.store(n);
            }
        }
    }
    std::vector<std::thread> pool_;
    std::mutex mu_;
    std::condition_variable wake_, idle_;
    std::function<void(uint32_t)>* job_ = nullptr;
    uint32_t n_ = 0, busy_ = 0;
    uint64_t gen_ = 0;
    std::atomic<uint32_t> next_{0};
    std::exception_ptr err_;
    bool quit_ = false;
};

// A byte range [at, at + bytes) of a device region and the host memory it
// is copied from or to.  Pieces of one region are sorted by `at` and do not
// overlap; bytes between them are padding.
struct Piece {
    size_t at;
    size_t bytes;
    char* host;
};

// Two pinned host buffers for copies between a device region and many host
// pieces: the DMA of one chunk runs while the team copies the other chunk's
// pieces, so pageable copies (the runtime's own staging, one host thread)
// and a whole-region host buffer are both avoided.
class Staging {
  public:
    static constexpr size_t kChunk = size_t(16) << 20;
    static constexpr size_t kJob = size_t(1) << 20;  // bytes per host copy job

    explicit Staging(hipStream_t s) : s_(s) {}
    ~Staging() {
        if (buf_[0] || buf_[1]) (void)hipStreamSynchronize(s_);  // no DMA may still use the buffers
        for (int i = 0; i < 2; i++) {
            if (buf_[i]) (void)hipHostFree(buf_[i]);
            if (done_[i]) (void)hipEventDestroy(done_[i]);
        }
    }
    hipError_t init() {
        hipError_t e = hipSuccess;
        for (int i = 0; i < 2 && e == hipSuccess; i++) {
            e = hipHostMalloc(reinterpret_cast<void**>(&buf_[i]), kChunk, hipHostMallocDefault);
            if (e == hipSuccess) e = hipEventCreateWithFlags(&done_[i], hipEventDisableTiming);
        }
        return e;
    }

    // Device region [dev, dev + size) into the pieces' host memory.
    hipError_t down(Team& team, const char* dev, size_t size, const std::vector<Piece>& pieces) {
        const size_t chunks = (size + kChunk - 1) / kChunk;
        auto issue = [&](size_t k) {
            const size_t c0 = k * kChunk;
            hipError_t e = hipMemcpyAsync(buf_[k & 1], dev + c0, std::min(kChunk, size - c0), hipMemcpyDeviceToHost, s_);
            return e == hipSuccess ? hipEventRecord(done_[k & 1], s_) : e;
        };
        hipError_t e = chunks ? issue(0) : hipSuccess;
        size_t p = 0;
        for (size_t k = 0; k < chunks && e == hipSuccess; k++) {
            // buffer (k + 1) & 1 was emptied by the team in step k - 1
            if (k + 1 < chunks) e = issue(k + 1);
            if (e == hipSuccess) e = hipEventSynchronize(done_[k & 1]);
            if (e != hipSuccess) break;
            split(pieces, k, std::min(size, (k + 1) * kChunk), p, false);
            const char* b = buf_[k & 1];
            team.run(uint32_t(jobs_.size()), [&](uint32_t j) {
                const Job& J = jobs_[j];
                std::memcpy(J.host, b + J.off, J.bytes);
            });
        }
        return e;
    }

    // The pieces' host memory into device region [dev, dev + size); padding
    // between pieces is written as zeros.  Returns once the region is written.
    hipError_t up(Team& team, char* dev, size_t size, const std::vector<Piece>& pieces) {
        const size_t chunks = (size + kChunk - 1) / kChunk;
        hipError_t e = hipSuccess;
        size_t p = 0;
        for (size_t k = 0; k < chunks && e == hipSuccess; k++) {
            // the DMA of chunk k - 2 out of this buffer has finished
            if (k >= 2) e = hipEventSynchronize(done_[k & 1]);
            if (e != hipSuccess) break;
            const size_t c0 = k * kChunk, c1 = std::min(size, c0 + kChunk);
            split(pieces, k, c1, p, true);
            char* b = buf_[k & 1];
            team.run(uint32_t(jobs_.size()), [&](uint32_t j) {
                const Job& J = jobs_[j];
                if (J.host)
                    std::memcpy(b + J.off, J.host, J.bytes);
                else
                    std::memset(b + J.off, 0, J.bytes);
            });
            e = hipMemcpyAsync(dev + c0, b, c1 - c0, hipMemcpyHostToDevice, s_);
            if (e == hipSuccess) e = hipEventRecord(done_[k & 1], s_);
        }
        return e == hipSuccess ? hipStreamSynchronize(s_) : e;
    }

  private:
    struct Job {
        char* host;  // nullptr: padding (zeros)
        size_t off;  // in the chunk
        size_t bytes;
    };

    // Jobs of chunk k = [k * kChunk, c1): the parts of the pieces inside it,
    // at most kJob bytes each, and with `pad` the padding between them.  p is
    // the first piece that may reach into chunk k, advanced for the next chunk.
    void split(const std::vector<Piece>& pieces, size_t k, size_t c1, size_t& p, bool pad) {
        const size_t c0 = k * kChunk;
        jobs_.clear();
        while (p < pieces.size() && pieces[p].at + pieces[p].bytes <= c0) p++;
        size_t cursor = c0;
        for (size_t q = p; q < pieces.size() && pieces[q].at < c1; q++) {
            const size_t a = std::max(pieces[q].at, c0), z = std::min(pieces[q].at + pieces[q].bytes, c1);
            if (pad && a > cursor) jobs_.push_back(Job{nullptr, cursor - c0, a - cursor});
            for (size_t x = a; x < z; x += kJob)
                jobs_.push_back(Job{pieces[q].host + (x - pieces[q].at), x - c0, std::min(kJob, z - x)});
            cursor = z;
        }
        if (pad && c1 > cursor) jobs_.push_back(Job{nullptr, cursor - c0, c1 - cursor});
    }

    hipStream_t s_;
    char* buf_[2] = {nullptr, nullptr};
    hipEvent_t done_[2] = {nullptr, nullptr};
    std::vector<Job> jobs_;
};

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
    DevTemp orig_mem;
    if (e == hipSuccess) e = hipMalloc(&orig_mem.p, total * sizeof(uint32_t));
    if (e != hipSuccess) return hip_fail(e, "atlas_rt_blas_build_device (outputs)");
    uint32_t* const orig = static_cast<uint32_t*>(orig_mem.p);
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
    // the caller's triangles, meshes one after another, up through pinned staging
    Team team(default_threads());
    Staging stage(s);
    DevTemp verts_mem;
    std::vector<Piece> tri_in(n);
    for (uint32_t m = 0; m < n; m++)  // read only: up() copies from the pieces
        tri_in[m] = Piece{9 * off_o[m] * sizeof(float), 9 * size_t(c[m]) * sizeof(float),
                          reinterpret_cast<char*>(const_cast<float*>(v[m]))};
    e = stage.init();
    if (e == hipSuccess) e = hipMalloc(&verts_mem.p, std::max<size_t>(9 * total, 1) * sizeof(float));
    if (e == hipSuccess) e = stage.up(team, static_cast<char*>(verts_mem.p), 9 * total * sizeof(float), tri_in);
    if (e != hipSuccess) return hip_fail(e, "atlas_rt_blas_build_device (triangle upload)");
    // the host mirrors are sized (and their pages touched) by the team while
    // this thread waits for the GPU build
    std::exception_ptr sizing_err;
    std::thread sizing([&] {
        try {
            team.run(n, [&](uint32_t m) {
                atlas_rt_blas* b = made[first + m];
                b->nodes.resize(size_t(c[m]) - 1);
                b->tris.resize(c[m]);
                b->orig_idx.resize(c[m]);
            });
        } catch (...) {
            sizing_err = std::current_exception();
        }
    });
    std::vector<uint32_t> depth;
    atlas_rt_status st;
    try {
        st = lbvh_build(static_cast<const float*>(verts_mem.p), c, o, depth, s, ms);
    } catch (...) {
        sizing.join();
        throw;
    }
    sizing.join();
    if (sizing_err) std::rethrow_exception(sizing_err);
    verts_mem.reset();
    if (st != ATLAS_RT_OK) return st;
    for (uint32_t m = 0; m < n; m++) made[first + m]->max_depth = depth[m];
    // host mirrors, read back through the staging straight into each BLAS's
    // vectors: node and vertex records from the build arena, the caller's
    // triangle indices from `orig`
    std::vector<Piece> recs, idx;
    recs.reserve(2 * size_t(n));
    idx.reserve(n);
    for (uint32_t m = 0; m < n; m++) {
        atlas_rt_blas* b = made[first + m];
        if (c[m] > 1) recs.push_back(Piece{off_n[m], b->nodes.size() * sizeof(atlas_rt_gpu_bvh_node),
                                           reinterpret_cast<char*>(b->nodes.data())});
    }
    for (uint32_t m = 0; m < n; m++) {
        atlas_rt_blas* b = made[first + m];
        recs.push_back(Piece{off_v[m], b->tris.size() * sizeof(atlas_rt_gpu_bvh_triangle),
                             reinterpret_cast<char*>(b->tris.data())});
        idx.push_back(Piece{off_o[m] * sizeof(uint32_t), size_t(c[m]) * sizeof(uint32_t),
                            reinterpret_cast<char*>(made[first + m]->orig_idx.data())});
    }
    e = stage.down(team, base, tris_begin, recs);
    if (e == hipSuccess) e = stage.down(team, reinterpret_cast<const char*>(orig), total * sizeof(uint32_t), idx);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    orig_mem.reset();
    if (e != hipSuccess) return hip_fail(e, "atlas_rt_blas_build_device (host mirror)");
    // the wide collapse, one mesh per host thread
    std::vector<WideTree> w4(n), w8(n);
    team.run(n, [&](uint32_t m) {
        const atlas_rt_blas* b = made[first + m];
        std::vector<uint8_t> end_flags(c[m]);
        for (size_t i = 0; i < c[m]; i++) end_flags[i] = b->tris[i].v0[3] > 0.0f ? 1 : 0;
        w4[m] = collapse_wide(b->nodes, end_flags, 4, 4);
        w8[m] = collapse_wide(b->nodes, end_flags, 8, 4);
    });
    // wide arena: [4-wide nodes of every mesh][8-wide nodes]
    std::vector<size_t> off4(n), off8(n);
    std::vector<Piece> wide;
    wide.reserve(2 * size_t(n));
    size_t wat = 0;
    for (int pass = 0; pass < 2; pass++)
        for (uint32_t m = 0; m < n; m++) {
            std::vector<float>& d = pass == 0 ? w4[m].data : w8[m].data;
            (pass == 0 ? off4 : off8)[m] = wat;
            if (!d.empty()) wide.push_back(Piece{wat, d.size() * sizeof(float), reinterpret_cast<char*>(d.data())});
            wat = align_up(wat + d.size() * sizeof(float));
        }
    e = hipMalloc(&arena->wide, std::max<size_t>(wat, 1));
    if (e == hipSuccess) e = stage.up(team, static_cast<char*>(arena->wide), wat, wide);
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
        atlas_rt_status st;
        try {
            st = build_batch(verts + first, counts + first, n, dev, s, made, device_ms ? &ms : nullptr);
        } catch (const std::bad_alloc&) {
            st = fail(ATLAS_RT_E_OUT_OF_MEMORY, "atlas_rt_blas_build_device: host allocation failed");
        } catch (const std::exception& ex) {  // e.g. no host thread could be started
            st = fail(ATLAS_RT_E_OUT_OF_MEMORY, std::string("atlas_rt_blas_build_device: ") + ex.what());
        }
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
