// Internal host-side declarations shared by the C ABI, the BLAS/world objects
// and the kernel launchers.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <memory>
#include <string>
#include <vector>

#include "../../include/atlas_rt.h"

namespace atlas_rt {

// Thread-local last error, set by fail().
void set_error(const std::string& msg);
atlas_rt_status fail(atlas_rt_status code, const std::string& msg);
atlas_rt_status hip_fail(hipError_t e, const char* what);

#define ATLAS_HIP_TRY(expr)                                  \
    do {                                                     \
        hipError_t _e = (expr);                              \
        if (_e != hipSuccess) return ::atlas_rt::hip_fail(_e, #expr); \
    } while (0)

// Device view of one BLAS (also the element type of BLAS tables).
struct BlasDev {
    const float4* nodes;  // engine GPUBVHNode, 4 float4 per node
    const float4* tris;   // 3 float4 per slot: (v0, caller idx bits), (v1-v0, end flag), (v2-v0, 0)
    const float4* vtx;    // engine GPUBVHTriangle, 3 float4 per slot (v0, v1, v2)
    const float4* nodes4; // 4-wide nodes collapsed from `nodes` (wide.h), root 0
    const float4* nodes8; // 8-wide nodes
    uint32_t node_count;  // binary nodes; 0 = the BLAS is a single leaf (root ~0)
    uint32_t tri_count;
};

// Stack spill scratch, grown on demand and owned by a handle.
struct Scratch {
    void* ptr = nullptr;
    size_t bytes = 0;
    int device = -1;
    atlas_rt_status ensure(size_t need);
    void release();
};

struct LaunchTiming {
    hipEvent_t start = nullptr, stop = nullptr;
    bool pending = false;
    float last_ms = 0.0f;
    atlas_rt_status begin(hipStream_t s);
    atlas_rt_status end(hipStream_t s);
    atlas_rt_status read(float* ms);
    void release();
};

}  // namespace atlas_rt

struct atlas_rt_blas {
    // Host copies in the engine's layout (what BVH::GetTree() and the packed
    // MeshData arrays hold).
    std::vector<atlas_rt_gpu_bvh_node> nodes;
    std::vector<atlas_rt_gpu_bvh_triangle> tris;
    std::vector<uint32_t> orig_idx;
    uint32_t max_depth = 0;

    // Device arrays:
    //  d_nodes: engine GPUBVHNode layout, 4 x float4 per node (data0..data3,
    //           structures.hsh:179-184);
    //  d_tris : 3 x float4 per BVH slot: (v0.xyz, caller index bits),
    //           (e0 = v1 - v0, end-of-leaf +1/-1), (e1 = v2 - v0, 0);
    //  d_vtx  : engine GPUBVHTriangle array, read once per ray for the
    //           reported (t, u, v) of the final hit.
    int device = -1;
    bool on_device = false;
    float4* d_nodes = nullptr;
    float4* d_tris = nullptr;
    float4* d_vtx = nullptr;
    //  d_shade: optional, 3 x float4 per slot: (n0, material index bits),
    //           (n1, baked opacity), (n2, 0) (the path tracer's shading data);
    //  d_uv   : optional, 2 x float4 per slot: (uv0, uv1), (uv2, 0, 0).
    float4* d_shade = nullptr;
    float4* d_uv = nullptr;
    //  d_tangent: optional, 2 x float4 per slot: (tangent, 0), (bitangent, 0).
    float4* d_tangent = nullptr;
    //  d_color: optional, 3 x float4 per slot: the vertex colours (RGBA)
    //           unpacked from GPUTriangle records (atlas_rt_blas_set_triangles).
    float4* d_color = nullptr;
    bool has_alpha = false;             // some baked opacity is not 1
    std::vector<float> shade_normals;   // host copy, 9 floats per slot
    std::vector<int32_t> shade_mat;     // host copy, 1 per slot
    //  d_nodes4 / d_nodes8: wide nodes (wide.h); cap4 / cap8 bound the
    //  traversal stack of the wide kernels.
    float4* d_nodes4 = nullptr;
    float4* d_nodes8 = nullptr;
    uint32_t cap4 = 1, cap8 = 1;
    //  arena: set for BLASes of one device build, whose d_nodes, d_tris,
    //  d_vtx, d_nodes4 and d_nodes8 are slices of shared allocations that the
    //  last BLAS of the batch frees.
    std::shared_ptr<void> arena;
    atlas_rt::Scratch spill;
    atlas_rt::Scratch flags;  // [0] ray counter, [1] finished waves, [2] stack overflow flag
    atlas_rt::LaunchTiming timing;
    // Most traversal stack entries per ray (atlas_rt_blas_set_stack_limit);
    // 0: sized from the tree, which never overflows.
    uint32_t stack_limit = 0;
};

namespace atlas_rt {

// Resident (persistent) kernels hand out their work from kPullRanges
// contiguous ranges, one counter each: a wave starts on the range of its XCD
// and moves on when that runs dry, spreading the counter updates over eight
// addresses (one per XCD of the MI355X) instead of one.
constexpr uint32_t kPullRanges = 8;
// BLAS flag block: [0] unused, [1] finished waves, [2] stack overflow flag,
// and range r's counter at word kRangeWord * (r + 1) (128 B apart).
constexpr uint32_t kRangeWord = 32;
constexpr size_t kFlagBytes = size_t(kRangeWord) * (kPullRanges + 1) * sizeof(uint32_t);

// Kernel launchers (trace_blas.hip).
struct TraceLaunch {
    const void* rays = nullptr;  // atlas_rt_ray[count]
    void* hits = nullptr;        // atlas_rt_hit[count] (closest) or uint32_t[count] (any)
    int32_t* slots = nullptr;
    uint32_t* visits = nullptr;  // optional uint2 per ray: (inner nodes fetched, triangles tested)
    uint32_t count = 0;
    int32_t kernel = ATLAS_RT_KERNEL_AUTO;
    uint32_t refill = 0;  // persistent kernels: refill a wave below this many active lanes (0 = default)
    hipStream_t stream = nullptr;
};

int device_cu_count();

// The stack entries a trace may use: what the tree needs, or fewer when the
// caller set a limit (limit 0: none).
inline uint32_t limit_cap(uint32_t need, uint32_t limit) { return limit && limit < need ? limit : need; }

// Waits for `stream`, then reads and clears the stack-overflow word
// flags[2] of a flag block (NULL: no trace has run yet).
atlas_rt_status take_overflow_flag(void* flags, void* stream, const char* who);

atlas_rt_status launch_blas_closest(atlas_rt_blas& blas, const TraceLaunch& L);
// BVH::GetIntersection / GetIntersectionAny on the host over the BLAS's host
// copies (host_trace.cpp).
void host_trace_closest(const atlas_rt_blas& b, const atlas_rt_ray* rays, atlas_rt_hit* hits, uint32_t count);
void host_trace_any(const atlas_rt_blas& b, const atlas_rt_ray* rays, uint32_t* hit, uint32_t count);
// The binary persistent kernels with their counters and stack spill in the
// library's fixed device scratch (no allocation).
atlas_rt_status launch_blas_closest_noalloc(const atlas_rt_blas& blas, const TraceLaunch& L);
atlas_rt_status launch_blas_any(atlas_rt_blas& blas, const TraceLaunch& L);

// Upload the engine-layout arrays of `blas` to the current device.
atlas_rt_status upload_blas(atlas_rt_blas& blas);
// The 4- and 8-wide node arrays of a BLAS whose host copy is filled.
atlas_rt_status upload_wide(atlas_rt_blas& blas);

}  // namespace atlas_rt
