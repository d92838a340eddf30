/*
 * atlas_rt.h - C ABI of the MI355X (gfx950) ray-tracing hot path.
 *
 * This is the drop-in boundary that replaces the GPU side of Atlas Engine's
 * software ray tracer.  The engine has no FFI for this path: it sits behind the
 * C++ classes Volume::BVH (src/engine/volume/BVH.h:114-138),
 * RayTracing::RayTracingWorld (src/engine/raytracing/RayTracingWorld.h:31-99),
 * Renderer::Helper::RayTracingHelper (src/engine/renderer/helper/RayTracingHelper.h:16-44)
 * and PathTracingRenderer (src/engine/renderer/PathTracingRenderer.h:20-46), plus a
 * GPU buffer ABI (src/engine/raytracing/RTStructures.h:14-110,
 * data/shader/raytracer/structures.hsh:9-208).  Every entry point below names
 * the reference interface it stands in for.  The C++ facades in
 * atlas-engine_amd/facade/ keep those class signatures and call this layer.
 *
 * Conventions
 *  - Every call returns atlas_rt_status; ATLAS_RT_OK is 0.  On failure,
 *    atlas_rt_last_error() returns a thread-local description.  No C++
 *    exception crosses this boundary.
 *  - Pointers named d_* are device pointers owned by the caller (borrowed for
 *    the duration of the call).  Handles (atlas_rt_blas*, ...) are owned by the
 *    library and released with the matching *_destroy.
 *  - `stream` is a hipStream_t passed as void* (NULL = the null stream).  Trace
 *    calls are asynchronous on that stream; handles are not internally
 *    synchronised (one thread per handle at a time).
 */
#ifndef ATLAS_RT_H
#define ATLAS_RT_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------------- */
/* Status codes                                                              */
/* ------------------------------------------------------------------------- */
typedef int32_t atlas_rt_status;
enum {
    ATLAS_RT_OK = 0,
    ATLAS_RT_E_INVALID_ARG = 1,    /* null pointer, zero count where >0 needed, bad flag */
    ATLAS_RT_E_SIZE_MISMATCH = 2,  /* aabbs/data sizes differ (BVH.cpp:18-19 silently returned) */
    ATLAS_RT_E_EMPTY = 3,          /* nothing to build (no triangles / no instances) */
    ATLAS_RT_E_HIP = 4,            /* a HIP runtime call failed */
    ATLAS_RT_E_STACK_OVERFLOW = 5, /* tree deeper than the configured maximum stack */
    ATLAS_RT_E_OUT_OF_MEMORY = 6,
    ATLAS_RT_E_UNSUPPORTED = 7,    /* variant not compiled / not available */
    ATLAS_RT_E_NO_DEVICE = 8       /* no HIP device visible */
};

/* Thread-local text for the last non-OK status returned on this thread. */
const char* atlas_rt_last_error(void);
/* Library version string. */
const char* atlas_rt_version(void);

/* ------------------------------------------------------------------------- */
/* GPU buffer ABI, byte-compatible with the engine                           */
/* ------------------------------------------------------------------------- */

/* GPUBVHNode / PackedBVHNode (RTStructures.h:95-103, structures.hsh:179-184,
 * unpacked at bvh.hsh:21-37).  ptr >= 0: inner node index; ptr < 0: leaf, ~ptr
 * is the first triangle (BLAS) or instance (TLAS). */
typedef struct atlas_rt_gpu_bvh_node {
    float left_min[3];
    float left_max[3];
    float right_min[3];
    float right_max[3];
    int32_t left_ptr;
    int32_t right_ptr;
    int32_t padding0;
    int32_t padding1;
} atlas_rt_gpu_bvh_node; /* 64 B */

/* GPUBVHTriangle / PackedBVHTriangle (RTStructures.h:23-27, packed at
 * MeshData.cpp:242-247): v0.w = +1 end of leaf / -1, v1.w = material index bits,
 * v2.w = opacity (-1 = opacity map). */
typedef struct atlas_rt_gpu_bvh_triangle {
    float v0[4];
    float v1[4];
    float v2[4];
} atlas_rt_gpu_bvh_triangle; /* 48 B */

/* GPUBVHInstance / Instance (RTStructures.h:85-93, structures.hsh:169-177).
 * inverse_matrix is mat3x4 = transpose(inverseGlobalMatrix) (RayTracingWorld.cpp:100):
 * column c (4 floats) is row c of the inverse object-to-world matrix, so the
 * object-space point is p_obj[c] = dot(vec4(p,1), inverse_matrix[c]). */
typedef struct atlas_rt_gpu_bvh_instance {
    float inverse_matrix[3][4];
    int32_t mesh_offset;     /* BLAS index in the world's BLAS table */
    int32_t material_offset;
    int32_t next_instance;   /* -1 at the end of a TLAS leaf (RayTracingWorld.cpp:294) */
    uint32_t mask;           /* ATLAS_RT_MASK_* bits */
} atlas_rt_gpu_bvh_instance; /* 64 B */

#define ATLAS_RT_MASK_ALL (1u << 7)    /* InstanceCullMasks::MaskAll  (RTStructures.h:10) */
#define ATLAS_RT_MASK_SHADOW (1u << 6) /* InstanceCullMasks::MaskShadow (RTStructures.h:11) */

/* PackedRay (structures.hsh:9-13, common.hsh:46-73): origin.w = ray ID bits,
 * hit = (t, hitID bits, hitInstanceID bits, unused). */
typedef struct atlas_rt_packed_ray {
    float origin[4];
    float direction[4];
    float hit[4];
} atlas_rt_packed_ray; /* 48 B */

/* Volume::BVHTriangle input (BVH.h:26-33): vertices + caller's triangle index. */
typedef struct atlas_rt_bvh_triangle {
    float v0[3];
    float v1[3];
    float v2[3];
    uint32_t idx;
} atlas_rt_bvh_triangle; /* 40 B */

/* Volume::AABB (AABB.h:102-103). */
typedef struct atlas_rt_aabb {
    float min[3];
    float max[3];
} atlas_rt_aabb; /* 24 B */

/* Compact ray for bulk tracing: 32 B in. */
typedef struct atlas_rt_ray {
    float origin[3];
    float tmin;
    float direction[3];
    float tmax;
} atlas_rt_ray;

/* Compact hit: 16 B out.  prim_id = caller's original triangle index
 * (BVHTriangle::idx), -1 on a miss (then t = tmax, u = v = 0). */
typedef struct atlas_rt_hit {
    float t;
    float u;
    float v;
    int32_t prim_id;
} atlas_rt_hit;

/* ------------------------------------------------------------------------- */
/* Device                                                                    */
/* ------------------------------------------------------------------------- */
/* Number of visible HIP devices (0 when none).  Never fails. */
int32_t atlas_rt_device_count(void);

/* ------------------------------------------------------------------------- */
/* BLAS: bottom-level BVH over one triangle mesh                             */
/* ------------------------------------------------------------------------- */
typedef struct atlas_rt_blas atlas_rt_blas;

#define ATLAS_RT_BUILD_SERIAL 0u      /* single-threaded build */
#define ATLAS_RT_BUILD_PARALLEL 1u    /* parallel subtree build (BVH.cpp:310, parallelBuild=true) */
#define ATLAS_RT_BUILD_HOST_ONLY 2u   /* keep the result on the host (no device upload) */

/* Replaces Volume::BVH::BVH(aabbs, data, parallelBuild) (BVH.cpp:14-56) followed by
 * MeshData::BuildBVH's packing (MeshData.cpp:172-269) and Mesh::BuildBVH's upload
 * (Mesh.cpp:97-108).  Root SBVH + binned SAH below + median fallback, one
 * reference per leaf, flattened pre-order with the larger-area child left.
 * `aabbs` may be NULL (then each triangle's vertex bounds are used, as
 * MeshData.cpp:143-153 computes them). */
atlas_rt_status atlas_rt_blas_build(const atlas_rt_aabb* aabbs, const atlas_rt_bvh_triangle* tris,
                                    uint32_t count, uint32_t flags, atlas_rt_blas** out);

/* Wraps an already packed BLAS in the engine's layout (what MeshData keeps in
 * gpuBvhNodes / gpuBvhTriangles, MeshData.h:120-122).  orig_idx[i] is the caller's
 * index of BVH slot i (may be NULL: then the slot is the index). */
atlas_rt_status atlas_rt_blas_create(const atlas_rt_gpu_bvh_node* nodes, uint32_t node_count,
                                     const atlas_rt_gpu_bvh_triangle* tris, const uint32_t* orig_idx,
                                     uint32_t tri_count, atlas_rt_blas** out);

atlas_rt_status atlas_rt_blas_destroy(atlas_rt_blas* blas);

/* Sizes of the packed arrays (nodes may be 0 when the mesh is one triangle). */
atlas_rt_status atlas_rt_blas_info(const atlas_rt_blas* blas, uint32_t* node_count, uint32_t* tri_count,
                                   uint32_t* max_depth);

/* Copies the engine-layout arrays out (BVH::GetTree(), BVH.cpp:219-223, and the
 * BVH-ordered data/refs).  Any output pointer may be NULL. */
atlas_rt_status atlas_rt_blas_export(const atlas_rt_blas* blas, atlas_rt_gpu_bvh_node* nodes,
                                     atlas_rt_gpu_bvh_triangle* tris, uint32_t* orig_idx);

/* ------------------------------------------------------------------------- */
/* Bulk tracing against one BLAS                                             */
/* ------------------------------------------------------------------------- */

/* Traversal kernels.  ATLAS_RT_KERNEL_AUTO picks the autotuned variant for the
 * current device and BLAS shape. */
#define ATLAS_RT_KERNEL_AUTO 0
#define ATLAS_RT_KERNEL_BVH2_EXACT 1    /* engine order: near child first, left on ties */
#define ATLAS_RT_KERNEL_BVH2_PERSISTENT 2
#define ATLAS_RT_KERNEL_BVH4 3
#define ATLAS_RT_KERNEL_BVH8 4

/* Closest hit, engine semantics of HitClosest (bvh.hsh:191-273, CheckLeafClosest
 * :44-72, IntersectAABB/IntersectTriangle intersections.hsh:18-58): accept
 * t > tmin, t < tmax, t < current t; first accepted wins equal t.  d_slot (may
 * be NULL) receives the BVH slot (the engine's hitID) per ray. */
atlas_rt_status atlas_rt_trace_closest_rays(const atlas_rt_blas* blas, const atlas_rt_ray* d_rays,
                                            atlas_rt_hit* d_hits, int32_t* d_slot, uint32_t count,
                                            int32_t kernel, void* stream);

/* Same trace, additionally writing per ray (inner nodes fetched, triangles
 * tested) to d_visits (2 x uint32 per ray).  The counts define the
 * algorithmic bytes of a trace: 64 B per node (GPUBVHNode) and 48 B per
 * triangle test (GPUBVHTriangle).  Used to price the roofline; the profiler
 * counters of this profiling launch are not representative of the plain one. */
atlas_rt_status atlas_rt_trace_closest_stats(const atlas_rt_blas* blas, const atlas_rt_ray* d_rays,
                                             atlas_rt_hit* d_hits, int32_t* d_slot, uint32_t* d_visits,
                                             uint32_t count, void* stream);

/* Any hit, HitAny semantics (bvh.hsh:359-441, CheckLeaf :74-104): left child
 * first, stop at the first accepted triangle.  d_hit[i] = 1 if occluded. */
atlas_rt_status atlas_rt_trace_any_rays(const atlas_rt_blas* blas, const atlas_rt_ray* d_rays,
                                        uint32_t* d_hit, uint32_t count, int32_t kernel, void* stream);

/* Host-buffer convenience for CPU callers (BVH::GetIntersection, BVH.cpp:103-174):
 * copies rays in, traces on the device, copies hits back, synchronises. */
atlas_rt_status atlas_rt_trace_closest_host(const atlas_rt_blas* blas, const atlas_rt_ray* rays,
                                            atlas_rt_hit* hits, uint32_t count);
atlas_rt_status atlas_rt_trace_any_host(const atlas_rt_blas* blas, const atlas_rt_ray* rays,
                                        uint32_t* hit, uint32_t count);

/* Per-launch timing of the last trace call on this BLAS (device time between
 * HIP events recorded around the kernel on its own stream), milliseconds. */
atlas_rt_status atlas_rt_last_kernel_ms(const atlas_rt_blas* blas, float* ms);

#ifdef __cplusplus
}
#endif

#endif /* ATLAS_RT_H */
