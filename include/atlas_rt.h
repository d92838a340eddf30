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

/* GPUTriangle / PackedTriangle (RTStructures.h:14-21), packed per BVH slot by
 * MeshData::BuildBVH (MeshData.cpp:172-240) and unpacked by UnpackTriangle
 * (raytracer/common.hsh:17-44):
 *   v0..v2 = (vertex xyz, bits of the vertex normal as 10-10-10-2,
 *            Packing.cpp:24-35),
 *   d0 = (bits of half2 uv0, uv1, uv2, bits of the material index),
 *   d1 = (bits of the 10-10-10-2 tangent, of the bitangent, +1 end of leaf / -1, 0),
 *   d2 = (bits of RGBA8 colour 0, 1, 2, opacity). */
typedef struct atlas_rt_gpu_triangle {
    float v0[4];
    float v1[4];
    float v2[4];
    float d0[4];
    float d1[4];
    float d2[4];
} atlas_rt_gpu_triangle; /* 96 B */

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

/* Traversal kernels.  ATLAS_RT_KERNEL_AUTO picks the default (single BLAS:
 * leaf-batched resident waves, BVH2_PERSISTENT | LEAF_BATCH, from 65,536 rays
 * up; one ray per lane, BVH2_EXACT, below) or the autotuned variant (batch
 * sets).
 *  BVH2_EXACT       one ray per lane over the engine's binary nodes, engine order
 *  BVH2_PERSISTENT  resident waves refilled with new rays, same traversal
 *  BVH4 / BVH8      resident waves over 4- / 8-wide nodes collapsed from the
 *                   engine's tree: hit children visited nearest first
 * Or'ing ATLAS_RT_KERNEL_STACK_REGISTER keeps the top of each lane's
 * traversal stack in registers (the rest in LDS, then global memory) instead
 * of all of it in LDS.  Every variant returns the engine's nearest hit
 * distance bit for bit; BVH4/BVH8 may report a different triangle where two
 * are hit at exactly the same t (the engine reports the first it tests). */
#define ATLAS_RT_KERNEL_AUTO 0
#define ATLAS_RT_KERNEL_BVH2_EXACT 1    /* engine order: near child first, left on ties */
#define ATLAS_RT_KERNEL_BVH2_PERSISTENT 2
#define ATLAS_RT_KERNEL_BVH4 3
#define ATLAS_RT_KERNEL_BVH8 4
#define ATLAS_RT_KERNEL_VARIANT_MASK 0xF
#define ATLAS_RT_KERNEL_STACK_REGISTER 0x10
/* With the binary kernels (1, 2) and, for single-BLAS traces, the 4-wide
 * kernel (3): the widened float slab test.  Slab distances
 * use the float reciprocal of the direction and a box counts as hit within
 * 2^-20 of its entry / exit distances, so every box the engine-exact test
 * enters is entered; triangles stay exact.  Same nearest hit as the engine
 * order; which triangle wins an exact equal-t tie may differ, and a hit that
 * the engine's slab test culls at a box face may be found. */
#define ATLAS_RT_KERNEL_SLAB_FLOAT 0x20
/* Or'ing ATLAS_RT_KERNEL_LEAF_BATCH into ATLAS_RT_KERNEL_BVH2_PERSISTENT
 * (with or without the two bits above), or, for single-BLAS traces, into
 * ATLAS_RT_KERNEL_BVH4, makes a wave take triangle steps only
 * once a quarter of its live lanes wait at a leaf, inner-node steps
 * otherwise.  Every ray's own traversal is unchanged, so results are those of
 * the same kernel without the bit. */
#define ATLAS_RT_KERNEL_LEAF_BATCH 0x40

/* Closest hit, engine semantics of HitClosest (bvh.hsh:191-273, CheckLeafClosest
 * :44-72, IntersectAABB/IntersectTriangle intersections.hsh:18-58): accept
 * t > tmin, t < tmax, t < current t; first accepted wins equal t.  d_slot (may
 * be NULL) receives the BVH slot (the engine's hitID) per ray.
 * Ordering: the traces of one BLAS (closest, any, stats) keep their ray and
 * wave counters, traversal stack spill and overflow flag in the BLAS, so they
 * are stream-ordered: traces of the same BLAS on different streams must not
 * overlap (make one stream wait on an event recorded after the other's
 * trace).  Traces of different BLASes are independent. */
atlas_rt_status atlas_rt_trace_closest_rays(const atlas_rt_blas* blas, const atlas_rt_ray* d_rays,
                                            atlas_rt_hit* d_hits, int32_t* d_slot, uint32_t count,
                                            int32_t kernel, void* stream);

/* No counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  The same closest-hit trace and results as
 * atlas_rt_trace_closest_rays with the same kernel, restricted to the binary
 * persistent kernels (ATLAS_RT_KERNEL_BVH2_PERSISTENT, optionally with the
 * STACK_REGISTER, SLAB_FLOAT and LEAF_BATCH bits; AUTO = persistent with leaf
 * batching).  The kernels' ray and wave counters and their traversal stack
 * spill live in a fixed 16 MiB + 2 KiB device array that each GPU holds until
 * the process ends; nothing is allocated or freed, and the counters are
 * cleared by a small kernel, not a memset.  A BLAS whose stack bound (depth +
 * 1, or its stack limit) exceeds what that array holds (24 + 16 MiB /
 * (4 B x 1024 lanes per CU x CUs), i.e. 40 entries on 256 CUs) is rejected with
 * ATLAS_RT_E_STACK_OVERFLOW.  A ray that would exceed a stack limit set with
 * atlas_rt_blas_set_stack_limit ends as a miss, and no overflow status is
 * kept.  Every call on one GPU shares that GPU's array: calls that overlap on
 * one GPU (other streams, other threads) race. */
atlas_rt_status atlas_rt_trace_closest_rays_noalloc(const atlas_rt_blas* blas, const atlas_rt_ray* d_rays,
                                                    atlas_rt_hit* d_hits, int32_t* d_slot, uint32_t count,
                                                    int32_t kernel, void* stream);

/* Same trace, additionally writing per ray (inner nodes fetched, triangles
 * tested) to d_visits (2 x uint32 per ray).  The counts define the
 * algorithmic bytes of a trace: 64 B per node (GPUBVHNode) and 48 B per
 * triangle test (GPUBVHTriangle).  Used to price the roofline; the profiler
 * counters of this profiling launch are not representative of the plain one. */
/* Ray coherence sorting (raytracer/binning.csh, binningOffset.csh; bin of
 * raytracer/tracing.hsh:18-21): a stable counting sort of `count` device rays
 * into d_out by the octahedral bin of their direction (common/octahedron.hsh:
 * 23-37; 8 x 8 cells, bin = cy * 8 + cx).  The engine truncates a map
 * coordinate of exactly 1 to cell 8, past its grid; here the cell is 7.  A
 * NaN direction falls in bin 0.  The engine's order inside a bin comes from
 * atomics; here rays keep their input order, so the result is deterministic.
 * d_perm (optional): d_out[i] = d_rays[d_perm[i]], to map hits back.
 * d_bin_offsets (optional, 65 entries): first output index of each bin, then
 * count.  d_workspace: atlas_rt_sort_rays_workspace(count) bytes of device
 * memory; the call allocates nothing.  Asynchronous on `stream`. */
atlas_rt_status atlas_rt_sort_rays_workspace(uint32_t count, size_t* bytes);
atlas_rt_status atlas_rt_sort_rays_by_direction(const atlas_rt_ray* d_rays, uint32_t count, atlas_rt_ray* d_out,
                                                uint32_t* d_perm, uint32_t* d_bin_offsets, void* d_workspace,
                                                size_t workspace_bytes, void* stream);

atlas_rt_status atlas_rt_trace_closest_stats(const atlas_rt_blas* blas, const atlas_rt_ray* d_rays,
                                             atlas_rt_hit* d_hits, int32_t* d_slot, uint32_t* d_visits,
                                             uint32_t count, void* stream);

/* Any hit, HitAny semantics (bvh.hsh:359-441, CheckLeaf :74-104): left child
 * first, stop at the first accepted triangle.  d_hit[i] = 1 if occluded. */
atlas_rt_status atlas_rt_trace_any_rays(const atlas_rt_blas* blas, const atlas_rt_ray* d_rays,
                                        uint32_t* d_hit, uint32_t count, int32_t kernel, void* stream);

/* The engine's CPU traversal for CPU callers, on the host, with host buffers:
 * BVH::GetIntersection / GetIntersectionAny (BVH.cpp:103-217) over Ray::Intersects
 * (Ray.cpp:28-76).  These are the engine's CPU semantics, not HitClosest's:
 * float-reciprocal slab test, a stack of (node, entry distance) with entries
 * beyond the current hit skipped, both hit children pushed nearer last (left
 * last on ties), triangles accepted at t >= 0 (tMin bounds only the boxes)
 * and t < current t; any hit pops left first and stops at the first t <
 * tMax.  prim_id is the caller's index, (t, u, v) the FP32 solution; a hit is
 * reported when t < the ray's tMax (the engine compares against the tMax it
 * overwrote during the walk, BVH.cpp:141,172).  Works for host-only BLASes;
 * large batches run on the host's hardware threads. */
atlas_rt_status atlas_rt_trace_closest_host(const atlas_rt_blas* blas, const atlas_rt_ray* rays,
                                            atlas_rt_hit* hits, uint32_t count);
atlas_rt_status atlas_rt_trace_any_host(const atlas_rt_blas* blas, const atlas_rt_ray* rays,
                                        uint32_t* hit, uint32_t count);

/* Traversal stack bound.  By default (entries = 0) every trace of the BLAS
 * sizes its stack from the tree's depth (LDS rows, then a global spill), so
 * no ray ever runs out of stack.  entries > 0 caps the stack at that many
 * entries, as the engine's fixed 32-entry stack does (bvh.hsh:16, which has
 * no overflow check): a ray that would push past the cap stops, is reported
 * as a miss (prim_id -1, t = tmax, slot -1), and raises the BLAS's overflow
 * flag.  Set it while no trace of the BLAS is in flight. */
atlas_rt_status atlas_rt_blas_set_stack_limit(atlas_rt_blas* blas, uint32_t entries);

/* Waits for `stream`, then reads and clears the BLAS's overflow flag:
 * ATLAS_RT_E_STACK_OVERFLOW if a trace of this BLAS since the last call
 * stopped a ray at the stack limit, ATLAS_RT_OK otherwise.  The engine has
 * no counterpart: its stack overflow goes unnoticed. */
atlas_rt_status atlas_rt_blas_trace_status(atlas_rt_blas* blas, void* stream);

/* Per-launch timing of the last trace call on this BLAS (device time between
 * HIP events recorded around the kernel on its own stream), milliseconds. */
atlas_rt_status atlas_rt_last_kernel_ms(const atlas_rt_blas* blas, float* ms);

/* Volume::BVH::BVH(aabbs, parallelBuild) (BVH.cpp:58-101): the AABB-only
 * tree the engine builds over instances (64 bins, one item per leaf, the
 * single-item special node of BVH.cpp:80-88).  Host only.  On input
 * *node_count / *ref_count are the capacities of nodes / ref_idx+end_of_node;
 * on output the sizes (count - 1 nodes and count references for count >= 2;
 * one node and two references for count == 1).  Too small a capacity returns
 * ATLAS_RT_E_INVALID_ARG with the required sizes written. */
atlas_rt_status atlas_rt_aabb_tree_build(const atlas_rt_aabb* aabbs, uint32_t count, uint32_t flags,
                                         atlas_rt_gpu_bvh_node* nodes, uint32_t* node_count, uint32_t* ref_idx,
                                         uint8_t* end_of_node, uint32_t* ref_count);

/* Builds `count` BLASes at once, one host thread per mesh (the engine builds
 * each mesh's BVH as its own job, editor/App.cpp:210-222 -> Mesh::BuildBVH).
 * tris[k] holds counts[k] triangles; out[k] receives the handle. */
atlas_rt_status atlas_rt_blas_build_many(const atlas_rt_bvh_triangle* const* tris, const uint32_t* counts,
                                         uint32_t count, uint32_t flags, atlas_rt_blas** out);

/* GPU builder (no engine counterpart; it serves the same callers as
 * atlas_rt_blas_build_many: Mesh::BuildBVH jobs, editor/App.cpp:210-222).
 * verts[k]: counts[k] triangles of 9 floats (v0, v1, v2) in the caller's
 * order, host memory; at most 2^24 triangles per mesh.  A linear BVH (30-bit
 * Morton codes over the centroid bounds, Karras radix tree), flattened like
 * BVHBuilder::Flatten (BVH.cpp:408-441): pre-order, the larger-area child
 * left, one triangle per leaf.  Same node and triangle format as
 * atlas_rt_blas_build, so every trace entry point accepts the result; the
 * tree differs from the SAH/SBVH one (nearest hits are the same, slots and
 * traversal cost differ).  Host copies are mirrored for export.  device_ms
 * (may be NULL) receives the GPU time of the build kernels, uploads excluded.
 * Synchronous on `stream`.  During the call the library uses up to 16 host
 * threads and two 16 MiB pinned host buffers (the triangle upload, the host
 * mirrors' readback and the wide nodes go through them), released on return;
 * ATLAS_RT_E_OUT_OF_MEMORY if host memory or threads run out. */
atlas_rt_status atlas_rt_blas_build_device(const float* const* verts, const uint32_t* counts, uint32_t count,
                                           uint32_t flags, void* stream, atlas_rt_blas** out, float* device_ms);

/* ------------------------------------------------------------------------- */
/* Batched traces: many independent BLAS / ray-batch pairs in one launch     */
/* ------------------------------------------------------------------------- */
/* No engine counterpart: the engine traces one ray buffer against one world
 * per dispatch (RayTracingHelper::DispatchHitClosest).  A batch set is a
 * table of (BLAS, ray range) pairs traced by one persistent launch: waves pull
 * rays from a global counter over the concatenated ranges, so uneven batch
 * sizes keep every CU busy. */
typedef struct atlas_rt_batch {
    uint32_t blas;       /* index into the BLAS table given at creation */
    uint32_t ray_offset; /* first ray of the batch in the ray / hit arrays */
    uint32_t ray_count;
    uint32_t reserved;
} atlas_rt_batch;

typedef struct atlas_rt_batch_set atlas_rt_batch_set;

/* Launch configuration chosen by the autotuner or forced by the caller. */
typedef struct atlas_rt_launch_config {
    int32_t kernel;       /* ATLAS_RT_KERNEL_*; AUTO = tuned */
    uint32_t block;       /* work-group size: 64, 128 or 256 (0 = default) */
    uint32_t refill;      /* persistent kernels: refill a wave below this many live lanes (0 = default) */
    uint32_t reserved;
} atlas_rt_launch_config;

atlas_rt_status atlas_rt_batch_create(const atlas_rt_blas* const* blases, uint32_t blas_count,
                                      const atlas_rt_batch* batches, uint32_t batch_count,
                                      atlas_rt_batch_set** out);
atlas_rt_status atlas_rt_batch_destroy(atlas_rt_batch_set* set);

/* Closest hit for every ray of every batch: hits[i] for rays[i] (16 B
 * records as atlas_rt_trace_closest_rays).  d_count (may be NULL) caps the
 * total number of rays processed (read on the device), so one captured graph
 * can serve uneven ray counts. */
atlas_rt_status atlas_rt_batch_trace_closest(atlas_rt_batch_set* set, const atlas_rt_ray* d_rays,
                                             atlas_rt_hit* d_hits, const uint32_t* d_count,
                                             const atlas_rt_launch_config* config, void* stream);

/* The autotuner's choice for this set and its trials, timed with HIP events
 * on the first AUTO launch: node width (2, 4, 8) x stack placement (LDS or
 * register top) x work-group size (64, 128, 256), binary nodes also one ray
 * per lane and at three refill thresholds.  Wide nodes run in resident waves
 * only (a grid request returns ATLAS_RT_E_UNSUPPORTED).
 * *n_trials is the capacity of trials/trial_ms on input, the trial count on
 * output.  Any output may be NULL. */
atlas_rt_status atlas_rt_batch_tuning(const atlas_rt_batch_set* set, atlas_rt_launch_config* best,
                                      uint32_t* n_trials, atlas_rt_launch_config* trials, float* trial_ms);

/* Device time of the last batched launch (HIP events on its stream), ms. */
atlas_rt_status atlas_rt_batch_last_kernel_ms(atlas_rt_batch_set* set, float* ms);

/* Records the batched trace (same arguments) into a HIP graph owned by the
 * set; atlas_rt_batch_graph_launch replays it on a stream. */
atlas_rt_status atlas_rt_batch_graph_capture(atlas_rt_batch_set* set, const atlas_rt_ray* d_rays,
                                             atlas_rt_hit* d_hits, const uint32_t* d_count,
                                             const atlas_rt_launch_config* config);
atlas_rt_status atlas_rt_batch_graph_launch(atlas_rt_batch_set* set, void* stream);

/* ------------------------------------------------------------------------- */
/* World: TLAS over instances of BLASes (RayTracingWorld)                    */
/* ------------------------------------------------------------------------- */
typedef struct atlas_rt_world atlas_rt_world;

/* One instance, as RayTracingWorld::Update reads it from a mesh + transform
 * component (RayTracingWorld.cpp:91-114). */
typedef struct atlas_rt_instance_desc {
    float object_to_world[3][4]; /* rows 0..2 of TransformComponent::globalMatrix */
    float world_to_object[3][4]; /* rows 0..2 of inverseGlobalMatrix (packed as GPUBVHInstance.inverseMatrix) */
    uint32_t blas;               /* index into the BLAS table passed to update */
    int32_t material_offset;
    uint32_t mask;               /* ATLAS_RT_MASK_ALL | (castShadow ? ATLAS_RT_MASK_SHADOW : 0) */
    uint32_t padding;
} atlas_rt_instance_desc; /* 112 B */

atlas_rt_status atlas_rt_world_create(atlas_rt_world** out);
atlas_rt_status atlas_rt_world_destroy(atlas_rt_world* world);

/* Replaces RayTracingWorld::Update + UpdateForSoftwareRayTracing
 * (RayTracingWorld.cpp:32-162, 267-307): builds the TLAS over the instances'
 * world AABBs with the engine's AABB builder (BVH.cpp:58-101, 64 bins, one
 * instance per leaf), packs GPUBVHInstance records in TLAS leaf order and
 * uploads TLAS nodes, instances and the BLAS table.  world_aabbs may be NULL:
 * then each instance's AABB is its BLAS root bounds transformed by
 * object_to_world (8 corners, AABB::Transform, AABB.cpp:35-60).  Not
 * concurrent with traces on the same world. */
atlas_rt_status atlas_rt_world_update(atlas_rt_world* world, const atlas_rt_instance_desc* instances,
                                      uint32_t instance_count, const atlas_rt_aabb* world_aabbs,
                                      const atlas_rt_blas* const* blases, uint32_t blas_count);

/* Per-frame update without a rebuild (the engine rebuilds its TLAS every
 * frame in RayTracingWorld::Update, RayTracingWorld.cpp:32-162; this keeps
 * the last update's tree and instance order).  d_instances: device memory,
 * the same instances in the caller's order of the last atlas_rt_world_update;
 * only object_to_world and world_to_object are read.  On `stream`: every
 * instance record takes its new inverse matrix, every instance its new world
 * box (as atlas_rt_world_update computes it with world_aabbs NULL), and every
 * TLAS node's child boxes the union of the boxes below, level by level.
 * Nearest hits are those of a rebuild; traversal cost grows as instances move
 * away from where the tree was built, so rebuild now and then.  A path tracer
 * created on the world sees the new transforms; its triangle lights move only
 * after atlas_rt_pathtracer_set_emissive_lights.  Stream-ordered: traces on
 * other streams must not overlap it. */
atlas_rt_status atlas_rt_world_refit(atlas_rt_world* world, const atlas_rt_instance_desc* d_instances, void* stream);

/* The traversal stack bound and overflow flag of a world's traces (TLAS and
 * BLAS levels share one stack, as in bvh.hsh), with the meaning of
 * atlas_rt_blas_set_stack_limit / atlas_rt_blas_trace_status.  The path
 * tracer's traces follow the limit of its world. */
atlas_rt_status atlas_rt_world_set_stack_limit(atlas_rt_world* world, uint32_t entries);
atlas_rt_status atlas_rt_world_trace_status(atlas_rt_world* world, void* stream);

/* Object history of the last update, RayTracingWorld::Update with
 * includeObjectHistory (RayTracingWorld.cpp:116-117, read as
 * instanceLastMatrices by rayHit.csh:80-101): last_object_to_world holds
 * instance_count x 12 floats, rows 0..2 of each instance's
 * TransformComponent::lastGlobalMatrix, in the caller's order of the last
 * atlas_rt_world_update.  A realtime path tracer's next frame takes its
 * velocities from them.  Without a history for the current update, a
 * realtime frame after an update matches instances by their index in the
 * update's array (an instance at an index the previous array did not have
 * starts without motion).  NULL clears the history. */
atlas_rt_status atlas_rt_world_set_history(atlas_rt_world* world, const float* last_object_to_world,
                                           uint32_t instance_count);

atlas_rt_status atlas_rt_world_info(const atlas_rt_world* world, uint32_t* tlas_node_count,
                                    uint32_t* instance_count, uint32_t* max_depth);
/* Copies the packed TLAS nodes and GPUBVHInstance records (TLAS leaf order)
 * and, per packed instance, the caller's instance index. */
atlas_rt_status atlas_rt_world_export(const atlas_rt_world* world, atlas_rt_gpu_bvh_node* tlas_nodes,
                                      atlas_rt_gpu_bvh_instance* instances, uint32_t* instance_order);

/* Engine-shaped trace, RayTracingHelper::DispatchHitClosest + traceClosest.csh
 * (RayTracingHelper.cpp:224-364, traceClosest.csh:12-36): for index <
 * *d_count (read on the device; no host sync), unpack the PackedRay, set
 * hitID = -1, and if ray.ID >= 0 run HitClosest(ray, mask, tmin, tmax);
 * write PackedRay with hit = (t, hitID = BVH slot, hitInstanceID).
 * max_count bounds the grid (the engine's ray buffer size).  d_count may be
 * NULL, then max_count rays are traced. */
atlas_rt_status atlas_rt_trace_closest(const atlas_rt_world* world, const atlas_rt_packed_ray* d_in,
                                       atlas_rt_packed_ray* d_out, const uint32_t* d_count, uint32_t max_count,
                                       float tmin, float tmax, uint32_t mask, void* stream);

/* DispatchHitClosest with opacityCheck = true (RayTracingHelper.cpp:351-353):
 * as atlas_rt_trace_closest, but HitClosestTransparency (bvh.hsh:275-357)
 * drops candidates whose opacity is <= 0.  Opacity is the BLASes' baked
 * per-slot value (atlas_rt_blas_set_opacity); slots baked -1 (opacity map)
 * read as 1 here, since only the path tracer holds materials and textures.
 * Every BLAS of the world needs shading data (atlas_rt_blas_set_shading). */
atlas_rt_status atlas_rt_trace_closest_transparency(const atlas_rt_world* world, const atlas_rt_packed_ray* d_in,
                                                    atlas_rt_packed_ray* d_out, const uint32_t* d_count,
                                                    uint32_t max_count, float tmin, float tmax, uint32_t mask,
                                                    void* stream);

/* HitAnyTransparency (bvh.hsh:443-524) per PackedRay: d_transparency[i] =
 * the transparency left along the ray (1 = unoccluded, 0 = blocked), with
 * the engine's per-leaf squaring (bvh.hsh:496) and its 1e-6 cut-off; 1 for
 * a NaN direction.  Opacity as for atlas_rt_trace_closest_transparency. */
atlas_rt_status atlas_rt_trace_transparency(const atlas_rt_world* world, const atlas_rt_packed_ray* d_in,
                                            float* d_transparency, const uint32_t* d_count, uint32_t max_count,
                                            float tmin, float tmax, uint32_t mask, void* stream);

/* Any-hit traversal flavours.
 * ATLAS_RT_ANY_ENGINE_HITANY: HitAny (bvh.hsh:359-441) bit for bit, including
 *   that it restores the world-space ray only when returning from an entered
 *   instance (bvh.hsh:384-387): after an instance rejected by the mask, the
 *   TLAS walk continues with that instance's object-space ray.
 * ATLAS_RT_ANY_RESTORE_ALWAYS: HitAnyTransparency's traversal (bvh.hsh:466-470),
 *   which restores on every TLAS step; for opaque triangles its answer is
 *   "any accepted triangle in (tmin, tmax) among admitted instances".  The
 *   path tracer's shadow rays use it (rayHit.csh:333). */
#define ATLAS_RT_ANY_ENGINE_HITANY 0u
#define ATLAS_RT_ANY_RESTORE_ALWAYS 1u

/* Any hit per PackedRay; d_occluded[i] = 1 if occluded. */
atlas_rt_status atlas_rt_trace_any(const atlas_rt_world* world, const atlas_rt_packed_ray* d_in,
                                   uint32_t* d_occluded, const uint32_t* d_count, uint32_t max_count, float tmin,
                                   float tmax, uint32_t mask, uint32_t flags, void* stream);

/* Compact two-level traces (per-ray tmin/tmax): hits carry the caller's
 * triangle index in the hit BLAS, d_instance (may be NULL) the caller's
 * instance index, -1 on a miss.  Like BLAS traces, the traces of one world
 * share the world's stack spill and overflow flag: traces of the same world
 * on different streams must not overlap. */
atlas_rt_status atlas_rt_world_trace_closest_rays(const atlas_rt_world* world, const atlas_rt_ray* d_rays,
                                                  atlas_rt_hit* d_hits, int32_t* d_instance, uint32_t count,
                                                  uint32_t mask, void* stream);
atlas_rt_status atlas_rt_world_trace_any_rays(const atlas_rt_world* world, const atlas_rt_ray* d_rays,
                                              uint32_t* d_hit, uint32_t count, uint32_t mask, uint32_t flags,
                                              void* stream);

/* ------------------------------------------------------------------------- */
/* Path tracer (PathTracingRenderer + RayTracingHelper bounce loop)          */
/* ------------------------------------------------------------------------- */

/* Per-slot shading data of a BLAS: vertex normals (GPUTriangle n0..n2,
 * MeshData.cpp:222-237, kept in float32 here) and the material index
 * (GPUTriangle.d0.w).  normals: 9 floats per triangle in the caller's
 * original order; material: one index per triangle or NULL (all 0). */
atlas_rt_status atlas_rt_blas_set_shading(atlas_rt_blas* blas, const float* normals, const int32_t* material);

/* The engine's own shading records of a BLAS: count atlas_rt_gpu_triangle in
 * BVH slot order, as MeshData::BuildBVH leaves them in gpuTriangles and
 * Mesh::BuildBVH uploads them (MeshData.cpp:172-240, Mesh.cpp:97-108); count
 * must be the BLAS's slot count and each record's end-of-leaf flag (d1.z)
 * the slot's.  They are unpacked as UnpackTriangle does
 * (raytracer/common.hsh:17-44: unpackUnitVector, unpackHalf2x16,
 * unpackUnorm4x8; the material index and opacity as stored), and the path
 * tracer shades from the unpacked values: vertex normals, material index,
 * opacity and texture coordinates (the opacity-aware traversal reads these
 * two), the tangent frame of normal maps, and vertex colours for materials
 * with use_vertex_colors.  Replaces atlas_rt_blas_set_shading, _set_opacity
 * and _set_tangents for the BLAS. */
atlas_rt_status atlas_rt_blas_set_triangles(atlas_rt_blas* blas, const atlas_rt_gpu_triangle* tris, uint32_t count);

/* Per-triangle opacity and texture coordinates of a BLAS, used by the path
 * tracer's opacity-aware traversal (HitClosestTransparency /
 * HitAnyTransparency, raytracer/bvh.hsh:106-170,275-357,443-524).
 * opacity: one value per triangle in the caller's original order, baked as
 * MeshData::BuildBVH does (MeshData.cpp:141): the material's opacity, or -1
 * where the material has an opacity map (the traversal then samples
 * GetOpacity, raytracer/surface.hsh:147-160); NULL = all 1 (opaque).
 * uvs: 6 floats per triangle (uv0, uv1, uv2; GPUTriangle d1/d2 texture
 * coordinates, MeshData.cpp:222-237, kept in float32 here), or NULL (all 0).
 * Call after atlas_rt_blas_set_shading. */
atlas_rt_status atlas_rt_blas_set_opacity(atlas_rt_blas* blas, const float* opacity, const float* uvs);

/* Per-triangle tangent frame for normal maps (GPUTriangle's tangent and
 * bitangent, MeshData.cpp:222-237; transformed and normalised per instance as
 * TransformTriangle does, surface.hsh:46-61): 6 floats per triangle (t, bt) in
 * the caller's original order.  Without it both are zero.  Call after
 * atlas_rt_blas_set_shading. */
atlas_rt_status atlas_rt_blas_set_tangents(atlas_rt_blas* blas, const float* tangents);

/* RaytraceMaterial / GPUMaterial subset used by the bounce shader
 * (RTStructures.h:51-78, raytracer/structures.hsh:107-135, surface.hsh:63-140).
 * The *_texture fields index the path tracer's texture table
 * (atlas_rt_pathtracer_set_textures); -1, or any index outside the table,
 * means no map (the Sample*Bilinear functions of raytracer/texture.hsh
 * return 1).  Maps are sampled at the triangle's texture coordinate
 * (atlas_rt_blas_set_opacity's uvs) as GetSurfaceParameters does
 * (raytracer/surface.hsh:86-130): base colour .rgb, opacity / roughness /
 * metalness / AO .r multiply the material's values; the normal map's
 * 2 * rgb - 1 goes through (tangent, +-bitangent, normal)
 * (atlas_rt_blas_set_tangents) and is mixed in by saturate(normal_scale).
 * Vertex colours are not supported. */
typedef struct atlas_rt_material {
    float base_color[3];
    float opacity;
    float emissive[3];
    float roughness;
    float metalness;
    float ao;
    float reflectance;
    int32_t two_sided;
    int32_t opacity_texture;
    int32_t invert_uvs;
    int32_t id;           /* RaytraceMaterial.ID (hash % 65535, RayTracingWorld.cpp:194): the
                             realtime G-buffer's material index (16 bits kept) */
    float normal_scale;   /* normalScale: weight of the normal map */
    int32_t base_color_texture;  /* texture table indices, -1 or outside the table = none */
    int32_t roughness_texture;
    int32_t metalness_texture;
    int32_t ao_texture;
    int32_t normal_texture;
    int32_t use_vertex_colors;  /* GPUMaterial.useVertexColors (RTStructures.h:72): base colour times the
                                   interpolated vertex colour (surface.hsh:98); needs atlas_rt_blas_set_triangles */
    int32_t reserved[2];
} atlas_rt_material; /* 96 B */

/* A punctual light as the engine's LightComponent describes it
 * (RayTracingHelper::UpdateLights, RayTracingHelper.cpp:436-500): directional
 * only; `direction` is the direction the light travels (L = -direction),
 * radiance = linear colour * intensity. */
#define ATLAS_RT_LIGHT_DIRECTIONAL 0u /* DIRECTIONAL_LIGHT (raytracer/common.hsh:11) */
#define ATLAS_RT_LIGHT_TRIANGLE 1u    /* TRIANGLE_LIGHT (raytracer/common.hsh:12) */
typedef struct atlas_rt_light {
    float direction[3];
    uint32_t type;
    float radiance[3];
    float reserved;
} atlas_rt_light; /* 32 B */

/* GPULight / PackedLight (RTStructures.h:105-110, raytracer/structures.hsh,
 * unpacked by UnpackLight raytracer/common.hsh:100-121): the light list the
 * bounce shader reads.  data[0] = bits(type << 28 | BVH slot of a triangle
 * light), data[1] = normalised selection weight, data[2] = area,
 * data[3] = bits(instance slot, in the world's TLAS order). */
typedef struct atlas_rt_gpu_light {
    float P[4];
    float N[4];
    float color[4];
    float data[4];
} atlas_rt_gpu_light; /* 64 B */

/* A material texture: UNORM8 texels with `channels` channels, row-major,
 * width * height * channels bytes in host memory (copied).  Sampled
 * bilinearly with repeat addressing at level 0 (the engine's global sampler,
 * MainRenderer.cpp:922-929). */
typedef struct atlas_rt_texture {
    uint32_t width;
    uint32_t height;
    const uint8_t* texels;
    uint32_t channels;  /* 1 (or 0) .. 4 UNORM8 channels per texel; missing g, b read 0 and a reads 1 */
    uint32_t reserved;
} atlas_rt_texture;

/* RayGenUniforms (PathTracingRenderer.h:48-57): origin/right/bottom are the
 * near-plane top-left corner and the spans to the top-right and bottom-left
 * corners (PathTracingRenderer.cpp:157-162); location is the camera position. */
typedef struct atlas_rt_camera {
    float location[3];
    float origin[3];
    float right[3];
    float bottom[3];
} atlas_rt_camera;

/* atlas_rt_pt_params.flags: sort every bounce's rays by direction bin before
 * tracing (the engine's useRayBinning, RayTracingHelper.h:53, switched off
 * there; see atlas_rt_sort_rays_by_direction).  The image is unchanged. */
#define ATLAS_RT_PT_RAY_BINNING 0x1u
/* atlas_rt_pt_params.flags: run every bounce as separate closest-hit,
 * shade and shadow-ray launches over compacted path queues (the engine's
 * dispatch shape, RayTracingHelper.cpp:224-405) instead of the default
 * resident frame kernel, which walks each path from its camera ray to its
 * last bounce in one lane.  Both give the same image bit for bit.  The
 * radiance cache and ray binning always use the bounce-by-bounce launches. */
#define ATLAS_RT_PT_WAVEFRONT 0x2u
/* atlas_rt_pt_params.flags bits 8-11: progressive frames in flight (1-4; 0
 * picks 3).  No engine counterpart: the engine renders one frame per call.
 * Each frame in flight has its own stream and path buffers; the last
 * bounces of a frame trace few rays, so the next frames' first bounces keep
 * the GPU busy meanwhile.  Events order the frames' seeds, accumulation and
 * cache training, so the image is the same bit for bit at every setting. */
#define ATLAS_RT_PT_FRAMES_IN_FLIGHT_SHIFT 8
#define ATLAS_RT_PT_FRAMES_IN_FLIGHT_MASK 0xF00u
#define ATLAS_RT_PT_FRAMES_IN_FLIGHT(n) ((uint32_t)(n) << ATLAS_RT_PT_FRAMES_IN_FLIGHT_SHIFT)

typedef struct atlas_rt_pt_params {
    uint32_t width;
    uint32_t height;
    int32_t bounces;          /* maxBounces; bounces + 1 trace passes per frame */
    float exposure;
    float env_radiance[3];    /* constant environment (SampleEnvironmentMap) */
    uint32_t flags;           /* ATLAS_RT_PT_* */
} atlas_rt_pt_params;

typedef struct atlas_rt_pathtracer atlas_rt_pathtracer;

/* The world, its BLASes (with shading data) and the material table must
 * outlive the path tracer; materials are indexed by instance
 * material_offset + triangle material index. */
atlas_rt_status atlas_rt_pathtracer_create(const atlas_rt_world* world, const atlas_rt_pt_params* params,
                                           const atlas_rt_material* materials, uint32_t material_count,
                                           const atlas_rt_light* lights, uint32_t light_count,
                                           atlas_rt_pathtracer** out);
atlas_rt_status atlas_rt_pathtracer_destroy(atlas_rt_pathtracer* pt);

/* Renders `frames` progressive samples per pixel (one path per pixel per
 * frame: raygen, then bounces + 1 x (closest-hit trace, shade, shadow-ray
 * any-hit), then accumulate), asynchronously on `stream`.  The per-bounce
 * RNG seeds follow Random::SampleFastUniformFloat's PCG32 stream
 * (RandomHelper.cpp:8-51) and the sub-pixel jitter rayGen.csh:33-34.
 * Frames run on the path tracer's own streams (ATLAS_RT_PT_FRAMES_IN_FLIGHT),
 * after the work already queued on `stream` and after the frames of earlier
 * calls, and `stream` waits for them: work queued on `stream` afterwards
 * sees the image.  The resident frame kernel's grid shape (three or four
 * groups per CU) is chosen by timing: of a path tracer's first six plain
 * resident launches, launches two to five alternate the shapes, each after
 * the previous launch has ended, and the sixth waits for them once on the
 * host; the image does not depend on the shape. */
atlas_rt_status atlas_rt_pathtracer_render(atlas_rt_pathtracer* pt, const atlas_rt_camera* camera, uint32_t frames,
                                           void* stream);

/* Realtime mode (PathTracingRenderer with realTime, PathTracingRenderer.cpp:
 * 29-292): every frame traces samples_per_frame paths per pixel through the
 * pixel centre plus the camera jitter (rayGen.csh:39-40), writes a G-buffer at
 * the first bounce (rayHit.csh:76-117: velocity from this and last frame's
 * matrices, view-space depth, octahedral geometry normal, material ID), sums
 * the samples (float radiance for one sample, else the engine's quantised
 * sum, rayHit.csh:133-149) and resolves them against the history
 * (pathtracer/temporal.csh).  Defaults in PathTracingRenderer.h:38-43. */
typedef struct atlas_rt_realtime_params {
    int32_t samples_per_frame;   /* realTimeSamplesPerFrame (4) */
    int32_t history_length_max;  /* historyLengthMax (32) */
    float history_clip_max;      /* historyClipMax (0.2) */
    float current_clip_factor;   /* currentClipFactor (2.0) */
    float max_radiance;          /* maxRadiance (65535) */
} atlas_rt_realtime_params;

/* The frame's camera matrices as the engine's globalData holds them: glm
 * column-major mat4 (m[column][row]); jitter = jitterCurrent. */
typedef struct atlas_rt_frame_matrices {
    float view[4][4];
    float projection[4][4];
    float last_view_projection[4][4];
    float jitter[2];
    float reserved[2];
} atlas_rt_frame_matrices;

#define ATLAS_RT_RT_RESOLVED 0u  /* float4 per pixel: resolved radiance, history length (outAccumImage) */
#define ATLAS_RT_RT_FRAME 1u     /* float4 per pixel: this frame's radiance as temporal.csh fetches it */
#define ATLAS_RT_RT_DEPTH 2u     /* float per pixel: view-space z, +inf on a miss */
#define ATLAS_RT_RT_VELOCITY 3u  /* 2 x binary16 per pixel (R16G16_SFLOAT) */
#define ATLAS_RT_RT_NORMAL 4u    /* 2 x binary16 per pixel: EncodeNormal(geometry normal) */
#define ATLAS_RT_RT_MATERIAL 5u  /* uint16 per pixel: material ID, 65535 on a miss */

/* Switches realtime mode on (params) or off (NULL); resets the history.
 * Not with a radiance cache attached. */
atlas_rt_status atlas_rt_pathtracer_set_realtime(atlas_rt_pathtracer* pt, const atlas_rt_realtime_params* params);
/* One realtime frame on `stream`.  Instances' last-frame matrices are the
 * world's transforms before its latest atlas_rt_world_refit (the current ones
 * after an update). */
atlas_rt_status atlas_rt_pathtracer_render_realtime(atlas_rt_pathtracer* pt, const atlas_rt_camera* camera,
                                                    const atlas_rt_frame_matrices* matrices, void* stream);
/* Copies one of this frame's ATLAS_RT_RT_* buffers (width * height elements)
 * to device memory d_out. */
atlas_rt_status atlas_rt_pathtracer_read_realtime(atlas_rt_pathtracer* pt, uint32_t buffer, void* d_out,
                                                  void* stream);

/* Opacity textures referenced by atlas_rt_material.opacity_texture; replaces
 * the previous table (the scene's bindless textures). */
atlas_rt_status atlas_rt_pathtracer_set_textures(atlas_rt_pathtracer* pt, const atlas_rt_texture* textures,
                                                 uint32_t count);

/* PathTracingRenderer::sampleEmissives (PathTracingRenderer.h:46, passed to
 * UpdateLights at PathTracingRenderer.cpp:65): adds one triangle light per
 * (instance, BVH slot) whose emitter has a positive brightness, as
 * RayTracingWorld::BuildTriangleLightsForMesh / UpdateTriangleLights
 * (RayTracingWorld.cpp:322-407) build them, and renormalises the whole list
 * (UpdateLights).  emitters: 4 floats per entry of the material table (the
 * material's sRGB emissive colour and emissive intensity, which
 * GPUMaterial.emissiveColor = linear(colour) * intensity comes from); the
 * emitter of a triangle is its material index + the instance's
 * material_offset.  emitters = NULL or count = 0 turns triangle lights off.
 * Light order: directional lights, then BLAS by BLAS, slot by slot, the
 * instances of that BLAS in TLAS order. */
atlas_rt_status atlas_rt_pathtracer_set_emissive_lights(atlas_rt_pathtracer* pt, const float* emitters,
                                                        uint32_t count);

/* The light list in use (host copy).  *count receives its length; at most
 * `cap` entries are written to `out` (which may be NULL when cap = 0). */
atlas_rt_status atlas_rt_pathtracer_lights(const atlas_rt_pathtracer* pt, atlas_rt_gpu_light* out, uint32_t cap,
                                           uint32_t* count);

/* PathTracingRenderer::ResetSampleCount: clears the accumulator. */
atlas_rt_status atlas_rt_pathtracer_reset(atlas_rt_pathtracer* pt);

/* Accumulated radiance sum (rgb) and sample count (a) per pixel, row-major,
 * width*height float4; d_out is a device buffer.  Synchronous on `stream`. */
atlas_rt_status atlas_rt_pathtracer_read(atlas_rt_pathtracer* pt, float* d_out, void* stream);
atlas_rt_status atlas_rt_pathtracer_sample_count(const atlas_rt_pathtracer* pt, int32_t* count);
/* No engine counterpart: traversal work counters, the measure of a frame's
 * algorithmic bytes (64 B per inner node, TLAS or BLAS, 48 B per triangle
 * test, 64 B per instance record).  enable = 1 starts counting from zero,
 * 0 stops; frames run as before (same walks, same image) through a counting
 * instance of the resident frame kernel.  atlas_rt_pathtracer_work_counts
 * waits for every frame and writes counts[0..5]: closest-hit rays, shadow
 * rays, closest hits, inner nodes fetched, triangles tested and instances
 * entered since counting started; ATLAS_RT_E_UNSUPPORTED if some frame ran
 * bounce-by-bounce launches (ATLAS_RT_PT_WAVEFRONT, binning, the radiance
 * cache) or transparency walks, which do not count. */
atlas_rt_status atlas_rt_pathtracer_count_work(atlas_rt_pathtracer* pt, int32_t enable);
/* No engine counterpart (the engine records its frame into a command list
 * each frame, PathTracingRenderer.cpp:51-208): the kernels of `frames`
 * progressive frames (1 to 8) of `camera` captured once as a HIP graph.
 * atlas_rt_pathtracer_replay launches it on `stream`, after every frame of
 * earlier calls; each replay renders what atlas_rt_pathtracer_render(camera,
 * frames) would at that point (the seed stream and sample count continue),
 * the same image bit for bit, and later render calls order after it.  Plain
 * progressive path tracing only: realtime mode, the radiance cache, binning
 * and ATLAS_RT_PT_WAVEFRONT are refused.  A replay after anything that
 * changes what the kernels were given (a world update or stack limit,
 * materials, lights, textures, work counting) returns
 * ATLAS_RT_E_INVALID_ARG: capture again.  The graph clears its counters with
 * kernels, not memsets. */
atlas_rt_status atlas_rt_pathtracer_capture(atlas_rt_pathtracer* pt, const atlas_rt_camera* camera, uint32_t frames);
atlas_rt_status atlas_rt_pathtracer_replay(atlas_rt_pathtracer* pt, void* stream);
atlas_rt_status atlas_rt_pathtracer_work_counts(atlas_rt_pathtracer* pt, uint64_t* counts);
/* Waits for `stream` and every frame in flight, then reads and clears the
 * path tracer's overflow flags: ATLAS_RT_E_STACK_OVERFLOW if a frame since
 * the last call stopped a path at its world's stack limit
 * (atlas_rt_world_set_stack_limit; such a path ends as a miss). */
atlas_rt_status atlas_rt_pathtracer_status(atlas_rt_pathtracer* pt, void* stream);

/* ------------------------------------------------------------------------- */
/* Radiance cache (no engine counterpart)                                    */
/* ------------------------------------------------------------------------- */
/* An online-trained neural radiance cache: a multi-resolution hash-grid
 * encoding of position (up to 12 levels x 2 features, 2^log2_table entries
 * per level, FP16 table) concatenated with the outgoing direction, the
 * shading normal, the roughness and a constant 1 (32 inputs) feeding a
 * 5 x 64 ReLU MLP with 3 linear outputs (no biases).  Weights are BF16 on
 * the matrix cores with FP32 accumulation; FP32 master weights and tables
 * with Adam moments.
 * Input vector: [0, 2*levels) hash features (level-major), zeros up to 24,
 *   24..26 outgoing direction, 27..29 normal, 30 roughness, 31 = 1.
 * Queries: 12 floats each, three float4: (position xyz, roughness),
 *   (outgoing direction xyz, unused), (normal xyz, unused).
 * Targets and outputs: 4 floats each (rgb radiance, unused). */
typedef struct atlas_rt_nrc atlas_rt_nrc;

typedef struct atlas_rt_nrc_params {
    float bounds_min[3];   /* positions are normalised to this box */
    float bounds_max[3];
    uint32_t levels;       /* 1..12 */
    uint32_t log2_table;   /* 10..24, e.g. 19 */
    uint32_t base_resolution;   /* e.g. 16 */
    uint32_t max_resolution;    /* e.g. 2048; level l has base * b^l cells, b = (max/base)^(1/(levels-1)) */
    float learning_rate;   /* Adam step size, e.g. 1e-2 */
    uint32_t seed;         /* initialisation seed */
} atlas_rt_nrc_params;

atlas_rt_status atlas_rt_nrc_create(const atlas_rt_nrc_params* params, atlas_rt_nrc** out);
atlas_rt_status atlas_rt_nrc_destroy(atlas_rt_nrc* nrc);

/* Predicted rgb radiance (4 floats per query, 4th unused) on `stream`. */
atlas_rt_status atlas_rt_nrc_infer(atlas_rt_nrc* nrc, const float* d_queries, float* d_out, uint32_t count,
                                   void* stream);

/* One Adam step (beta1 0.9, beta2 0.99, eps 1e-15; table entries whose
 * gradient is exactly zero are left untouched) on the relative L2 loss
 * sum((y - t)^2 / (sg(y)^2 + 0.01)) / (3 count), sg = no gradient through
 * the denominator.  If d_loss != NULL the batch loss is written there. */
atlas_rt_status atlas_rt_nrc_train(atlas_rt_nrc* nrc, const float* d_queries, const float* d_targets,
                                   uint32_t count, float* d_loss, void* stream);

/* Gradients of the same loss without an update (for checking): d_grad_mlp
 * receives the FP32 weight gradients in the master layout (layer by layer,
 * K x N row-major), d_grad_table the FP32 table gradients (levels x 2^log2 x 2). */
atlas_rt_status atlas_rt_nrc_gradients(atlas_rt_nrc* nrc, const float* d_queries, const float* d_targets,
                                       uint32_t count, float* d_grad_mlp, float* d_grad_table, void* stream);

/* Parameter access: FP32 master MLP weights (same layout as the gradients)
 * and the FP16 table as uint16 bits.  Either pointer may be NULL.
 * Device pointers.  set_params also resets the Adam moments and step. */
atlas_rt_status atlas_rt_nrc_get_params(atlas_rt_nrc* nrc, float* d_mlp, uint16_t* d_table);
atlas_rt_status atlas_rt_nrc_set_params(atlas_rt_nrc* nrc, const float* d_mlp, const uint16_t* d_table);
/* Sizes in elements: MLP weights, table entries (levels x 2^log2 x 2). */
atlas_rt_status atlas_rt_nrc_sizes(const atlas_rt_nrc* nrc, uint32_t* mlp_params, uint32_t* table_params);

/* Path tracing with the cache.  With a cache attached, every camera path is
 * traced to its second surface vertex and ended there with the cache's
 * prediction of the radiance leaving that vertex towards the first (clamped
 * at 0, times the path throughput).  About one pixel in train_stride per
 * frame (a different random set each frame) instead traces the full
 * `bounces` path as in plain mode; every surface vertex of such a path
 * becomes a training sample: the radiance the rest of the path brought back
 * to the pixel divided by the throughput that reached the vertex.  After the
 * frame's bounces the cache takes train_steps Adam steps on those samples
 * (train_steps = 0 freezes it).  cache = NULL detaches it: plain path
 * tracing, the fallback.  Needs bounces >= 2.  The cache must outlive its
 * use by the path tracer.
 * query_lag 0: a frame queries the network as the previous frame's training
 * left it, so each frame's queries wait for that training.  query_lag 1:
 * frame i of a render call queries the network as frame i - 2's training
 * left it (frames 0 and 1 of the call: as the call found it), so a frame's
 * training overlaps the next frame's bounces when frames are in flight.
 * Both give the same image at every ATLAS_RT_PT_FRAMES_IN_FLIGHT setting. */
typedef struct atlas_rt_pt_cache_params {
    uint32_t train_stride;   /* e.g. 32 */
    uint32_t train_steps;    /* Adam steps per frame, e.g. 4 */
    uint32_t query_lag;      /* 0 or 1 */
    uint32_t reserved;       /* 0 */
} atlas_rt_pt_cache_params;

atlas_rt_status atlas_rt_pathtracer_set_cache(atlas_rt_pathtracer* pt, atlas_rt_nrc* cache,
                                              const atlas_rt_pt_cache_params* params);

/* Counts of the last rendered frame (synchronous): training samples built,
 * cache queries issued, and the loss of its last training step. */
atlas_rt_status atlas_rt_pathtracer_cache_stats(atlas_rt_pathtracer* pt, uint32_t* train_samples, uint32_t* queries,
                                                float* loss);

/* The last frame's training samples (synchronous): up to max_count queries
 * (12 floats each) and targets (4 floats each) into device buffers, in no
 * particular order; *count receives how many were copied. */
atlas_rt_status atlas_rt_pathtracer_cache_samples(atlas_rt_pathtracer* pt, float* d_queries, float* d_targets,
                                                  uint32_t max_count, uint32_t* count);

#ifdef __cplusplus
}
#endif

#endif /* ATLAS_RT_H */
