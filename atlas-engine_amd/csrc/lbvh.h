// GPU BVH builder: linear BVH (Morton order, Karras radix tree) emitted in
// the engine's BLAS layout, so the traversal kernels run on it unchanged.
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

#include "rt_internal.h"

namespace atlas_rt {

// One mesh's device outputs (allocated by the caller, sizes from `count`):
// nodes: count - 1 engine GPUBVHNode records (64 B); tris: 3 float4 per slot
// in the traversal layout; vtx: engine GPUBVHTriangle per slot; orig: caller
// triangle index per slot; depth: the deepest leaf's number of inner-node
// ancestors (1 for a single triangle).
struct LbvhMeshOut {
    float4* nodes;
    float4* tris;
    float4* vtx;
    uint32_t* orig;
};

// Builds every mesh of the batch with one set of launches on `stream`.
// verts: device memory, 9 floats per triangle (v0, v1, v2), the meshes one
// after another in batch order.
// Per mesh at most 2^24 triangles; at most 1024 meshes per call.
// depth_out: per mesh (host), filled after the stream has been synchronised.
atlas_rt_status lbvh_build(const float* verts, const std::vector<uint32_t>& counts,
                           const std::vector<LbvhMeshOut>& outs, std::vector<uint32_t>& depth_out,
                           hipStream_t stream, float* device_ms);

}  // namespace atlas_rt
