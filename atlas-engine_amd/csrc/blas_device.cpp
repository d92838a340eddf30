// BLAS objects built on the GPU (lbvh.hip): device arrays written by the
// builder, host mirrors copied back for export, world bounds, shading and
// the wide-node collapse.
#include <algorithm>
#include <cstring>

#include "lbvh.h"
#include "rt_internal.h"
#include "wide.h"

using namespace atlas_rt;

namespace {

void free_partial(std::vector<atlas_rt_blas*>& made) {
    for (atlas_rt_blas* b : made) atlas_rt_blas_destroy(b);
    made.clear();
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
        std::vector<const float*> v(verts + first, verts + first + n);
        std::vector<uint32_t> c(counts + first, counts + first + n);
        std::vector<LbvhMeshOut> o(n);
        std::vector<uint32_t*> orig(n, nullptr);
        atlas_rt_status st = ATLAS_RT_OK;
        for (uint32_t m = 0; m < n && st == ATLAS_RT_OK; m++) {
            auto* b = new atlas_rt_blas();
            made.push_back(b);
            b->device = dev;
            b->on_device = true;
            const size_t nt = c[m], nn = nt - 1;
            hipError_t e = hipSuccess;
            if (nn) e = hipMalloc(reinterpret_cast<void**>(&b->d_nodes), nn * sizeof(atlas_rt_gpu_bvh_node));
            if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void**>(&b->d_tris), nt * 3 * sizeof(float4));
            if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void**>(&b->d_vtx), nt * 3 * sizeof(float4));
            if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void**>(&orig[m]), nt * sizeof(uint32_t));
            if (e != hipSuccess) st = hip_fail(e, "atlas_rt_blas_build_device (outputs)");
            o[m] = LbvhMeshOut{b->d_nodes, b->d_tris, b->d_vtx, orig[m]};
        }
        std::vector<uint32_t> depth;
        float ms = 0.0f;
        if (st == ATLAS_RT_OK) st = lbvh_build(v, c, o, depth, s, device_ms ? &ms : nullptr);
        total_ms += ms;
        // host mirrors
        for (uint32_t m = 0; m < n && st == ATLAS_RT_OK; m++) {
            atlas_rt_blas* b = made[first + m];
            const size_t nt = c[m], nn = nt - 1;
            b->nodes.resize(nn);
            b->tris.resize(nt);
            b->orig_idx.resize(nt);
            hipError_t e = hipSuccess;
            if (nn) e = hipMemcpy(b->nodes.data(), b->d_nodes, nn * sizeof(atlas_rt_gpu_bvh_node), hipMemcpyDeviceToHost);
            if (e == hipSuccess)
                e = hipMemcpy(b->tris.data(), b->d_vtx, nt * sizeof(atlas_rt_gpu_bvh_triangle), hipMemcpyDeviceToHost);
            if (e == hipSuccess) e = hipMemcpy(b->orig_idx.data(), orig[m], nt * sizeof(uint32_t), hipMemcpyDeviceToHost);
            if (e != hipSuccess) st = hip_fail(e, "atlas_rt_blas_build_device (host mirror)");
            b->max_depth = depth[m];
            if (st == ATLAS_RT_OK) st = upload_wide(*b);
        }
        for (uint32_t* p : orig)
            if (p) (void)hipFree(p);
        if (st != ATLAS_RT_OK) {
            free_partial(made);
            return st;
        }
    }
    for (uint32_t k = 0; k < count; k++) out[k] = made[k];
    if (device_ms) *device_ms = total_ms;
    return ATLAS_RT_OK;
}

}  // extern "C"
