// BLAS objects: host build (builder.cpp), packing into the engine layout
// (MeshData::BuildBVH, src/engine/mesh/MeshData.cpp:172-269) and device upload
// (Mesh::BuildBVH, src/engine/mesh/Mesh.cpp:97-108), plus the small runtime
// helpers shared by every handle (errors, scratch, timing events).
#include <cmath>
#include <cstring>
#include <thread>

#include "builder.h"
#include "rt_hostmath.h"
#include "packing.h"
#include "rt_internal.h"
#include "wide.h"

namespace atlas_rt {

static thread_local std::string g_last_error;

void set_error(const std::string& msg) { g_last_error = msg; }
const char* last_error() { return g_last_error.c_str(); }

atlas_rt_status fail(atlas_rt_status code, const std::string& msg) {
    set_error(msg);
    return code;
}

atlas_rt_status hip_fail(hipError_t e, const char* what) {
    set_error(std::string(what) + ": " + hipGetErrorString(e));
    if (e == hipErrorOutOfMemory) return ATLAS_RT_E_OUT_OF_MEMORY;
    if (e == hipErrorNoDevice || e == hipErrorInvalidDevice) return ATLAS_RT_E_NO_DEVICE;
    return ATLAS_RT_E_HIP;
}

atlas_rt_status Scratch::ensure(size_t need) {
    if (need <= bytes && ptr) return ATLAS_RT_OK;
    release();
    size_t sz = need < 256 ? 256 : need;
    hipError_t e = hipMalloc(&ptr, sz);
    if (e != hipSuccess) {
        ptr = nullptr;
        return hip_fail(e, "hipMalloc(scratch)");
    }
    bytes = sz;
    hipGetDevice(&device);
    return ATLAS_RT_OK;
}

void Scratch::release() {
    if (ptr) hipFree(ptr);
    ptr = nullptr;
    bytes = 0;
}

atlas_rt_status LaunchTiming::begin(hipStream_t s) {
    if (!start) {
        ATLAS_HIP_TRY(hipEventCreate(&start));
        ATLAS_HIP_TRY(hipEventCreate(&stop));
    }
    ATLAS_HIP_TRY(hipEventRecord(start, s));
    return ATLAS_RT_OK;
}

atlas_rt_status LaunchTiming::end(hipStream_t s) {
    ATLAS_HIP_TRY(hipEventRecord(stop, s));
    pending = true;
    return ATLAS_RT_OK;
}

atlas_rt_status LaunchTiming::read(float* ms) {
    if (!pending) {
        *ms = last_ms;
        return ATLAS_RT_OK;
    }
    ATLAS_HIP_TRY(hipEventSynchronize(stop));
    ATLAS_HIP_TRY(hipEventElapsedTime(&last_ms, start, stop));
    pending = false;
    *ms = last_ms;
    return ATLAS_RT_OK;
}

void LaunchTiming::release() {
    if (start) hipEventDestroy(start);
    if (stop) hipEventDestroy(stop);
    start = stop = nullptr;
}

atlas_rt_status upload_blas(atlas_rt_blas& b) {
    int dev = 0;
    ATLAS_HIP_TRY(hipGetDevice(&dev));
    const size_t nn = b.nodes.size(), nt = b.tris.size();
    // Kernel triangle layout: (v0, index bits), (v1 - v0, end flag), (v2 - v0, 0).
    std::vector<float> packed(12 * nt);
    for (size_t i = 0; i < nt; i++) {
        const atlas_rt_gpu_bvh_triangle& t = b.tris[i];
        float* o = &packed[12 * i];
        const uint32_t idx = b.orig_idx.empty() ? uint32_t(i) : b.orig_idx[i];
        o[0] = t.v0[0];
        o[1] = t.v0[1];
        o[2] = t.v0[2];
        std::memcpy(&o[3], &idx, 4);
        o[4] = t.v1[0] - t.v0[0];
        o[5] = t.v1[1] - t.v0[1];
        o[6] = t.v1[2] - t.v0[2];
        o[7] = t.v0[3] > 0.0f ? 1.0f : -1.0f;
        o[8] = t.v2[0] - t.v0[0];
        o[9] = t.v2[1] - t.v0[1];
        o[10] = t.v2[2] - t.v0[2];
        o[11] = 0.0f;
    }
    void* dn = nullptr;
    void* dt = nullptr;
    void* dv = nullptr;
    hipError_t e = hipSuccess;
    if (nn) e = hipMalloc(&dn, nn * sizeof(atlas_rt_gpu_bvh_node));
    if (e == hipSuccess) e = hipMalloc(&dt, packed.size() * sizeof(float));
    if (e == hipSuccess) e = hipMalloc(&dv, nt * sizeof(atlas_rt_gpu_bvh_triangle));
    if (e == hipSuccess && nn)
        e = hipMemcpy(dn, b.nodes.data(), nn * sizeof(atlas_rt_gpu_bvh_node), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(dt, packed.data(), packed.size() * sizeof(float), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(dv, b.tris.data(), nt * sizeof(atlas_rt_gpu_bvh_triangle), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        if (dn) hipFree(dn);
        if (dt) hipFree(dt);
        if (dv) hipFree(dv);
        return hip_fail(e, "upload_blas");
    }
    b.d_nodes = reinterpret_cast<float4*>(dn);
    b.d_tris = reinterpret_cast<float4*>(dt);
    b.d_vtx = reinterpret_cast<float4*>(dv);
    b.device = dev;
    b.on_device = true;
    return upload_wide(b);
}

// Wide nodes for the 4- and 8-wide kernels, collapsed from the host copy.
atlas_rt_status upload_wide(atlas_rt_blas& b) {
    const size_t nt = b.tris.size();
    hipError_t e = hipSuccess;
    std::vector<uint8_t> end_flags(nt);
    for (size_t i = 0; i < nt; i++) end_flags[i] = b.tris[i].v0[3] > 0.0f ? 1 : 0;
    for (int width : {4, 8}) {
        const WideTree w = collapse_wide(b.nodes, end_flags, width, 4);
        void* dw = nullptr;
        if (w.node_count) {
            e = hipMalloc(&dw, w.data.size() * sizeof(float));
            if (e == hipSuccess) e = hipMemcpy(dw, w.data.data(), w.data.size() * sizeof(float), hipMemcpyHostToDevice);
            if (e != hipSuccess) {
                if (dw) hipFree(dw);
                return hip_fail(e, "upload_blas (wide nodes)");
            }
        }
        (width == 4 ? b.d_nodes4 : b.d_nodes8) = reinterpret_cast<float4*>(dw);
        (width == 4 ? b.cap4 : b.cap8) = w.max_stack + 1;
    }
    return ATLAS_RT_OK;
}

int default_threads() {
    unsigned h = std::thread::hardware_concurrency();
    if (h == 0) h = 1;
    if (h > 16) h = 16;
    return int(h);
}

}  // namespace atlas_rt

using namespace atlas_rt;

static void pack_triangles(atlas_rt_blas& b, const atlas_rt_bvh_triangle* tris, const BuildOutput& out) {
    const size_t n = out.ref_idx.size();
    b.tris.resize(n);
    b.orig_idx.resize(n);
    for (size_t i = 0; i < n; i++) {
        const atlas_rt_bvh_triangle& s = tris[out.ref_idx[i]];
        atlas_rt_gpu_bvh_triangle& g = b.tris[i];
        std::memcpy(g.v0, s.v0, 12);
        std::memcpy(g.v1, s.v1, 12);
        std::memcpy(g.v2, s.v2, 12);
        g.v0[3] = out.end_of_node[i] ? 1.0f : -1.0f;  // MeshData.cpp:243
        const int32_t material = 0;
        std::memcpy(&g.v1[3], &material, 4);          // MeshData.cpp:244
        g.v2[3] = 1.0f;                                // opacity, MeshData.cpp:245
        b.orig_idx[i] = s.idx;
    }
}

extern "C" {

atlas_rt_status atlas_rt_blas_build(const atlas_rt_aabb* aabbs, const atlas_rt_bvh_triangle* tris, uint32_t count,
                                    uint32_t flags, atlas_rt_blas** out) {
    if (!out) return fail(ATLAS_RT_E_INVALID_ARG, "atlas_rt_blas_build: out is NULL");
    *out = nullptr;
    if (!tris) return fail(ATLAS_RT_E_INVALID_ARG, "atlas_rt_blas_build: tris is NULL");
    if (count == 0) return fail(ATLAS_RT_E_EMPTY, "atlas_rt_blas_build: no triangles");
    if (flags & ~(ATLAS_RT_BUILD_PARALLEL | ATLAS_RT_BUILD_HOST_ONLY))
        return fail(ATLAS_RT_E_INVALID_ARG, "atlas_rt_blas_build: unknown flag bits");
    std::vector<float> verts(9 * size_t(count));
    std::vector<float> boxes(6 * size_t(count));
    for (uint32_t i = 0; i < count; i++) {
        float* v = &verts[9 * size_t(i)];
        std::memcpy(v, tris[i].v0, 12);
        std::memcpy(v + 3, tris[i].v1, 12);
        std::memcpy(v + 6, tris[i].v2, 12);
        float* bx = &boxes[6 * size_t(i)];
        if (aabbs) {
            std::memcpy(bx, aabbs[i].min, 12);
            std::memcpy(bx + 3, aabbs[i].max, 12);
        } else {
            // glm::min(glm::min(v0, v1), v2) / max (MeshData.cpp:143-146).
            for (int k = 0; k < 3; k++) {
                bx[k] = gl_min(gl_min(v[k], v[3 + k]), v[6 + k]);
                bx[3 + k] = gl_max(gl_max(v[k], v[3 + k]), v[6 + k]);
            }
        }
    }
    BuildOutput bo;
    const int threads = (flags & ATLAS_RT_BUILD_PARALLEL) ? default_threads() : 1;
    if (!build_triangle_bvh(boxes.data(), verts.data(), count, threads, bo))
        return fail(ATLAS_RT_E_EMPTY, "atlas_rt_blas_build: build failed");
    auto* b = new atlas_rt_blas();
    b->nodes = std::move(bo.nodes);
    pack_triangles(*b, tris, bo);
    b->max_depth = bo.max_depth;
    if (!(flags & ATLAS_RT_BUILD_HOST_ONLY)) {
        atlas_rt_status st = upload_blas(*b);
        if (st != ATLAS_RT_OK) {
            delete b;
            return st;
        }
    }
    *out = b;
    return ATLAS_RT_OK;
}

atlas_rt_status atlas_rt_blas_create(const atlas_rt_gpu_bvh_node* nodes, uint32_t node_count,
                                     const atlas_rt_gpu_bvh_triangle* tris, const uint32_t* orig_idx,
                                     uint32_t tri_count, atlas_rt_blas** out) {
    if (!out) return fail(ATLAS_RT_E_INVALID_ARG, "atlas_rt_blas_create: out is NULL");
    *out = nullptr;
    if (!tris || (node_count && !nodes)) return fail(ATLAS_RT_E_INVALID_ARG, "atlas_rt_blas_create: NULL array");
    if (tri_count == 0) return fail(ATLAS_RT_E_EMPTY, "atlas_rt_blas_create: no triangles");
    std::vector<uint8_t> eon(tri_count);
    for (uint32_t i = 0; i < tri_count; i++) eon[i] = tris[i].v0[3] > 0.0f ? 1 : 0;
    uint32_t depth = 0;
    if (!validate_tree(nodes, node_count, tri_count, eon.data(), &depth))
        return fail(ATLAS_RT_E_INVALID_ARG,
                    "atlas_rt_blas_create: malformed tree (pointer out of range, not pre-order, or leaf without end flag)");
    if (node_count == 0) {
        uint32_t j = 0;
        while (j < tri_count && !eon[j]) j++;
        if (j >= tri_count) return fail(ATLAS_RT_E_INVALID_ARG, "atlas_rt_blas_create: root leaf without end flag");
    }
    auto* b = new atlas_rt_blas();
    b->nodes.assign(nodes, nodes + node_count);
    b->tris.assign(tris, tris + tri_count);
    if (orig_idx) b->orig_idx.assign(orig_idx, orig_idx + tri_count);
    else {
        b->orig_idx.resize(tri_count);
        for (uint32_t i = 0; i < tri_count; i++) b->orig_idx[i] = i;
    }
    b->max_depth = depth;
    atlas_rt_status st = upload_blas(*b);
    if (st != ATLAS_RT_OK) {
        delete b;
        return st;
    }
    *out = b;
    return ATLAS_RT_OK;
}

atlas_rt_status atlas_rt_blas_destroy(atlas_rt_blas* b) {
    if (!b) return ATLAS_RT_OK;
    if (!b->arena) {
        if (b->d_nodes) hipFree(b->d_nodes);
        if (b->d_tris) hipFree(b->d_tris);
        if (b->d_vtx) hipFree(b->d_vtx);
        if (b->d_nodes4) hipFree(b->d_nodes4);
        if (b->d_nodes8) hipFree(b->d_nodes8);
    }
    if (b->d_shade) hipFree(b->d_shade);
    if (b->d_uv) hipFree(b->d_uv);
    if (b->d_tangent) hipFree(b->d_tangent);
    if (b->d_color) hipFree(b->d_color);
    b->spill.release();
    b->flags.release();
    b->timing.release();
    delete b;
    return ATLAS_RT_OK;
}

atlas_rt_status atlas_rt_blas_set_shading(atlas_rt_blas* b, const float* normals, const int32_t* material) {
    if (!b || !normals) return fail(ATLAS_RT_E_INVALID_ARG, "atlas_rt_blas_set_shading: NULL argument");
    if (!b->on_device) return fail(ATLAS_RT_E_INVALID_ARG, "atlas_rt_blas_set_shading: BLAS was built host-only");
    // Reorder to BVH slot order (slot i holds caller triangle orig_idx[i]).
    const size_t n = b->tris.size();
    std::vector<float> packed(12 * n);
    std::vector<float> host_n(9 * n);
    std::vector<int32_t> host_m(n);
    for (size_t i = 0; i < n; i++) {
        const uint32_t src = b->orig_idx[i];
        const float* nv = normals + 9 * size_t(src);
        const int32_t mat = material ? material[src] : 0;
        std::memcpy(&host_n[9 * i], nv, 9 * sizeof(float));
        host_m[i] = mat;
        float* o = &packed[12 * i];
        o[0] = nv[0];
        o[1] = nv[1];
        o[2] = nv[2];
        std::memcpy(&o[3], &mat, 4);
        o[4] = nv[3];
        o[5] = nv[4];
        o[6] = nv[5];
        o[7] = 1.0f;  // baked opacity (atlas_rt_blas_set_opacity)
        o[8] = nv[6];
        o[9] = nv[7];
        o[10] = nv[8];
        o[11] = 0.0f;
    }
    void* d = nullptr;
    ATLAS_HIP_TRY(hipMalloc(&d, packed.size() * sizeof(float)));
    hipError_t e = hipMemcpy(d, packed.data(), packed.size() * sizeof(float), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        hipFree(d);
        return hip_fail(e, "atlas_rt_blas_set_shading");
    }
    if (b->d_shade) hipFree(b->d_shade);
    if (b->d_uv) hipFree(b->d_uv);
    if (b->d_color) hipFree(b->d_color);
    b->d_shade = reinterpret_cast<float4*>(d);
    b->d_uv = nullptr;
    b->d_color = nullptr;
    b->has_alpha = false;
    b->shade_normals.swap(host_n);
    b->shade_mat.swap(host_m);
    return ATLAS_RT_OK;
}

atlas_rt_status atlas_rt_blas_set_triangles(atlas_rt_blas* b, const atlas_rt_gpu_triangle* tris, uint32_t count) {
    if (!b || !tris) return fail(ATLAS_RT_E_INVALID_ARG, "atlas_rt_blas_set_triangles: NULL argument");
    if (!b->on_device) return fail(ATLAS_RT_E_INVALID_ARG, "atlas_rt_blas_set_triangles: BLAS was built host-only");
    const size_t n = b->tris.size();
    if (count != n)
        return fail(ATLAS_RT_E_SIZE_MISMATCH, "atlas_rt_blas_set_triangles: " + std::to_string(count) +
                                                  " records for " + std::to_string(n) + " BVH slots");
    // Per slot: shading (n0 + material bits, n1 + opacity, n2), texture
    // coordinates (uv0 uv1, uv2), tangent frame (t, bt) and vertex colours
    // (three RGBA), the layouts the kernels read.
    std::vector<float> shade(12 * n), uv(8 * n, 0.0f), tan(8 * n, 0.0f), col(12 * n);
    std::vector<float> host_n(9 * n);
    std::vector<int32_t> host_m(n);
    bool alpha = false;
    for (size_t i = 0; i < n; i++) {
        const atlas_rt_gpu_triangle& r = tris[i];
        if ((r.d1[2] > 0.0f) != (b->tris[i].v0[3] > 0.0f))
            return fail(ATLAS_RT_E_INVALID_ARG, "atlas_rt_blas_set_triangles: record " + std::to_string(i) +
                                                    "'s end-of-leaf flag is not its BVH slot's (records must be in "
                                                    "BVH slot order)");
        const float* v[3] = {r.v0, r.v1, r.v2};
        const float* bv[3] = {b->tris[i].v0, b->tris[i].v1, b->tris[i].v2};
        for (int k = 0; k < 3; k++)
            if (std::memcmp(v[k], bv[k], 12) != 0)
                return fail(ATLAS_RT_E_INVALID_ARG, "atlas_rt_blas_set_triangles: record " + std::to_string(i) +
                                                        "'s vertices are not its BVH slot's triangle");
        float* o = &shade[12 * i];
        for (int k = 0; k < 3; k++) {
            float nk[3];
            unpack_unit_vector(float_bits(v[k][3]), nk);
            std::memcpy(&o[4 * k], nk, 12);
            std::memcpy(&host_n[9 * i + 3 * k], nk, 12);
        }
        const int32_t mat = int32_t(float_bits(r.d0[3]));
        std::memcpy(&o[3], &mat, 4);
        o[7] = r.d2[3];
        o[11] = 0.0f;
        host_m[i] = mat;
        alpha = alpha || r.d2[3] != 1.0f;
        for (int k = 0; k < 3; k++) unpack_half2(float_bits(r.d0[k]), &uv[8 * i + 2 * k]);
        unpack_unit_vector(float_bits(r.d1[0]), &tan[8 * i]);
        unpack_unit_vector(float_bits(r.d1[1]), &tan[8 * i + 4]);
        for (int k = 0; k < 3; k++) unpack_unorm4x8(float_bits(r.d2[k]), &col[12 * i + 4 * k]);
    }
    float4* dev[4] = {nullptr, nullptr, nullptr, nullptr};
    const std::vector<float>* src[4] = {&shade, &uv, &tan, &col};
    for (int a = 0; a < 4; a++) {
        hipError_t e = hipMalloc(reinterpret_cast<void**>(&dev[a]), src[a]->size() * sizeof(float));
        if (e == hipSuccess) e = hipMemcpy(dev[a], src[a]->data(), src[a]->size() * sizeof(float), hipMemcpyHostToDevice);
        if (e != hipSuccess) {
            for (float4* d : dev)
                if (d) hipFree(d);
            return hip_fail(e, "atlas_rt_blas_set_triangles");
        }
    }
    float4** mine[4] = {&b->d_shade, &b->d_uv, &b->d_tangent, &b->d_color};
    for (int a = 0; a < 4; a++) {
        if (*mine[a]) hipFree(*mine[a]);
        *mine[a] = dev[a];
    }
    b->has_alpha = alpha;
    b->shade_normals.swap(host_n);
    b->shade_mat.swap(host_m);
    return ATLAS_RT_OK;
}

atlas_rt_status atlas_rt_blas_set_tangents(atlas_rt_blas* b, const float* tangents) {
    if (!b || !tangents) return fail(ATLAS_RT_E_INVALID_ARG, "atlas_rt_blas_set_tangents: NULL argument");
    if (!b->d_shade)
        return fail(ATLAS_RT_E_INVALID_ARG, "atlas_rt_blas_set_tangents: call atlas_rt_blas_set_shading first");
    const size_t n = b->tris.size();
    std::vector<float> packed(8 * n, 0.0f);
    for (size_t i = 0; i < n; i++) {
        const float* t = tangents + 6 * size_t(b->orig_idx[i]);
        std::memcpy(&packed[8 * i], t, 12);
        std::memcpy(&packed[8 * i + 4], t + 3, 12);
    }
    void* d = nullptr;
    ATLAS_HIP_TRY(hipMalloc(&d, packed.size() * sizeof(float)));
    const hipError_t e = hipMemcpy(d, packed.data(), packed.size() * sizeof(float), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        hipFree(d);
        return hip_fail(e, "atlas_rt_blas_set_tangents");
    }
    if (b->d_tangent) hipFree(b->d_tangent);
    b->d_tangent = reinterpret_cast<float4*>(d);
    return ATLAS_RT_OK;
}

atlas_rt_status atlas_rt_blas_set_opacity(atlas_rt_blas* b, const float* opacity, const float* uvs) {
    if (!b) return fail(ATLAS_RT_E_INVALID_ARG, "atlas_rt_blas_set_opacity: NULL handle");
    if (!b->d_shade)
        return fail(ATLAS_RT_E_INVALID_ARG, "atlas_rt_blas_set_opacity: call atlas_rt_blas_set_shading first");
    const size_t n = b->tris.size();
    // baked opacity goes to the w lane of each slot's second shading float4
    std::vector<float> op(n, 1.0f);
    bool alpha = false;
    for (size_t i = 0; i < n; i++) {
        if (opacity) op[i] = opacity[b->orig_idx[i]];
        alpha = alpha || op[i] != 1.0f;
    }
    float4* d_uv = nullptr;
    if (uvs) {
        std::vector<float> packed(8 * n, 0.0f);
        for (size_t i = 0; i < n; i++) std::memcpy(&packed[8 * i], uvs + 6 * size_t(b->orig_idx[i]), 6 * sizeof(float));
        ATLAS_HIP_TRY(hipMalloc(reinterpret_cast<void**>(&d_uv), packed.size() * sizeof(float)));
        hipError_t e = hipMemcpy(d_uv, packed.data(), packed.size() * sizeof(float), hipMemcpyHostToDevice);
        if (e != hipSuccess) {
            hipFree(d_uv);
            return hip_fail(e, "atlas_rt_blas_set_opacity");
        }
    }
    hipError_t e = hipMemcpy2D(reinterpret_cast<char*>(b->d_shade) + 7 * sizeof(float), 12 * sizeof(float), op.data(),
                               sizeof(float), sizeof(float), n, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        if (d_uv) hipFree(d_uv);
        return hip_fail(e, "atlas_rt_blas_set_opacity");
    }
    if (b->d_uv) hipFree(b->d_uv);
    b->d_uv = d_uv;
    b->has_alpha = alpha;
    return ATLAS_RT_OK;
}

atlas_rt_status atlas_rt_blas_info(const atlas_rt_blas* b, uint32_t* node_count, uint32_t* tri_count,
                                   uint32_t* max_depth) {
    if (!b) return fail(ATLAS_RT_E_INVALID_ARG, "atlas_rt_blas_info: NULL handle");
    if (node_count) *node_count = uint32_t(b->nodes.size());
    if (tri_count) *tri_count = uint32_t(b->tris.size());
    if (max_depth) *max_depth = b->max_depth;
    return ATLAS_RT_OK;
}

atlas_rt_status atlas_rt_blas_export(const atlas_rt_blas* b, atlas_rt_gpu_bvh_node* nodes,
                                     atlas_rt_gpu_bvh_triangle* tris, uint32_t* orig_idx) {
    if (!b) return fail(ATLAS_RT_E_INVALID_ARG, "atlas_rt_blas_export: NULL handle");
    if (nodes && !b->nodes.empty()) std::memcpy(nodes, b->nodes.data(), b->nodes.size() * sizeof(*nodes));
    if (tris) std::memcpy(tris, b->tris.data(), b->tris.size() * sizeof(*tris));
    if (orig_idx) std::memcpy(orig_idx, b->orig_idx.data(), b->orig_idx.size() * sizeof(uint32_t));
    return ATLAS_RT_OK;
}

}  // extern "C"
