// extern "C" trace entry points over one BLAS, and library-level queries.
#include <cstring>

#include "binning.h"
#include "rt_internal.h"

namespace atlas_rt {
const char* last_error();
}

using namespace atlas_rt;

atlas_rt_status atlas_rt::take_overflow_flag(void* flags, void* stream, const char* who) {
    if (!flags) return ATLAS_RT_OK;
    ATLAS_HIP_TRY(hipStreamSynchronize(reinterpret_cast<hipStream_t>(stream)));
    uint32_t* word = reinterpret_cast<uint32_t*>(flags) + 2;
    uint32_t v = 0;
    ATLAS_HIP_TRY(hipMemcpy(&v, word, sizeof(v), hipMemcpyDeviceToHost));
    if (!v) return ATLAS_RT_OK;
    const uint32_t zero = 0;
    ATLAS_HIP_TRY(hipMemcpy(word, &zero, sizeof(zero), hipMemcpyHostToDevice));
    return fail(ATLAS_RT_E_STACK_OVERFLOW, std::string(who) +
                                               ": a ray needed more traversal stack entries than the limit; it was "
                                               "reported as a miss");
}

static atlas_rt_status check_blas_ready(const atlas_rt_blas* b, const char* who) {
    if (!b) return fail(ATLAS_RT_E_INVALID_ARG, std::string(who) + ": NULL BLAS handle");
    if (!b->on_device) return fail(ATLAS_RT_E_INVALID_ARG, std::string(who) + ": BLAS was built host-only");
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev != b->device)
        return fail(ATLAS_RT_E_INVALID_ARG, std::string(who) + ": BLAS lives on another device");
    return ATLAS_RT_OK;
}

static atlas_rt_status check_kernel(int32_t kernel, const char* who) {
    const int32_t base = kernel & ATLAS_RT_KERNEL_VARIANT_MASK;
    const int32_t extra = kernel & ~int32_t(ATLAS_RT_KERNEL_VARIANT_MASK | ATLAS_RT_KERNEL_STACK_REGISTER |
                                            ATLAS_RT_KERNEL_SLAB_FLOAT | ATLAS_RT_KERNEL_LEAF_BATCH);
    const bool fslab_ok = !(kernel & ATLAS_RT_KERNEL_SLAB_FLOAT) || base == ATLAS_RT_KERNEL_BVH2_EXACT ||
                          base == ATLAS_RT_KERNEL_BVH2_PERSISTENT || base == ATLAS_RT_KERNEL_BVH4;
    const bool batch_ok = !(kernel & ATLAS_RT_KERNEL_LEAF_BATCH) || base == ATLAS_RT_KERNEL_BVH2_PERSISTENT ||
                          base == ATLAS_RT_KERNEL_BVH4;
    if (extra == 0 && fslab_ok && batch_ok && base >= ATLAS_RT_KERNEL_AUTO && base <= ATLAS_RT_KERNEL_BVH8 &&
        !(base == ATLAS_RT_KERNEL_AUTO && kernel != ATLAS_RT_KERNEL_AUTO))
        return ATLAS_RT_OK;
    return fail(ATLAS_RT_E_UNSUPPORTED, std::string(who) + ": kernel variant not available");
}

extern "C" {

const char* atlas_rt_last_error(void) { return last_error(); }

const char* atlas_rt_version(void) { return "atlas_rt 0.1.0 (gfx950)"; }

int32_t atlas_rt_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

atlas_rt_status atlas_rt_trace_closest_rays(const atlas_rt_blas* blas, const atlas_rt_ray* d_rays,
                                            atlas_rt_hit* d_hits, int32_t* d_slot, uint32_t count, int32_t kernel,
                                            void* stream) {
    atlas_rt_status st = check_blas_ready(blas, "atlas_rt_trace_closest_rays");
    if (st != ATLAS_RT_OK) return st;
    if (count && (!d_rays || !d_hits)) return fail(ATLAS_RT_E_INVALID_ARG, "atlas_rt_trace_closest_rays: NULL buffer");
    st = check_kernel(kernel, "atlas_rt_trace_closest_rays");
    if (st != ATLAS_RT_OK) return st;
    TraceLaunch L;
    L.rays = d_rays;
    L.hits = d_hits;
    L.slots = d_slot;
    L.count = count;
    L.kernel = kernel;
    L.stream = reinterpret_cast<hipStream_t>(stream);
    return launch_blas_closest(*const_cast<atlas_rt_blas*>(blas), L);
}

atlas_rt_status atlas_rt_trace_closest_rays_noalloc(const atlas_rt_blas* blas, const atlas_rt_ray* d_rays,
                                                    atlas_rt_hit* d_hits, int32_t* d_slot, uint32_t count,
                                                    int32_t kernel, void* stream) {
    atlas_rt_status st = check_blas_ready(blas, "atlas_rt_trace_closest_rays_noalloc");
    if (st != ATLAS_RT_OK) return st;
    if (count && (!d_rays || !d_hits))
        return fail(ATLAS_RT_E_INVALID_ARG, "atlas_rt_trace_closest_rays_noalloc: NULL buffer");
    st = check_kernel(kernel, "atlas_rt_trace_closest_rays_noalloc");
    if (st != ATLAS_RT_OK) return st;
    TraceLaunch L;
    L.rays = d_rays;
    L.hits = d_hits;
    L.slots = d_slot;
    L.count = count;
    L.kernel = kernel;
    L.stream = reinterpret_cast<hipStream_t>(stream);
    return launch_blas_closest_noalloc(*blas, L);
}

atlas_rt_status atlas_rt_sort_rays_workspace(uint32_t count, size_t* bytes) {
    if (!bytes) return fail(ATLAS_RT_E_INVALID_ARG, "atlas_rt_sort_rays_workspace: NULL argument");
    *bytes = bin_sort_workspace(count);
    return ATLAS_RT_OK;
}

atlas_rt_status atlas_rt_sort_rays_by_direction(const atlas_rt_ray* d_rays, uint32_t count, atlas_rt_ray* d_out,
                                                uint32_t* d_perm, uint32_t* d_bin_offsets, void* d_workspace,
                                                size_t workspace_bytes, void* stream) {
    if (count && (!d_rays || !d_out || !d_workspace))
        return fail(ATLAS_RT_E_INVALID_ARG, "atlas_rt_sort_rays_by_direction: NULL buffer");
    if (count && d_rays < d_out + count && d_out < d_rays + count)
        return fail(ATLAS_RT_E_INVALID_ARG, "atlas_rt_sort_rays_by_direction: input and output overlap");
    if (workspace_bytes < bin_sort_workspace(count))
        return fail(ATLAS_RT_E_INVALID_ARG, "atlas_rt_sort_rays_by_direction: workspace smaller than "
                                            "atlas_rt_sort_rays_workspace(count)");
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    const hipError_t e = launch_bin_sort(reinterpret_cast<const float4*>(d_rays), reinterpret_cast<float4*>(d_out), 2,
                                         nullptr, count, count, d_workspace, d_perm, d_bin_offsets, s);
    if (e != hipSuccess) return hip_fail(e, "atlas_rt_sort_rays_by_direction");
    return ATLAS_RT_OK;
}

atlas_rt_status atlas_rt_trace_closest_stats(const atlas_rt_blas* blas, const atlas_rt_ray* d_rays,
                                             atlas_rt_hit* d_hits, int32_t* d_slot, uint32_t* d_visits,
                                             uint32_t count, void* stream) {
    atlas_rt_status st = check_blas_ready(blas, "atlas_rt_trace_closest_stats");
    if (st != ATLAS_RT_OK) return st;
    if (count && (!d_rays || !d_hits || !d_visits))
        return fail(ATLAS_RT_E_INVALID_ARG, "atlas_rt_trace_closest_stats: NULL buffer");
    TraceLaunch L;
    L.rays = d_rays;
    L.hits = d_hits;
    L.slots = d_slot;
    L.visits = d_visits;
    L.count = count;
    L.kernel = ATLAS_RT_KERNEL_BVH2_EXACT;
    L.stream = reinterpret_cast<hipStream_t>(stream);
    return launch_blas_closest(*const_cast<atlas_rt_blas*>(blas), L);
}

atlas_rt_status atlas_rt_trace_any_rays(const atlas_rt_blas* blas, const atlas_rt_ray* d_rays, uint32_t* d_hit,
                                        uint32_t count, int32_t kernel, void* stream) {
    atlas_rt_status st = check_blas_ready(blas, "atlas_rt_trace_any_rays");
    if (st != ATLAS_RT_OK) return st;
    if (count && (!d_rays || !d_hit)) return fail(ATLAS_RT_E_INVALID_ARG, "atlas_rt_trace_any_rays: NULL buffer");
    if (kernel != ATLAS_RT_KERNEL_AUTO && kernel != ATLAS_RT_KERNEL_BVH2_EXACT &&
        kernel != ATLAS_RT_KERNEL_BVH2_PERSISTENT)
        return fail(ATLAS_RT_E_UNSUPPORTED, "atlas_rt_trace_any_rays: any hit runs the binary engine-order kernel");
    if (st != ATLAS_RT_OK) return st;
    TraceLaunch L;
    L.rays = d_rays;
    L.hits = d_hit;
    L.count = count;
    L.kernel = kernel;
    L.stream = reinterpret_cast<hipStream_t>(stream);
    return launch_blas_any(*const_cast<atlas_rt_blas*>(blas), L);
}

atlas_rt_status atlas_rt_trace_closest_host(const atlas_rt_blas* blas, const atlas_rt_ray* rays, atlas_rt_hit* hits,
                                            uint32_t count) {
    if (!blas) return fail(ATLAS_RT_E_INVALID_ARG, "atlas_rt_trace_closest_host: NULL BLAS handle");
    if (count == 0) return ATLAS_RT_OK;
    if (!rays || !hits) return fail(ATLAS_RT_E_INVALID_ARG, "atlas_rt_trace_closest_host: NULL buffer");
    host_trace_closest(*blas, rays, hits, count);
    return ATLAS_RT_OK;
}

atlas_rt_status atlas_rt_trace_any_host(const atlas_rt_blas* blas, const atlas_rt_ray* rays, uint32_t* hit,
                                        uint32_t count) {
    if (!blas) return fail(ATLAS_RT_E_INVALID_ARG, "atlas_rt_trace_any_host: NULL BLAS handle");
    if (count == 0) return ATLAS_RT_OK;
    if (!rays || !hit) return fail(ATLAS_RT_E_INVALID_ARG, "atlas_rt_trace_any_host: NULL buffer");
    host_trace_any(*blas, rays, hit, count);
    return ATLAS_RT_OK;
}

atlas_rt_status atlas_rt_blas_set_stack_limit(atlas_rt_blas* blas, uint32_t entries) {
    if (!blas) return fail(ATLAS_RT_E_INVALID_ARG, "atlas_rt_blas_set_stack_limit: NULL BLAS handle");
    blas->stack_limit = entries;
    return ATLAS_RT_OK;
}

atlas_rt_status atlas_rt_blas_trace_status(atlas_rt_blas* blas, void* stream) {
    if (!blas) return fail(ATLAS_RT_E_INVALID_ARG, "atlas_rt_blas_trace_status: NULL BLAS handle");
    return take_overflow_flag(blas->flags.ptr, stream, "atlas_rt_blas_trace_status");
}

atlas_rt_status atlas_rt_last_kernel_ms(const atlas_rt_blas* blas, float* ms) {
    if (!blas || !ms) return fail(ATLAS_RT_E_INVALID_ARG, "atlas_rt_last_kernel_ms: NULL argument");
    return const_cast<atlas_rt_blas*>(blas)->timing.read(ms);
}

}  // extern "C"
