// Path tracer objects: PathTracingRenderer's host side (buffers, light list,
// per-frame seeds, progressive accumulation) over the wavefront kernels.
#include "pathtracer.h"
#include "nrc.h"

#include <algorithm>
#include <cstring>

using namespace atlas_rt;
using namespace atlas_rt::pt;

namespace {

template <class T>
atlas_rt_status dev_alloc(T*& p, size_t count) {
    void* v = nullptr;
    ATLAS_HIP_TRY(hipMalloc(&v, std::max<size_t>(count, 1) * sizeof(T)));
    p = reinterpret_cast<T*>(v);
    return ATLAS_RT_OK;
}

void free_all(PathTracer& P) {
    void* ptrs[] = {P.ctl, P.paths[0], P.paths[1], P.hits, P.shadow, P.Lpix, P.accum, P.xform,
                    const_cast<float4**>(P.shade_table), P.materials, P.lights};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    P.spill.release();
    P.flags.release();
    free_cache_buffers(P);
}

// Random::Init seeds PCG32 with uint64(SampleUniformFloat()) = 0
// (RandomHelper.cpp:26-37): state = increment, then one draw.
uint64_t pcg_initial_state() {
    uint64_t state = 0ull + 1442695040888963407ull;
    state = state * 6364136223846793005ull + 1442695040888963407ull;
    return state;
}

}  // namespace

void atlas_rt::pt::free_cache_buffers(PathTracer& P) {
    void* ptrs[] = {P.rec, P.q_buf, P.q_meta, P.q_out, P.s_q, P.s_t};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    P.rec = P.q_buf = P.q_meta = P.q_out = P.s_q = P.s_t = nullptr;
    P.nrc = nullptr;
    P.train_cap = 0;
    P.shade_args.cache = CacheArgs{};
}

extern "C" {

atlas_rt_status atlas_rt_pathtracer_create(const atlas_rt_world* world, const atlas_rt_pt_params* params,
                                           const atlas_rt_material* materials, uint32_t material_count,
                                           const atlas_rt_light* lights, uint32_t light_count,
                                           atlas_rt_pathtracer** out) {
    if (!out || !world || !params) return fail(ATLAS_RT_E_INVALID_ARG, "atlas_rt_pathtracer_create: NULL argument");
    *out = nullptr;
    if (!world->valid) return fail(ATLAS_RT_E_INVALID_ARG, "atlas_rt_pathtracer_create: world has no instances");
    if (!params->width || !params->height || params->bounces < 0 || params->bounces > 15)
        return fail(ATLAS_RT_E_INVALID_ARG, "atlas_rt_pathtracer_create: need width, height > 0 and 0 <= bounces <= 15");
    if (!materials || !material_count) return fail(ATLAS_RT_E_INVALID_ARG, "atlas_rt_pathtracer_create: no materials");
    if (light_count && !lights) return fail(ATLAS_RT_E_INVALID_ARG, "atlas_rt_pathtracer_create: NULL lights");
    for (const atlas_rt_blas* b : world->blas_handles)
        if (!b->d_shade)
            return fail(ATLAS_RT_E_INVALID_ARG, "atlas_rt_pathtracer_create: a BLAS has no shading data "
                                                "(atlas_rt_blas_set_shading)");
    for (uint32_t l = 0; l < light_count; l++)
        if (lights[l].type != ATLAS_RT_LIGHT_DIRECTIONAL)
            return fail(ATLAS_RT_E_UNSUPPORTED, "atlas_rt_pathtracer_create: only directional lights");
    auto* pt = new atlas_rt_pathtracer();
    PathTracer& P = pt->P;
    P.width = params->width;
    P.height = params->height;
    P.bounces = params->bounces;
    P.light_count = int32_t(light_count);
    P.world = world;
    P.world_dev = world->dev();
    P.cap = world->max_depth + 1;
    const size_t pixels = size_t(P.width) * P.height;
    atlas_rt_status st = ATLAS_RT_OK;
    if (st == ATLAS_RT_OK) st = dev_alloc(P.ctl, 1);
    if (st == ATLAS_RT_OK) st = dev_alloc(P.paths[0], 3 * pixels);
    if (st == ATLAS_RT_OK) st = dev_alloc(P.paths[1], 3 * pixels);
    if (st == ATLAS_RT_OK) st = dev_alloc(P.hits, pixels);
    if (st == ATLAS_RT_OK) st = dev_alloc(P.shadow, 3 * pixels);
    if (st == ATLAS_RT_OK) st = dev_alloc(P.Lpix, pixels);
    if (st == ATLAS_RT_OK) st = dev_alloc(P.accum, pixels);
    if (st == ATLAS_RT_OK) st = dev_alloc(P.xform, 3 * world->instances.size());
    if (st == ATLAS_RT_OK) st = dev_alloc(P.shade_table, world->blas_handles.size());
    if (st == ATLAS_RT_OK) st = dev_alloc(P.materials, material_count);
    if (st == ATLAS_RT_OK) st = dev_alloc(P.lights, std::max<uint32_t>(light_count, 1));
    const uint32_t grid_threads = uint32_t((pixels + 127) / 128 * 128);
    if (st == ATLAS_RT_OK && P.cap > 32) st = P.spill.ensure(size_t(P.cap - 32) * grid_threads * sizeof(int32_t));
    if (st == ATLAS_RT_OK) st = P.flags.ensure(256);
    if (st != ATLAS_RT_OK) {
        free_all(P);
        delete pt;
        return st;
    }
    // Light list (UpdateLights, RayTracingHelper.cpp:436-500): weights are the
    // brightness, floored at 0.5% of the maximum and normalised.
    std::vector<Light> L(std::max<uint32_t>(light_count, 1));
    float maxw = 0.0f, total = 0.0f;
    for (uint32_t l = 0; l < light_count; l++) {
        const float* r = lights[l].radiance;
        L[l].pdf = (r[0] * 0.3333f + r[1] * 0.3333f) + r[2] * 0.3333f;
        maxw = std::max(maxw, L[l].pdf);
    }
    for (uint32_t l = 0; l < light_count; l++) {
        L[l].pdf = std::max(L[l].pdf, 0.005f * maxw);
        total += L[l].pdf;
    }
    for (uint32_t l = 0; l < light_count; l++) {
        std::memcpy(L[l].N, lights[l].direction, 12);
        std::memcpy(L[l].radiance, lights[l].radiance, 12);
        L[l].type = lights[l].type;
        L[l].pdf /= total;
    }
    std::vector<const float4*> shade(world->blas_handles.size());
    for (size_t k = 0; k < shade.size(); k++) shade[k] = world->blas_handles[k]->d_shade;
    FrameCtl ctl{};
    ctl.pcg_state = pcg_initial_state();
    hipError_t e = hipSuccess;
    e = hipMemcpy(P.ctl, &ctl, sizeof(ctl), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemset(P.accum, 0, pixels * sizeof(float4));
    if (e == hipSuccess) e = hipMemset(P.flags.ptr, 0, 256);
    if (e == hipSuccess)
        e = hipMemcpy(P.xform, world->xform.data(), world->xform.size() * sizeof(float), hipMemcpyHostToDevice);
    if (e == hipSuccess)
        e = hipMemcpy(const_cast<float4**>(P.shade_table), shade.data(), shade.size() * sizeof(float4*),
                      hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(P.materials, materials, material_count * sizeof(Material), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(P.lights, L.data(), L.size() * sizeof(Light), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        free_all(P);
        delete pt;
        return hip_fail(e, "atlas_rt_pathtracer_create");
    }
    ShadeArgs& A = P.shade_args;
    A.world = P.world_dev;
    A.xform = P.xform;
    A.shade = P.shade_table;
    A.materials = P.materials;
    A.lights = P.lights;
    A.light_count = P.light_count;
    A.bounces = P.bounces;
    std::memcpy(A.env, params->env_radiance, 12);
    A.ctl = P.ctl;
    A.hits = P.hits;
    A.shadow = P.shadow;
    A.Lpix = P.Lpix;
    *out = pt;
    return ATLAS_RT_OK;
}

atlas_rt_status atlas_rt_pathtracer_destroy(atlas_rt_pathtracer* pt) {
    if (!pt) return ATLAS_RT_OK;
    free_all(pt->P);
    delete pt;
    return ATLAS_RT_OK;
}

atlas_rt_status atlas_rt_pathtracer_render(atlas_rt_pathtracer* pt, const atlas_rt_camera* camera, uint32_t frames,
                                           void* stream) {
    if (!pt || !camera) return fail(ATLAS_RT_E_INVALID_ARG, "atlas_rt_pathtracer_render: NULL argument");
    Camera cam;
    std::memcpy(cam.loc, camera->location, 12);
    std::memcpy(cam.origin, camera->origin, 12);
    std::memcpy(cam.right, camera->right, 12);
    std::memcpy(cam.bottom, camera->bottom, 12);
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    for (uint32_t f = 0; f < frames; f++) {
        atlas_rt_status st = launch_frame(pt->P, cam, s);
        if (st != ATLAS_RT_OK) return st;
        pt->P.host_samples++;
    }
    return ATLAS_RT_OK;
}

atlas_rt_status atlas_rt_pathtracer_reset(atlas_rt_pathtracer* pt) {
    if (!pt) return fail(ATLAS_RT_E_INVALID_ARG, "atlas_rt_pathtracer_reset: NULL");
    PathTracer& P = pt->P;
    ATLAS_HIP_TRY(hipDeviceSynchronize());
    ATLAS_HIP_TRY(hipMemset(P.accum, 0, size_t(P.width) * P.height * sizeof(float4)));
    // sample_count back to 0; the PCG stream continues (Random is global)
    ATLAS_HIP_TRY(hipMemset(reinterpret_cast<char*>(P.ctl) + offsetof(FrameCtl, sample_count), 0, sizeof(int32_t)));
    P.host_samples = 0;
    return ATLAS_RT_OK;
}

atlas_rt_status atlas_rt_pathtracer_read(atlas_rt_pathtracer* pt, float* d_out, void* stream) {
    if (!pt || !d_out) return fail(ATLAS_RT_E_INVALID_ARG, "atlas_rt_pathtracer_read: NULL argument");
    PathTracer& P = pt->P;
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    ATLAS_HIP_TRY(hipMemcpyAsync(d_out, P.accum, size_t(P.width) * P.height * sizeof(float4), hipMemcpyDeviceToDevice, s));
    ATLAS_HIP_TRY(hipStreamSynchronize(s));
    return ATLAS_RT_OK;
}

atlas_rt_status atlas_rt_pathtracer_sample_count(const atlas_rt_pathtracer* pt, int32_t* count) {
    if (!pt || !count) return fail(ATLAS_RT_E_INVALID_ARG, "atlas_rt_pathtracer_sample_count: NULL argument");
    *count = pt->P.host_samples;
    return ATLAS_RT_OK;
}

atlas_rt_status atlas_rt_pathtracer_set_cache(atlas_rt_pathtracer* pt, atlas_rt_nrc* cache,
                                              const atlas_rt_pt_cache_params* params) {
    if (!pt) return fail(ATLAS_RT_E_INVALID_ARG, "atlas_rt_pathtracer_set_cache: NULL path tracer");
    PathTracer& P = pt->P;
    ATLAS_HIP_TRY(hipDeviceSynchronize());
    free_cache_buffers(P);
    if (!cache) return ATLAS_RT_OK;
    if (!params) return fail(ATLAS_RT_E_INVALID_ARG, "atlas_rt_pathtracer_set_cache: NULL params");
    if (P.bounces < 2) return fail(ATLAS_RT_E_INVALID_ARG, "atlas_rt_pathtracer_set_cache: needs bounces >= 2");
    if (params->train_stride < 1 || params->train_steps > 64)
        return fail(ATLAS_RT_E_INVALID_ARG, "atlas_rt_pathtracer_set_cache: need train_stride >= 1, train_steps <= 64");
    int dev = -1;
    ATLAS_HIP_TRY(hipGetDevice(&dev));
    if (dev != cache->device) return fail(ATLAS_RT_E_INVALID_ARG, "atlas_rt_pathtracer_set_cache: cache on another device");
    const size_t pixels = size_t(P.width) * P.height;
    // about pixels / stride training paths per frame; the hash selection is
    // binomial, so leave twice that (the surplus pixels render normally)
    const size_t cap = std::min(pixels, 2 * (pixels / params->train_stride) + 256);
    const size_t recs = cap * size_t(P.bounces + 1);
    atlas_rt_status st = ATLAS_RT_OK;
    if (st == ATLAS_RT_OK) st = dev_alloc(P.rec, 5 * recs);
    if (st == ATLAS_RT_OK) st = dev_alloc(P.q_buf, 3 * pixels);
    if (st == ATLAS_RT_OK) st = dev_alloc(P.q_meta, pixels);
    if (st == ATLAS_RT_OK) st = dev_alloc(P.q_out, pixels);
    if (st == ATLAS_RT_OK) st = dev_alloc(P.s_q, 3 * recs);
    if (st == ATLAS_RT_OK) st = dev_alloc(P.s_t, recs);
    if (st != ATLAS_RT_OK) {
        free_cache_buffers(P);
        return st;
    }
    P.nrc = cache;
    P.train_steps = params->train_steps;
    P.train_cap = uint32_t(cap);
    CacheArgs& C = P.shade_args.cache;
    C.rec_q = P.rec;
    C.rec_beta = P.rec + 3 * recs;
    C.rec_add = P.rec + 4 * recs;
    C.q_buf = P.q_buf;
    C.q_meta = P.q_meta;
    C.train_cap = uint32_t(cap);
    C.train_stride = params->train_stride;
    C.enabled = 1;
    return ATLAS_RT_OK;
}

atlas_rt_status atlas_rt_pathtracer_cache_stats(atlas_rt_pathtracer* pt, uint32_t* train_samples, uint32_t* queries,
                                                float* loss) {
    if (!pt) return fail(ATLAS_RT_E_INVALID_ARG, "atlas_rt_pathtracer_cache_stats: NULL path tracer");
    FrameCtl ctl;
    ATLAS_HIP_TRY(hipDeviceSynchronize());
    ATLAS_HIP_TRY(hipMemcpy(&ctl, pt->P.ctl, sizeof(ctl), hipMemcpyDeviceToHost));
    if (train_samples) *train_samples = ctl.train_samples;
    if (queries) *queries = ctl.queries;
    if (loss) *loss = ctl.train_loss;
    return ATLAS_RT_OK;
}

atlas_rt_status atlas_rt_pathtracer_cache_samples(atlas_rt_pathtracer* pt, float* d_queries, float* d_targets,
                                                  uint32_t max_count, uint32_t* count) {
    if (!pt || !count) return fail(ATLAS_RT_E_INVALID_ARG, "atlas_rt_pathtracer_cache_samples: NULL argument");
    PathTracer& P = pt->P;
    *count = 0;
    if (!P.nrc) return fail(ATLAS_RT_E_INVALID_ARG, "atlas_rt_pathtracer_cache_samples: no cache attached");
    FrameCtl ctl;
    ATLAS_HIP_TRY(hipDeviceSynchronize());
    ATLAS_HIP_TRY(hipMemcpy(&ctl, P.ctl, sizeof(ctl), hipMemcpyDeviceToHost));
    const uint32_t n = std::min(std::min(ctl.train_samples, P.train_cap * uint32_t(P.bounces + 1)), max_count);
    if (n && (!d_queries || !d_targets))
        return fail(ATLAS_RT_E_INVALID_ARG, "atlas_rt_pathtracer_cache_samples: NULL buffer");
    if (n) {
        ATLAS_HIP_TRY(hipMemcpy(d_queries, P.s_q, size_t(n) * 3 * sizeof(float4), hipMemcpyDeviceToDevice));
        ATLAS_HIP_TRY(hipMemcpy(d_targets, P.s_t, size_t(n) * sizeof(float4), hipMemcpyDeviceToDevice));
    }
    *count = n;
    return ATLAS_RT_OK;
}

}  // extern "C"
