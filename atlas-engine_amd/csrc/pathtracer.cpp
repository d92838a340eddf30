// Path tracer objects: PathTracingRenderer's host side (buffers, light list,
// per-frame seeds, progressive accumulation) over the wavefront kernels.
#include "pathtracer.h"
#include "binning.h"
#include "nrc.h"

#include <algorithm>
#include <cmath>
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

// Slot k's buffers sized by the paths per frame: both path queues, hits,
// shadow rays, radiance per ray ID, the traversal stack spill and the
// binning buffers.
atlas_rt_status alloc_slot_rays(PathTracer& P, uint32_t k, uint32_t rays, uint32_t frames = 1) {
    FrameSlot& S = P.slot[k];
    void* old[] = {S.paths[0], S.paths[1], S.hits, S.shadow, S.Lpix, S.binned, S.order, S.order_cnt, S.order_cls};
    for (void* p : old)
        if (p) (void)hipFree(p);
    S.paths[0] = S.paths[1] = nullptr;
    S.hits = nullptr;
    S.shadow = S.Lpix = S.binned = nullptr;
    S.order = S.order_cnt = nullptr;
    S.order_cls = nullptr;
    S.order_cap = 0;
    const size_t paths = size_t(rays) * frames;
    const size_t span = (paths + kPullRanges - 1) / kPullRanges;
    const size_t order_blocks = (span + kOrderSpan - 1) / kOrderSpan;
    atlas_rt_status st = dev_alloc(S.paths[0], 3 * size_t(rays));
    if (st == ATLAS_RT_OK) st = dev_alloc(S.paths[1], 3 * size_t(rays));
    if (st == ATLAS_RT_OK) st = dev_alloc(S.hits, rays);
    if (st == ATLAS_RT_OK) st = dev_alloc(S.shadow, 3 * size_t(rays));
    if (st == ATLAS_RT_OK) st = dev_alloc(S.Lpix, size_t(rays) * frames);
    if (st == ATLAS_RT_OK) st = dev_alloc(S.order, paths);
    if (st == ATLAS_RT_OK) st = dev_alloc(S.order_cnt, kPullRanges * (order_blocks + 1));
    if (st == ATLAS_RT_OK) st = dev_alloc(S.order_cls, paths);
    if (st == ATLAS_RT_OK) S.order_cap = paths;
    if (st == ATLAS_RT_OK) st = S.spill.ensure(slot_spill_bytes(rays, P.cap));
    if (st == ATLAS_RT_OK && P.binning) st = dev_alloc(S.binned, 3 * size_t(rays));
    if (st == ATLAS_RT_OK && P.binning) st = S.bin_ws.ensure(bin_sort_workspace(rays));
    return st;
}

// A slot's frame block, flag words, stream and events.
atlas_rt_status alloc_slot(PathTracer& P, uint32_t k, uint32_t rays) {
    FrameSlot& S = P.slot[k];
    atlas_rt_status st = dev_alloc(S.ctl, kBatchFrames);
    if (st == ATLAS_RT_OK) st = S.flags.ensure(256);
    if (st == ATLAS_RT_OK) st = alloc_slot_rays(P, k, rays, P.batch);
    if (st != ATLAS_RT_OK) return st;
    ATLAS_HIP_TRY(hipMemset(S.ctl, 0, kBatchFrames * sizeof(FrameCtl)));
    ATLAS_HIP_TRY(hipMemset(S.flags.ptr, 0, 256));
    ATLAS_HIP_TRY(hipStreamCreateWithFlags(&S.stream, hipStreamNonBlocking));
    ATLAS_HIP_TRY(hipEventCreateWithFlags(&S.ev_begin, hipEventDisableTiming));
    ATLAS_HIP_TRY(hipEventCreateWithFlags(&S.ev_train, hipEventDisableTiming));
    ATLAS_HIP_TRY(hipEventCreateWithFlags(&S.ev_acc, hipEventDisableTiming));
    return ATLAS_RT_OK;
}

void free_slot_cache(FrameSlot& S) {
    void* ptrs[] = {S.rec, S.q_buf, S.q_meta, S.q_out, S.s_q, S.s_t, S.train_list};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    S.rec = S.q_buf = S.q_meta = S.q_out = S.s_q = S.s_t = nullptr;
    S.train_list = nullptr;
}

void free_slot(FrameSlot& S) {
    void* ptrs[] = {S.ctl, S.paths[0], S.paths[1], S.hits, S.shadow, S.Lpix, S.binned, S.order, S.order_cnt,
                    S.order_cls};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    S.spill.release();
    S.flags.release();
    S.bin_ws.release();
    free_slot_cache(S);
    if (S.stream) (void)hipStreamDestroy(S.stream);
    hipEvent_t evs[] = {S.ev_begin, S.ev_train, S.ev_acc};
    for (hipEvent_t e : evs)
        if (e) (void)hipEventDestroy(e);
    S = FrameSlot{};
}

void free_realtime(PathTracer& P) {
    void* ptrs[] = {P.rt_frame, P.rt_radiance[0], P.rt_radiance[1], P.rt_depth[0], P.rt_depth[1], P.rt_velocity[0],
                    P.rt_velocity[1], P.rt_normal[0], P.rt_normal[1], P.rt_material[0], P.rt_material[1],
                    P.rt_last_xform};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    P.rt_frame = P.rt_last_xform = nullptr;
    for (int k = 0; k < 2; k++) {
        P.rt_radiance[k] = nullptr;
        P.rt_depth[k] = nullptr;
        P.rt_velocity[k] = P.rt_normal[k] = nullptr;
        P.rt_material[k] = nullptr;
    }
    P.realtime = false;
}

void free_all(PathTracer& P) {
    void* ptrs[] = {P.ctl, P.accum, const_cast<float4**>(P.shade_table), const_cast<float4**>(P.uv_table),
                    const_cast<float4**>(P.tangent_table), const_cast<float4**>(P.color_table), P.materials,
                    P.lights, P.tex, P.texels, P.stats, P.hit0};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    free_cache_buffers(P);
    free_realtime(P);
    for (FrameSlot& S : P.slot) free_slot(S);
    if (P.graph) (void)hipGraphExecDestroy(P.graph);
    P.graph = nullptr;
    if (P.ev_start) (void)hipEventDestroy(P.ev_start);
    P.ev_start = nullptr;
    for (hipEvent_t& e : P.occ_ev) {
        if (e) (void)hipEventDestroy(e);
        e = nullptr;
    }
}

// Random::Init seeds PCG32 with uint64(SampleUniformFloat()) = 0
// (RandomHelper.cpp:26-37): state = increment, then one draw.
uint64_t pcg_initial_state() {
    uint64_t state = 0ull + 1442695040888963407ull;
    state = state * 6364136223846793005ull + 1442695040888963407ull;
    return state;
}

float bits_to_float(uint32_t u) {
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}

struct H3 {
    float x, y, z;
};
inline H3 operator+(H3 a, H3 b) { return H3{a.x + b.x, a.y + b.y, a.z + b.z}; }
inline H3 operator-(H3 a, H3 b) { return H3{a.x - b.x, a.y - b.y, a.z - b.z}; }
inline float hdot(H3 a, H3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
// glm's mat4 * vec4 over object-to-world rows (columns summed in order)
inline H3 hxf(const float* M, H3 p, float w) {
    return H3{((M[0] * p.x + M[1] * p.y) + M[2] * p.z) + M[3] * w, ((M[4] * p.x + M[5] * p.y) + M[6] * p.z) + M[7] * w,
              ((M[8] * p.x + M[9] * p.y) + M[10] * p.z) + M[11] * w};
}

// RayTracingWorld::BuildTriangleLightsForMesh (RayTracingWorld.cpp:322-376)
// per BLAS slot and UpdateTriangleLights (:378-407) per instance of the BLAS:
// centroid and mean vertex normal moved to world space, Heron's area of the
// object-space triangle, weight = area * brightness of linear(colour), radiance
// = linear(linear(colour)) * intensity.  The instance index is the world's
// TLAS slot of the instance.
void triangle_lights(const atlas_rt_world& W, const float* emitters, uint32_t n_emitters, std::vector<Light>& out) {
    for (size_t k = 0; k < W.blas_handles.size(); k++) {
        const atlas_rt_blas& b = *W.blas_handles[k];
        for (size_t i = 0; i < b.tris.size(); i++) {
            const atlas_rt_gpu_bvh_triangle& tri = b.tris[i];
            const float* ns = &b.shade_normals[9 * i];
            for (size_t j = 0; j < W.instances.size(); j++) {
                const atlas_rt_gpu_bvh_instance& in = W.instances[j];
                if (in.mesh_offset != int32_t(k)) continue;
                const int64_t e = int64_t(b.shade_mat[i]) + in.material_offset;
                if (e < 0 || e >= int64_t(n_emitters)) continue;
                const float* em = emitters + 4 * size_t(e);
                const H3 lin{std::pow(em[0], 2.2f), std::pow(em[1], 2.2f), std::pow(em[2], 2.2f)};
                const float brightness = hdot(lin, H3{0.3333f, 0.3333f, 0.3333f});
                if (!(brightness > 0.0f)) continue;
                const H3 v0{tri.v0[0], tri.v0[1], tri.v0[2]}, v1{tri.v1[0], tri.v1[1], tri.v1[2]},
                    v2{tri.v2[0], tri.v2[1], tri.v2[2]};
                const H3 c = (v0 + v1) + v2;
                const H3 nsum = (H3{ns[0], ns[1], ns[2]} + H3{ns[3], ns[4], ns[5]}) + H3{ns[6], ns[7], ns[8]};
                const H3 P{c.x / 3.0f, c.y / 3.0f, c.z / 3.0f}, N{nsum.x / 3.0f, nsum.y / 3.0f, nsum.z / 3.0f};
                const H3 ea = v1 - v0, eb = v2 - v0, ec = v1 - v2;
                const float a = std::sqrt(hdot(ea, ea)), bb = std::sqrt(hdot(eb, eb)), cc = std::sqrt(hdot(ec, ec));
                const float p = 0.5f * ((a + bb) + cc);
                const float area = std::sqrt(((p * (p - a)) * (p - bb)) * (p - cc));
                const float* M = &W.xform[12 * j];
                const H3 WP = hxf(M, P, 1.0f), WN = hxf(M, N, 0.0f);
                Light L{};
                L.P[0] = WP.x;
                L.P[1] = WP.y;
                L.P[2] = WP.z;
                L.P[3] = 1.0f;
                L.N[0] = WN.x;
                L.N[1] = WN.y;
                L.N[2] = WN.z;
                L.color[0] = std::pow(lin.x, 2.2f) * em[3];
                L.color[1] = std::pow(lin.y, 2.2f) * em[3];
                L.color[2] = std::pow(lin.z, 2.2f) * em[3];
                L.data[0] = bits_to_float((ATLAS_RT_LIGHT_TRIANGLE << 28) | uint32_t(i));
                L.data[1] = area * brightness;
                L.data[2] = area;
                L.data[3] = bits_to_float(uint32_t(j));
                out.push_back(L);
            }
        }
    }
}

atlas_rt_status upload_lights(PathTracer& P) {
    const uint32_t n = uint32_t(P.light_list.size());
    if (n > P.light_cap) {
        Light* d = nullptr;
        ATLAS_HIP_TRY(hipMalloc(reinterpret_cast<void**>(&d), n * sizeof(Light)));
        if (P.lights) (void)hipFree(P.lights);
        P.lights = d;
        P.light_cap = n;
    }
    if (n) ATLAS_HIP_TRY(hipMemcpy(P.lights, P.light_list.data(), n * sizeof(Light), hipMemcpyHostToDevice));
    P.light_count = int32_t(n);
    P.shade_args.lights = P.lights;
    P.shade_args.light_count = P.light_count;
    return ATLAS_RT_OK;
}

}  // namespace

void atlas_rt::pt::build_light_list(PathTracer& P, const float* emitters, uint32_t emitter_count) {
    // UpdateLights (RayTracingHelper.cpp:436-500): punctual lights weighted by
    // brightness, then the triangle lights; every weight floored at 0.5% of
    // the maximum, then normalised.
    std::vector<Light>& L = P.light_list;
    L.clear();
    for (size_t l = 0; l < P.punctual.size(); l++) {
        const atlas_rt_light& src = P.punctual[l];
        Light g{};
        g.P[3] = 1.0f;
        std::memcpy(g.N, src.direction, 12);
        std::memcpy(g.color, src.radiance, 12);
        g.data[0] = bits_to_float((ATLAS_RT_LIGHT_DIRECTIONAL << 28) | uint32_t(l));
        g.data[1] = (src.radiance[0] * 0.3333f + src.radiance[1] * 0.3333f) + src.radiance[2] * 0.3333f;
        L.push_back(g);
    }
    if (emitters && emitter_count) triangle_lights(*P.world, emitters, emitter_count, L);
    float maxw = 0.0f, total = 0.0f;
    for (const Light& g : L) maxw = std::max(maxw, g.data[1]);
    for (Light& g : L) {
        g.data[1] = std::max(g.data[1], 0.005f * maxw);
        total += g.data[1];
    }
    for (Light& g : L) g.data[1] /= total;
}

void atlas_rt::pt::free_cache_buffers(PathTracer& P) {
    for (FrameSlot& S : P.slot) free_slot_cache(S);
    for (int j = 0; j < 2; j++) {
        if (P.snap_table[j]) (void)hipFree(P.snap_table[j]);
        if (P.snap_wn[j]) (void)hipFree(P.snap_wn[j]);
        P.snap_table[j] = P.snap_wn[j] = nullptr;
    }
    P.query_lag = 0;
    P.nrc = nullptr;
    P.train_cap = 0;
    P.shade_args.cache = CacheArgs{};
}

size_t atlas_rt::pt::slot_spill_bytes(uint32_t rays, uint32_t cap) {
    size_t need = 0;
    if (cap > uint32_t(kGridLds)) {
        const size_t grid_threads = (size_t(rays) + kGridBlock - 1) / kGridBlock * kGridBlock;
        need = size_t(cap - kGridLds) * grid_threads * sizeof(int32_t);
    }
    if (cap > uint32_t(kResidentLds)) {
        const size_t lanes = size_t(device_cu_count()) * kResidentGroupsPerCu * kResidentBlock;
        need = std::max(need, size_t(cap - kResidentLds) * lanes * sizeof(int32_t));
    }
    return need;
}

atlas_rt_status atlas_rt::pt::apply_stack_cap(PathTracer& P) {
    const uint32_t cap = P.world->stack_cap();
    if (cap == P.cap && P.cap_bound) return ATLAS_RT_OK;
    // a larger spill replaces the old one, which frames in flight may still use
    if (P.cap_bound && cap > P.cap) ATLAS_HIP_TRY(hipDeviceSynchronize());
    const size_t pixels = size_t(P.width) * P.height;
    for (uint32_t k = 0; k < P.slots; k++) {
        if (!P.slot[k].ctl) continue;
        const size_t rays = k == 0 ? P.rays : pixels;  // realtime mode renders in slot 0
        atlas_rt_status st = P.slot[k].spill.ensure(slot_spill_bytes(uint32_t(rays), cap));
        if (st != ATLAS_RT_OK) return st;
    }
    P.cap = cap;
    P.cap_bound = true;
    return ATLAS_RT_OK;
}

atlas_rt_status atlas_rt::pt::bind_world(PathTracer& P) {
    const atlas_rt_world& W = *P.world;
    if (P.world_gen == W.generation && P.history_serial == W.history_serial) return ATLAS_RT_OK;
    if (!W.valid) return fail(ATLAS_RT_E_INVALID_ARG, "path tracer: the world has no instances (failed update?)");
    for (const atlas_rt_blas* b : W.blas_handles)
        if (!b->d_shade)
            return fail(ATLAS_RT_E_INVALID_ARG, "path tracer: a BLAS of the updated world has no shading data "
                                                "(atlas_rt_blas_set_shading)");
    // frames in flight still read the tables and the realtime history
    ATLAS_HIP_TRY(hipDeviceSynchronize());
    const size_t nb = W.blas_handles.size();
    if (nb != P.table_count) {
        const void* old[] = {P.shade_table, P.uv_table, P.tangent_table, P.color_table};
        for (const void* p : old)
            if (p) (void)hipFree(const_cast<void*>(p));
        P.shade_table = P.uv_table = P.tangent_table = P.color_table = nullptr;
        P.table_count = 0;
        atlas_rt_status st = dev_alloc(P.shade_table, nb);
        if (st == ATLAS_RT_OK) st = dev_alloc(P.uv_table, nb);
        if (st == ATLAS_RT_OK) st = dev_alloc(P.tangent_table, nb);
        if (st == ATLAS_RT_OK) st = dev_alloc(P.color_table, nb);
        if (st != ATLAS_RT_OK) return st;
        P.table_count = nb;
    }
    std::vector<const float4*> shade(nb), uv(nb), tangent(nb), color(nb);
    bool alpha = false;
    for (size_t k = 0; k < nb; k++) {
        shade[k] = W.blas_handles[k]->d_shade;
        uv[k] = W.blas_handles[k]->d_uv;
        tangent[k] = W.blas_handles[k]->d_tangent;
        color[k] = W.blas_handles[k]->d_color;
        alpha = alpha || W.blas_handles[k]->has_alpha;
    }
    ATLAS_HIP_TRY(hipMemcpy(const_cast<float4**>(P.shade_table), shade.data(), nb * sizeof(float4*),
                            hipMemcpyHostToDevice));
    ATLAS_HIP_TRY(hipMemcpy(const_cast<float4**>(P.uv_table), uv.data(), nb * sizeof(float4*), hipMemcpyHostToDevice));
    ATLAS_HIP_TRY(hipMemcpy(const_cast<float4**>(P.tangent_table), tangent.data(), nb * sizeof(float4*),
                            hipMemcpyHostToDevice));
    ATLAS_HIP_TRY(hipMemcpy(const_cast<float4**>(P.color_table), color.data(), nb * sizeof(float4*),
                            hipMemcpyHostToDevice));
    P.alpha = alpha;
    P.world_dev = W.dev();
    P.xform = reinterpret_cast<const float4*>(W.d_xform);  // the world's, kept current by refits
    {
        atlas_rt_status st = apply_stack_cap(P);
        if (st != ATLAS_RT_OK) return st;
    }
    ShadeArgs& A = P.shade_args;
    A.world = P.world_dev;
    A.xform = P.xform;
    A.shade = P.shade_table;
    A.tangent = P.tangent_table;
    A.color = P.color_table;
    A.alpha.shade = P.shade_table;
    A.alpha.uv = P.uv_table;
    if (P.realtime && P.rt_last_xform) {
        // The update re-packs the instances.  With an object history for this
        // update (atlas_rt_world_set_history) each slot takes its instance's
        // last matrix from it; otherwise the previous frame's transform of
        // each instance moves to its new slot, matched by caller index, and
        // an index the previous update did not have starts without motion.
        const size_t m_old = P.world_order.size(), m_new = W.order.size();
        std::vector<float> last(12 * m_old), moved(12 * m_new);
        const bool history = W.history_gen == W.generation && W.history.size() == 12 * m_new;
        if (m_old && !history)
            ATLAS_HIP_TRY(hipMemcpy(last.data(), P.rt_last_xform, 12 * m_old * sizeof(float), hipMemcpyDeviceToHost));
        std::vector<int64_t> where;
        for (size_t j = 0; j < m_old && !history; j++) {
            if (P.world_order[j] >= where.size()) where.resize(P.world_order[j] + 1, -1);
            where[P.world_order[j]] = int64_t(j);
        }
        for (size_t j = 0; j < m_new; j++) {
            const uint32_t id = W.order[j];
            if (history) {
                std::memcpy(&moved[12 * j], &W.history[12 * size_t(id)], 12 * sizeof(float));
                continue;
            }
            const int64_t src = id < where.size() ? where[id] : -1;
            std::memcpy(&moved[12 * j], src >= 0 ? &last[12 * size_t(src)] : &W.xform[12 * j], 12 * sizeof(float));
        }
        if (m_new != m_old) {
            (void)hipFree(P.rt_last_xform);
            P.rt_last_xform = nullptr;
            atlas_rt_status st = dev_alloc(P.rt_last_xform, 3 * m_new);
            if (st != ATLAS_RT_OK) return st;
        }
        ATLAS_HIP_TRY(hipMemcpy(P.rt_last_xform, moved.data(), 12 * m_new * sizeof(float), hipMemcpyHostToDevice));
    }
    // the triangle lights name instance slots and carry world positions
    atlas_rt_status st = ATLAS_RT_OK;
    if (!P.emitters.empty()) st = const_cast<atlas_rt_world&>(W).sync_host();
    if (st != ATLAS_RT_OK) return st;
    build_light_list(P, P.emitters.empty() ? nullptr : P.emitters.data(), uint32_t(P.emitters.size() / 4));
    st = upload_lights(P);
    if (st != ATLAS_RT_OK) return st;
    P.world_order = W.order;
    P.world_gen = W.generation;
    P.history_serial = W.history_serial;
    return ATLAS_RT_OK;
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
    if (params->flags & ~(ATLAS_RT_PT_RAY_BINNING | ATLAS_RT_PT_WAVEFRONT | ATLAS_RT_PT_FRAMES_IN_FLIGHT_MASK))
        return fail(ATLAS_RT_E_INVALID_ARG, "atlas_rt_pathtracer_create: unknown flag bits");
    const uint32_t slots = (params->flags & ATLAS_RT_PT_FRAMES_IN_FLIGHT_MASK) >> ATLAS_RT_PT_FRAMES_IN_FLIGHT_SHIFT;
    if (slots > kMaxFrameSlots)
        return fail(ATLAS_RT_E_INVALID_ARG, "atlas_rt_pathtracer_create: at most 4 frames in flight");
    auto* pt = new atlas_rt_pathtracer();
    PathTracer& P = pt->P;
    P.binning = (params->flags & ATLAS_RT_PT_RAY_BINNING) != 0;
    P.wavefront = (params->flags & ATLAS_RT_PT_WAVEFRONT) != 0;
    P.batch = (P.binning || P.wavefront) ? 1u : kBatchFrames;
    P.slots = slots ? slots : kDefaultFrameSlots;
    P.width = params->width;
    P.height = params->height;
    P.bounces = params->bounces;
    P.punctual.assign(lights, lights + light_count);
    P.material_count = material_count;
    P.world_gen = ~world->generation;
    P.world = world;
    P.cap = world->stack_cap();  // slot spills are sized for it; bind_world rebinds on every update
    const size_t pixels = size_t(P.width) * P.height;
    atlas_rt_status st = ATLAS_RT_OK;
    P.rays = uint32_t(pixels);
    if (st == ATLAS_RT_OK) st = dev_alloc(P.ctl, 1);
    for (uint32_t k = 0; k < P.slots && st == ATLAS_RT_OK; k++) st = alloc_slot(P, k, P.rays);
    if (st == ATLAS_RT_OK && hipEventCreateWithFlags(&P.ev_start, hipEventDisableTiming) != hipSuccess)
        st = fail(ATLAS_RT_E_HIP, "atlas_rt_pathtracer_create: hipEventCreateWithFlags failed");
    if (st == ATLAS_RT_OK) st = dev_alloc(P.accum, pixels);
    if (st == ATLAS_RT_OK) st = dev_alloc(P.hit0, pixels);
    if (st == ATLAS_RT_OK) st = dev_alloc(P.materials, material_count);
    if (st != ATLAS_RT_OK) {
        free_all(P);
        delete pt;
        return st;
    }
    FrameCtl ctl{};
    ctl.pcg_state = pcg_initial_state();
    hipError_t e = hipSuccess;
    e = hipMemcpy(P.ctl, &ctl, sizeof(ctl), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemset(P.accum, 0, pixels * sizeof(float4));
    if (e == hipSuccess) e = hipMemset(P.hit0, 0, pixels);
    if (e == hipSuccess) e = hipMemcpy(P.materials, materials, material_count * sizeof(Material), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        free_all(P);
        delete pt;
        return hip_fail(e, "atlas_rt_pathtracer_create");
    }
    ShadeArgs& A = P.shade_args;
    A.materials = P.materials;
    A.alpha = AlphaArgs{nullptr, nullptr, P.materials, nullptr, nullptr, 0u, 0u};
    A.bounces = P.bounces;
    std::memcpy(A.env, params->env_radiance, 12);
    st = bind_world(P);
    if (st != ATLAS_RT_OK) {
        free_all(P);
        delete pt;
        return st;
    }
    *out = pt;
    return ATLAS_RT_OK;
}

atlas_rt_status atlas_rt_pathtracer_destroy(atlas_rt_pathtracer* pt) {
    if (!pt) return ATLAS_RT_OK;
    (void)hipDeviceSynchronize();  // frames may still run on the slot streams
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
    PathTracer& P = pt->P;
    if (!frames) return ATLAS_RT_OK;
    atlas_rt_status bst = bind_world(P);
    if (bst == ATLAS_RT_OK) bst = apply_stack_cap(P);  // a limit set on the world since the bind
    if (bst != ATLAS_RT_OK) return bst;
    // frames rotate over the slots, each on its slot's stream: they start
    // after the caller's earlier work on `s`, and `s` waits for all of them
    const uint32_t n = P.realtime ? 1u : P.slots;  // realtime mode has one slot sized for its rays
    // the frames of earlier calls come first whatever stream they were queued on
    FrameOrder order;
    order.prev = P.hist_prev;
    order.prev2 = P.hist_prev2;
    if (P.nrc && P.query_lag) {
        // frames 0 and 1 of the call query the network as the call finds it
        if (order.prev >= 0) ATLAS_HIP_TRY(hipStreamWaitEvent(s, P.slot[order.prev].ev_train, 0));
        atlas_rt_status st = refresh_snapshots(P, s);
        if (st != ATLAS_RT_OK) return st;
    }
    ATLAS_HIP_TRY(hipEventRecord(P.ev_start, s));
    uint32_t used = 0;
    // resident frames go in batches of up to P.batch per launch
    const uint32_t per = frame_batch(P);
    for (uint32_t f = 0; f < frames;) {
        const uint32_t k = P.next_slot % n;
        if (!(used & (1u << k))) {
            ATLAS_HIP_TRY(hipStreamWaitEvent(P.slot[k].stream, P.ev_start, 0));
            used |= 1u << k;
        }
        order.index = f;
        const uint32_t nb = std::min(per, frames - f);
        atlas_rt_status st = launch_frame(P, cam, k, order, nb);
        if (st != ATLAS_RT_OK) {
            // leave `s` behind whatever was queued, then report
            for (uint32_t j = 0; j < n; j++)
                if (used & (1u << j)) (void)hipStreamWaitEvent(s, P.slot[j].ev_acc, 0);
            return st;
        }
        order.prev2 = order.prev;
        order.prev = int32_t(k);
        P.hist_prev2 = order.prev2;
        P.hist_prev = order.prev;
        P.last_slot = k;
        P.next_slot = (k + 1) % n;
        P.host_samples += int32_t(nb);
        f += nb;
    }
    for (uint32_t k = 0; k < n; k++)
        if (used & (1u << k)) ATLAS_HIP_TRY(hipStreamWaitEvent(s, P.slot[k].ev_acc, 0));
    return ATLAS_RT_OK;
}

atlas_rt_status atlas_rt_pathtracer_capture(atlas_rt_pathtracer* pt, const atlas_rt_camera* camera, uint32_t frames) {
    if (!pt || !camera) return fail(ATLAS_RT_E_INVALID_ARG, "atlas_rt_pathtracer_capture: NULL argument");
    PathTracer& P = pt->P;
    if (P.realtime || !resident_frame(P) || P.shade_args.cache.enabled)
        return fail(ATLAS_RT_E_UNSUPPORTED, "atlas_rt_pathtracer_capture: progressive frames of the resident frame "
                                            "kernel only (no realtime mode, radiance cache, binning or "
                                            "ATLAS_RT_PT_WAVEFRONT)");
    if (frames < 1 || frames > P.batch)
        return fail(ATLAS_RT_E_INVALID_ARG, "atlas_rt_pathtracer_capture: frames must be 1.." +
                                                std::to_string(P.batch));
    atlas_rt_status st = bind_world(P);
    if (st == ATLAS_RT_OK) st = apply_stack_cap(P);
    if (st != ATLAS_RT_OK) return st;
    Camera cam;
    std::memcpy(cam.loc, camera->location, 12);
    std::memcpy(cam.origin, camera->origin, 12);
    std::memcpy(cam.right, camera->right, 12);
    std::memcpy(cam.bottom, camera->bottom, 12);
    // slot 0 is idle from here on: no frame of an earlier call still runs
    ATLAS_HIP_TRY(hipDeviceSynchronize());
    if (P.graph) (void)hipGraphExecDestroy(P.graph);
    P.graph = nullptr;
    P.graph_frames = 0;
    hipStream_t cs = P.slot[0].stream;
    ATLAS_HIP_TRY(hipStreamBeginCapture(cs, hipStreamCaptureModeRelaxed));
    st = record_frames(P, cam, frames, cs);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(cs, &g);  // ends the capture whatever happened
    if (st != ATLAS_RT_OK) {
        if (g) (void)hipGraphDestroy(g);
        return st;
    }
    if (e != hipSuccess) return hip_fail(e, "hipStreamEndCapture");
    hipGraphExec_t x = nullptr;
    const hipError_t ei = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (ei != hipSuccess) return hip_fail(ei, "hipGraphInstantiate");
    P.graph = x;
    P.graph_frames = frames;
    P.graph_key = graph_key(P);
    return ATLAS_RT_OK;
}

atlas_rt_status atlas_rt_pathtracer_replay(atlas_rt_pathtracer* pt, void* stream) {
    if (!pt) return fail(ATLAS_RT_E_INVALID_ARG, "atlas_rt_pathtracer_replay: NULL path tracer");
    PathTracer& P = pt->P;
    if (!P.graph) return fail(ATLAS_RT_E_INVALID_ARG, "atlas_rt_pathtracer_replay: nothing captured");
    atlas_rt_status st = bind_world(P);
    if (st == ATLAS_RT_OK) st = apply_stack_cap(P);
    if (st != ATLAS_RT_OK) return st;
    const GraphKey now = graph_key(P);
    if (P.realtime || !resident_frame(P) || P.shade_args.cache.enabled ||
        std::memcmp(&now, &P.graph_key, sizeof(now)) != 0)
        return fail(ATLAS_RT_E_INVALID_ARG, "atlas_rt_pathtracer_replay: the path tracer changed since the capture "
                                            "(world update, stack limit, materials, lights, textures, mode or work "
                                            "counting); capture again");
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    // after every frame of earlier calls, on whatever slot it ran
    const uint32_t n = P.slots;
    for (uint32_t k = 0; k < n; k++) {
        ATLAS_HIP_TRY(hipStreamWaitEvent(s, P.slot[k].ev_begin, 0));
        ATLAS_HIP_TRY(hipStreamWaitEvent(s, P.slot[k].ev_acc, 0));
    }
    ATLAS_HIP_TRY(hipGraphLaunch(P.graph, s));
    // later frames order after these as after a frame of slot 0
    FrameSlot& S = P.slot[0];
    ATLAS_HIP_TRY(hipEventRecord(S.ev_begin, s));
    ATLAS_HIP_TRY(hipEventRecord(S.ev_train, s));
    ATLAS_HIP_TRY(hipEventRecord(S.ev_acc, s));
    P.hist_prev2 = P.hist_prev;
    P.hist_prev = 0;
    P.last_slot = 0;
    P.next_slot = 1 % n;
    P.host_samples += int32_t(P.graph_frames);
    return ATLAS_RT_OK;
}

atlas_rt_status atlas_rt_pathtracer_set_realtime(atlas_rt_pathtracer* pt, const atlas_rt_realtime_params* params) {
    if (!pt) return fail(ATLAS_RT_E_INVALID_ARG, "atlas_rt_pathtracer_set_realtime: NULL path tracer");
    PathTracer& P = pt->P;
    if (params && P.nrc)
        return fail(ATLAS_RT_E_UNSUPPORTED, "atlas_rt_pathtracer_set_realtime: not with a radiance cache");
    if (params && (params->samples_per_frame < 1 || params->samples_per_frame > 64 ||
                   params->history_length_max < 1 || !(params->max_radiance > 0.0f)))
        return fail(ATLAS_RT_E_INVALID_ARG, "atlas_rt_pathtracer_set_realtime: need 1 <= samples_per_frame <= 64, "
                                            "history_length_max >= 1, max_radiance > 0");
    atlas_rt_status bst = bind_world(P);
    if (bst != ATLAS_RT_OK) return bst;
    ATLAS_HIP_TRY(hipDeviceSynchronize());
    free_realtime(P);
    const size_t pixels = size_t(P.width) * P.height;
    const uint32_t rays = uint32_t(pixels * (params ? size_t(params->samples_per_frame) : 1u));
    P.rays = rays;
    P.next_slot = P.last_slot = 0;
    P.hist_prev = P.hist_prev2 = -1;
    // realtime frames run one at a time; progressive ones in batches
    atlas_rt_status st = alloc_slot_rays(P, 0, rays, params ? 1u : P.batch);
    if (st != ATLAS_RT_OK || !params) return st;
    const size_t inst = P.world->instances.size();
    if (st == ATLAS_RT_OK) st = dev_alloc(P.rt_frame, pixels);
    for (int k = 0; k < 2 && st == ATLAS_RT_OK; k++) {
        st = dev_alloc(P.rt_radiance[k], pixels);
        if (st == ATLAS_RT_OK) st = dev_alloc(P.rt_depth[k], pixels);
        if (st == ATLAS_RT_OK) st = dev_alloc(P.rt_velocity[k], pixels);
        if (st == ATLAS_RT_OK) st = dev_alloc(P.rt_normal[k], pixels);
        if (st == ATLAS_RT_OK) st = dev_alloc(P.rt_material[k], pixels);
    }
    if (st == ATLAS_RT_OK) st = dev_alloc(P.rt_last_xform, 3 * inst);
    if (st != ATLAS_RT_OK) {
        free_realtime(P);
        return st;
    }
    // empty history: zero radiance and length, G-buffer of zeros
    for (int k = 0; k < 2; k++) {
        ATLAS_HIP_TRY(hipMemset(P.rt_radiance[k], 0, pixels * sizeof(float4)));
        ATLAS_HIP_TRY(hipMemset(P.rt_depth[k], 0, pixels * sizeof(float)));
        ATLAS_HIP_TRY(hipMemset(P.rt_velocity[k], 0, pixels * sizeof(uint32_t)));
        ATLAS_HIP_TRY(hipMemset(P.rt_normal[k], 0, pixels * sizeof(uint32_t)));
        ATLAS_HIP_TRY(hipMemset(P.rt_material[k], 0, pixels * sizeof(uint16_t)));
    }
    ATLAS_HIP_TRY(hipMemset(P.rt_frame, 0, pixels * sizeof(float4)));
    // no motion before the first frame
    ATLAS_HIP_TRY(hipMemcpy(P.rt_last_xform, P.world->d_xform, 3 * inst * sizeof(float4), hipMemcpyDeviceToDevice));
    P.rt = *params;
    P.rt_cur = 0;
    P.rt_frames = 0;
    P.realtime = true;
    return ATLAS_RT_OK;
}

atlas_rt_status atlas_rt_pathtracer_render_realtime(atlas_rt_pathtracer* pt, const atlas_rt_camera* camera,
                                                    const atlas_rt_frame_matrices* matrices, void* stream) {
    if (!pt || !camera || !matrices) return fail(ATLAS_RT_E_INVALID_ARG, "atlas_rt_pathtracer_render_realtime: NULL argument");
    if (!pt->P.realtime)
        return fail(ATLAS_RT_E_INVALID_ARG, "atlas_rt_pathtracer_render_realtime: realtime mode is off "
                                            "(atlas_rt_pathtracer_set_realtime)");
    Camera cam;
    std::memcpy(cam.loc, camera->location, 12);
    std::memcpy(cam.origin, camera->origin, 12);
    std::memcpy(cam.right, camera->right, 12);
    std::memcpy(cam.bottom, camera->bottom, 12);
    PathTracer& P = pt->P;
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    atlas_rt_status bst = bind_world(P);
    if (bst == ATLAS_RT_OK) bst = apply_stack_cap(P);
    if (bst != ATLAS_RT_OK) return bst;
    // slot 0's buffers and the PCG stream: after the frames before, on any stream
    if (P.hist_prev >= 0) ATLAS_HIP_TRY(hipStreamWaitEvent(s, P.slot[P.hist_prev].ev_acc, 0));
    const atlas_rt_status st = launch_frame_realtime(P, cam, *matrices, s);
    if (st != ATLAS_RT_OK) return st;
    // the frames after this one wait for all of it
    FrameSlot& S = P.slot[0];
    ATLAS_HIP_TRY(hipEventRecord(S.ev_begin, s));
    ATLAS_HIP_TRY(hipEventRecord(S.ev_train, s));
    ATLAS_HIP_TRY(hipEventRecord(S.ev_acc, s));
    P.hist_prev = P.hist_prev2 = 0;
    return ATLAS_RT_OK;
}

atlas_rt_status atlas_rt_pathtracer_read_realtime(atlas_rt_pathtracer* pt, uint32_t buffer, void* d_out, void* stream) {
    if (!pt || !d_out) return fail(ATLAS_RT_E_INVALID_ARG, "atlas_rt_pathtracer_read_realtime: NULL argument");
    PathTracer& P = pt->P;
    if (!P.realtime) return fail(ATLAS_RT_E_INVALID_ARG, "atlas_rt_pathtracer_read_realtime: realtime mode is off");
    const size_t pixels = size_t(P.width) * P.height;
    const uint32_t c = P.rt_cur;
    const void* src = nullptr;
    size_t bytes = 0;
    switch (buffer) {
        case ATLAS_RT_RT_RESOLVED: src = P.rt_radiance[c]; bytes = pixels * sizeof(float4); break;
        case ATLAS_RT_RT_FRAME: src = P.rt_frame; bytes = pixels * sizeof(float4); break;
        case ATLAS_RT_RT_DEPTH: src = P.rt_depth[c]; bytes = pixels * sizeof(float); break;
        case ATLAS_RT_RT_VELOCITY: src = P.rt_velocity[c]; bytes = pixels * sizeof(uint32_t); break;
        case ATLAS_RT_RT_NORMAL: src = P.rt_normal[c]; bytes = pixels * sizeof(uint32_t); break;
        case ATLAS_RT_RT_MATERIAL: src = P.rt_material[c]; bytes = pixels * sizeof(uint16_t); break;
        default: return fail(ATLAS_RT_E_INVALID_ARG, "atlas_rt_pathtracer_read_realtime: unknown buffer");
    }
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    ATLAS_HIP_TRY(hipMemcpyAsync(d_out, src, bytes, hipMemcpyDeviceToDevice, s));
    ATLAS_HIP_TRY(hipStreamSynchronize(s));
    return ATLAS_RT_OK;
}

atlas_rt_status atlas_rt_pathtracer_set_textures(atlas_rt_pathtracer* pt, const atlas_rt_texture* textures,
                                                 uint32_t count) {
    if (!pt || (count && !textures)) return fail(ATLAS_RT_E_INVALID_ARG, "atlas_rt_pathtracer_set_textures: NULL argument");
    PathTracer& P = pt->P;
    std::vector<TexDesc> desc(count);
    size_t bytes = 0;
    for (uint32_t t = 0; t < count; t++) {
        const atlas_rt_texture& T = textures[t];
        if (!T.width || !T.height || !T.texels)
            return fail(ATLAS_RT_E_INVALID_ARG, "atlas_rt_pathtracer_set_textures: empty texture");
        if (T.channels > 4) return fail(ATLAS_RT_E_INVALID_ARG, "atlas_rt_pathtracer_set_textures: 1 to 4 channels");
        const uint32_t ch = T.channels ? T.channels : 1u;
        if (bytes + size_t(T.width) * T.height * ch > 0xFFFFFFFFull)
            return fail(ATLAS_RT_E_INVALID_ARG, "atlas_rt_pathtracer_set_textures: more than 4 GiB of texels");
        desc[t] = TexDesc{uint32_t(bytes), T.width, T.height, ch};
        bytes += size_t(T.width) * T.height * ch;
    }
    ATLAS_HIP_TRY(hipDeviceSynchronize());
    if (P.tex) (void)hipFree(P.tex);
    if (P.texels) (void)hipFree(P.texels);
    P.tex = nullptr;
    P.texels = nullptr;
    P.tex_count = 0;
    P.shade_args.alpha.tex = nullptr;
    P.shade_args.alpha.texels = nullptr;
    P.shade_args.alpha.tex_count = 0;
    if (!count) return ATLAS_RT_OK;
    atlas_rt_status st = dev_alloc(P.tex, count);
    if (st == ATLAS_RT_OK) st = dev_alloc(P.texels, bytes);
    if (st != ATLAS_RT_OK) return st;
    ATLAS_HIP_TRY(hipMemcpy(P.tex, desc.data(), count * sizeof(TexDesc), hipMemcpyHostToDevice));
    for (uint32_t t = 0; t < count; t++)
        ATLAS_HIP_TRY(hipMemcpy(P.texels + desc[t].offset, textures[t].texels,
                                size_t(textures[t].width) * textures[t].height * desc[t].channels,
                                hipMemcpyHostToDevice));
    P.tex_count = count;
    P.shade_args.alpha.tex = P.tex;
    P.shade_args.alpha.texels = P.texels;
    P.shade_args.alpha.tex_count = count;
    return ATLAS_RT_OK;
}

atlas_rt_status atlas_rt_pathtracer_set_emissive_lights(atlas_rt_pathtracer* pt, const float* emitters,
                                                        uint32_t count) {
    if (!pt) return fail(ATLAS_RT_E_INVALID_ARG, "atlas_rt_pathtracer_set_emissive_lights: NULL handle");
    PathTracer& P = pt->P;
    if (emitters && count && count != P.material_count)
        return fail(ATLAS_RT_E_SIZE_MISMATCH,
                    "atlas_rt_pathtracer_set_emissive_lights: one emitter per material table entry expected");
    ATLAS_HIP_TRY(hipDeviceSynchronize());
    if (emitters && count)
        P.emitters.assign(emitters, emitters + 4 * size_t(count));
    else
        P.emitters.clear();
    atlas_rt_status st = bind_world(P);  // after an update this also builds the list
    if (st != ATLAS_RT_OK) return st;
    // the instance transforms of the world's last update or refit
    st = const_cast<atlas_rt_world*>(P.world)->sync_host();
    if (st != ATLAS_RT_OK) return st;
    build_light_list(P, P.emitters.empty() ? nullptr : P.emitters.data(), uint32_t(P.emitters.size() / 4));
    return upload_lights(P);
}

atlas_rt_status atlas_rt_pathtracer_lights(const atlas_rt_pathtracer* pt, atlas_rt_gpu_light* out, uint32_t cap,
                                           uint32_t* count) {
    if (!pt || !count || (cap && !out)) return fail(ATLAS_RT_E_INVALID_ARG, "atlas_rt_pathtracer_lights: NULL argument");
    const auto& L = pt->P.light_list;
    *count = uint32_t(L.size());
    std::memcpy(out, L.data(), std::min<size_t>(cap, L.size()) * sizeof(Light));
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

atlas_rt_status atlas_rt_pathtracer_status(atlas_rt_pathtracer* pt, void* stream) {
    if (!pt) return fail(ATLAS_RT_E_INVALID_ARG, "atlas_rt_pathtracer_status: NULL path tracer");
    ATLAS_HIP_TRY(hipStreamSynchronize(reinterpret_cast<hipStream_t>(stream)));
    ATLAS_HIP_TRY(hipDeviceSynchronize());  // frames in flight run on the slot streams
    atlas_rt_status out = ATLAS_RT_OK;
    for (uint32_t k = 0; k < kMaxFrameSlots; k++) {
        const atlas_rt_status st = take_overflow_flag(pt->P.slot[k].flags.ptr, nullptr, "atlas_rt_pathtracer_status");
        if (st != ATLAS_RT_OK) out = st;
    }
    return out;
}

atlas_rt_status atlas_rt_pathtracer_sample_count(const atlas_rt_pathtracer* pt, int32_t* count) {
    if (!pt || !count) return fail(ATLAS_RT_E_INVALID_ARG, "atlas_rt_pathtracer_sample_count: NULL argument");
    *count = pt->P.host_samples;
    return ATLAS_RT_OK;
}

atlas_rt_status atlas_rt_pathtracer_count_work(atlas_rt_pathtracer* pt, int32_t enable) {
    if (!pt) return fail(ATLAS_RT_E_INVALID_ARG, "atlas_rt_pathtracer_count_work: NULL path tracer");
    PathTracer& P = pt->P;
    ATLAS_HIP_TRY(hipDeviceSynchronize());
    if (P.stats) (void)hipFree(P.stats);
    P.stats = nullptr;
    P.stats_partial = false;
    if (!enable) return ATLAS_RT_OK;
    atlas_rt_status st = dev_alloc(P.stats, kWorkCounters);
    if (st != ATLAS_RT_OK) return st;
    ATLAS_HIP_TRY(hipMemset(P.stats, 0, kWorkCounters * sizeof(unsigned long long)));
    return ATLAS_RT_OK;
}

atlas_rt_status atlas_rt_pathtracer_work_counts(atlas_rt_pathtracer* pt, uint64_t* counts) {
    if (!pt || !counts) return fail(ATLAS_RT_E_INVALID_ARG, "atlas_rt_pathtracer_work_counts: NULL argument");
    const PathTracer& P = pt->P;
    if (!P.stats) return fail(ATLAS_RT_E_INVALID_ARG, "atlas_rt_pathtracer_work_counts: counting is off");
    if (P.stats_partial)
        return fail(ATLAS_RT_E_UNSUPPORTED, "atlas_rt_pathtracer_work_counts: some frames ran bounce-by-bounce "
                                            "launches or transparency walks, which are not counted");
    unsigned long long c[kWorkCounters];
    ATLAS_HIP_TRY(hipDeviceSynchronize());
    ATLAS_HIP_TRY(hipMemcpy(c, P.stats, sizeof(c), hipMemcpyDeviceToHost));
    for (int i = 0; i < kWorkCounters; i++) counts[i] = uint64_t(c[i]);
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
    if (P.realtime) return fail(ATLAS_RT_E_UNSUPPORTED, "atlas_rt_pathtracer_set_cache: not in realtime mode");
    if (P.bounces < 2) return fail(ATLAS_RT_E_INVALID_ARG, "atlas_rt_pathtracer_set_cache: needs bounces >= 2");
    if (params->train_stride < 1 || params->train_steps > 64)
        return fail(ATLAS_RT_E_INVALID_ARG, "atlas_rt_pathtracer_set_cache: need train_stride >= 1, train_steps <= 64");
    if (params->query_lag > 1 || params->reserved)
        return fail(ATLAS_RT_E_INVALID_ARG, "atlas_rt_pathtracer_set_cache: query_lag is 0 or 1, reserved is 0");
    int dev = -1;
    ATLAS_HIP_TRY(hipGetDevice(&dev));
    if (dev != cache->device) return fail(ATLAS_RT_E_INVALID_ARG, "atlas_rt_pathtracer_set_cache: cache on another device");
    const size_t pixels = size_t(P.width) * P.height;
    // about pixels / stride training paths per frame; the hash selection is
    // binomial, so leave twice that (the surplus pixels render normally)
    const size_t cap = std::min(pixels, 2 * (pixels / params->train_stride) + 256);
    const size_t recs = cap * size_t(P.bounces + 1);
    atlas_rt_status st = ATLAS_RT_OK;
    for (uint32_t k = 0; k < P.slots && st == ATLAS_RT_OK; k++) {
        FrameSlot& S = P.slot[k];
        st = dev_alloc(S.rec, 5 * recs);
        if (st == ATLAS_RT_OK) st = dev_alloc(S.q_buf, 3 * pixels);
        if (st == ATLAS_RT_OK) st = dev_alloc(S.q_meta, pixels);
        if (st == ATLAS_RT_OK) st = dev_alloc(S.q_out, pixels);
        if (st == ATLAS_RT_OK) st = dev_alloc(S.s_q, 3 * recs);
        if (st == ATLAS_RT_OK) st = dev_alloc(S.s_t, recs);
        if (st == ATLAS_RT_OK) st = dev_alloc(S.train_list, cap);
    }
    for (int j = 0; j < 2 && params->query_lag && st == ATLAS_RT_OK; j++) {
        st = dev_alloc(P.snap_table[j], cache->table_params);
        if (st == ATLAS_RT_OK) st = dev_alloc(P.snap_wn[j], size_t(nrc::kWnSize));
    }
    if (st != ATLAS_RT_OK) {
        free_cache_buffers(P);
        return st;
    }
    P.nrc = cache;
    P.train_steps = params->train_steps;
    P.train_cap = uint32_t(cap);
    P.query_lag = params->query_lag;
    // the slot's record and query buffers are bound per frame
    CacheArgs& C = P.shade_args.cache;
    C.train_cap = uint32_t(cap);
    C.train_stride = params->train_stride;
    C.enabled = 1;
    return ATLAS_RT_OK;
}

atlas_rt_status atlas_rt_pathtracer_cache_stats(atlas_rt_pathtracer* pt, uint32_t* train_samples, uint32_t* queries,
                                                float* loss) {
    if (!pt) return fail(ATLAS_RT_E_INVALID_ARG, "atlas_rt_pathtracer_cache_stats: NULL path tracer");
    // the last frame's statistics
    FrameCtl ctl;
    ATLAS_HIP_TRY(hipDeviceSynchronize());
    ATLAS_HIP_TRY(hipMemcpy(&ctl, pt->P.slot[pt->P.last_slot].ctl, sizeof(ctl), hipMemcpyDeviceToHost));
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
    // the last frame's samples
    const FrameSlot& S = P.slot[P.last_slot];
    FrameCtl ctl;
    ATLAS_HIP_TRY(hipDeviceSynchronize());
    ATLAS_HIP_TRY(hipMemcpy(&ctl, S.ctl, sizeof(ctl), hipMemcpyDeviceToHost));
    const uint32_t n = std::min(std::min(ctl.train_samples, P.train_cap * uint32_t(P.bounces + 1)), max_count);
    if (n && (!d_queries || !d_targets))
        return fail(ATLAS_RT_E_INVALID_ARG, "atlas_rt_pathtracer_cache_samples: NULL buffer");
    if (n) {
        ATLAS_HIP_TRY(hipMemcpy(d_queries, S.s_q, size_t(n) * 3 * sizeof(float4), hipMemcpyDeviceToDevice));
        ATLAS_HIP_TRY(hipMemcpy(d_targets, S.s_t, size_t(n) * sizeof(float4), hipMemcpyDeviceToDevice));
    }
    *count = n;
    return ATLAS_RT_OK;
}

}  // extern "C"
