// Host-side path tracer object and the argument block of the shade kernel.
#pragma once

#include <vector>

#include "pt_types.h"
#include "rt_internal.h"
#include "world.h"

namespace atlas_rt {
namespace pt {

struct ShadeArgs {
    WorldDev world;
    const float4* xform;          // object-to-world rows, 3 float4 per packed instance
    const float4* const* shade;   // per BLAS: 3 float4 per slot (n0 + material, n1 + opacity, n2)
    const float4* const* tangent; // per BLAS: 2 float4 per slot (tangent, bitangent), or null
    const Material* materials;
    const Light* lights;
    int32_t light_count;
    AlphaArgs alpha;
    int32_t bounces;
    float env[3];
    FrameCtl* ctl;
    const float4* paths_in;
    float4* paths_out;
    const int4* hits;
    float4* shadow;
    float4* Lpix;                 // radiance per ray ID
    CacheArgs cache;
    RealtimeArgs rt;
};

struct PathTracer {
    uint32_t width = 0, height = 0;
    int32_t bounces = 0;
    int32_t light_count = 0;
    const atlas_rt_world* world = nullptr;
    WorldDev world_dev{};
    uint32_t cap = 0;
    // device buffers
    FrameCtl* ctl = nullptr;
    float4* paths[2] = {nullptr, nullptr};
    int4* hits = nullptr;
    float4* shadow = nullptr;
    float4* Lpix = nullptr;
    float4* accum = nullptr;
    const float4* xform = nullptr;      // borrowed: the world's transform table
    const float4** shade_table = nullptr;
    const float4** uv_table = nullptr;
    const float4** tangent_table = nullptr;
    Material* materials = nullptr;
    uint32_t material_count = 0;
    Light* lights = nullptr;
    uint32_t light_cap = 0;
    bool alpha = false;                     // some BLAS has non-opaque triangles
    std::vector<atlas_rt_light> punctual;   // the lights given at creation
    std::vector<Light> light_list;          // host copy of the device list
    TexDesc* tex = nullptr;
    uint8_t* texels = nullptr;
    uint32_t tex_count = 0;
    Scratch spill;
    Scratch flags;
    // ray binning (ATLAS_RT_PT_RAY_BINNING): each bounce's paths sorted into
    // `binned` by direction bin before the trace
    bool binning = false;
    float4* binned = nullptr;
    Scratch bin_ws;
    ShadeArgs shade_args{};
    int32_t host_samples = 0;
    uint32_t rays = 0;            // paths per frame: pixels, times the samples per frame in realtime mode
    // realtime mode (atlas_rt_pathtracer_set_realtime)
    bool realtime = false;
    atlas_rt_realtime_params rt{};
    uint32_t rt_cur = 0;          // which set of the pairs below is this frame's
    uint32_t rt_frames = 0;
    float4* rt_frame = nullptr;   // this frame's radiance, as the temporal pass fetches it
    float4* rt_radiance[2] = {nullptr, nullptr};  // resolved radiance + history length
    float* rt_depth[2] = {nullptr, nullptr};
    uint32_t* rt_velocity[2] = {nullptr, nullptr};
    uint32_t* rt_normal[2] = {nullptr, nullptr};
    uint16_t* rt_material[2] = {nullptr, nullptr};
    float4* rt_last_xform = nullptr;  // the world's transforms of the previous frame
    // radiance cache (null: plain path tracing)
    atlas_rt_nrc* nrc = nullptr;
    uint32_t train_steps = 0;
    uint32_t train_cap = 0;       // training paths per frame
    float4* rec = nullptr;        // training records: 5 float4 each (query x3, beta, added radiance)
    float4* q_buf = nullptr;      // render-path queries, 3 float4 each
    float4* q_meta = nullptr;
    float4* q_out = nullptr;
    float4* s_q = nullptr;        // training samples: queries (3 float4) and targets
    float4* s_t = nullptr;
};

atlas_rt_status launch_frame(PathTracer& P, const Camera& cam, hipStream_t s);
// The temporal resolve's inputs and output (pt_temporal.hip,
// pathtracer/temporal.csh).
struct TemporalArgs {
    uint32_t width, height;
    const float4* frame;       // this frame's radiance (FetchTexel)
    const float* depth;
    const uint32_t* velocity;  // 2 x binary16
    const uint32_t* normal;    // 2 x binary16
    const uint16_t* material;
    const float* hdepth;       // last frame's G-buffer
    const uint32_t* hnormal;
    const uint16_t* hmaterial;
    const float4* hradiance;   // inAccumImage: last frame's resolve
    float4* out;               // outAccumImage
    float history_clip_max, current_clip_factor, max_history_length;
};
atlas_rt_status launch_temporal(const TemporalArgs& T, hipStream_t s);

// One realtime frame.
atlas_rt_status launch_frame_realtime(PathTracer& P, const Camera& cam, const atlas_rt_frame_matrices& M,
                                      hipStream_t s);
// The bounce loop shared by both modes: bounces + 1 x (trace, shade, shadow).
atlas_rt_status launch_bounces(PathTracer& P, hipStream_t s);
// Builds the light list (UpdateLights over the punctual lights and, with
// emitters, the triangle lights) into P.light_list.
void build_light_list(PathTracer& P, const float* emitters, uint32_t emitter_count);
void free_cache_buffers(PathTracer& P);

}  // namespace pt
}  // namespace atlas_rt

struct atlas_rt_pathtracer {
    atlas_rt::pt::PathTracer P;
};
