// Host-side path tracer object and the argument block of the shade kernel.
#pragma once

#include <vector>

#include "pt_types.h"
#include "rt_internal.h"
#include "nrc.h"
#include "world.h"

namespace atlas_rt {
namespace pt {

struct ShadeArgs {
    WorldDev world;
    const float4* xform;          // object-to-world rows, 3 float4 per packed instance
    const float4* const* shade;   // per BLAS: 3 float4 per slot (n0 + material, n1 + opacity, n2)
    const float4* const* tangent; // per BLAS: 2 float4 per slot (tangent, bitangent), or null
    const float4* const* color;   // per BLAS: 3 float4 per slot (vertex colours), or null
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

// One frame's working state.  Progressive rendering keeps several frames in
// flight, each on its own stream with its own set of these buffers: the last
// bounces of a frame trace few rays and wait on their traversal latency, so
// the next frame's first bounces fill the GPU meanwhile.  Events keep the
// results those of frames run one after another (see launch_frame).
struct FrameSlot {
    FrameCtl* ctl = nullptr;      // this frame's seeds, counts and cache statistics
    float4* paths[2] = {nullptr, nullptr};
    int4* hits = nullptr;
    float4* shadow = nullptr;
    float4* Lpix = nullptr;
    Scratch spill;
    Scratch flags;
    float4* binned = nullptr;
    Scratch bin_ws;
    // resident launches: path numbers in pull order and the order's block
    // counts (k_order_* in pt_kernels.hip), sized for rays x frames paths
    uint32_t* order = nullptr;
    uint32_t* order_cnt = nullptr;
    uint8_t* order_cls = nullptr;
    size_t order_cap = 0;  // paths
    // radiance cache
    float4* rec = nullptr;        // training records: 5 float4 each (query x3, beta, added radiance)
    uint32_t* train_list = nullptr;  // resident frames: training pixels
    float4* q_buf = nullptr;      // render-path queries, 3 float4 each
    float4* q_meta = nullptr;
    float4* q_out = nullptr;
    float4* s_q = nullptr;        // training samples: queries (3 float4) and targets
    float4* s_t = nullptr;
    hipStream_t stream = nullptr;
    hipEvent_t ev_begin = nullptr;   // after k_frame_begin: the frame's seeds are drawn
    hipEvent_t ev_train = nullptr;   // after the frame's last training step
    hipEvent_t ev_acc = nullptr;     // after the frame's accumulation
};

constexpr uint32_t kMaxFrameSlots = 4;
constexpr uint32_t kDefaultFrameSlots = 3;
// Progressive frames one resident launch renders: the launch's tail (its
// slowest paths) is paid once per batch, not once per frame.
constexpr uint32_t kBatchFrames = 8;
// The resident frame kernel's range counters (kPullRanges) live in the
// slot's control blocks.
static_assert(kPullRanges <= kBatchFrames, "the range counters live in the control blocks");
// Counters of atlas_rt_pathtracer_work_counts.
constexpr int kWorkCounters = 6;

// The resident frame kernel's launch shape: 256-lane groups, a 24-entry LDS
// stack per lane (deeper entries spill to the slot's spill rows, one column
// per resident lane), and three or four groups per CU, one wave per SIMD
// each.  Three leave a lane 168 VGPRs (128 at four, with about twice the
// register spills) and win where most paths end early; four hide more
// memory latency where most paths are shaded at every bounce.  After one
// warm-up launch, kOccTrials plain launches alternate the two, timed, and
// the next one waits for them once and keeps the faster (resident_shape).
constexpr int kResidentBlock = 256;
constexpr int kResidentLds = 24;
constexpr uint32_t kResidentGroupsPerCu = 4;  // the most groups; spill rows are sized for it
constexpr uint32_t kOccTrials = 4;  // two pairs
// Path numbers per block of the path-order kernels (k_order_*).
constexpr uint32_t kOrderSpan = 2048;
static_assert(kOccTrials % 2 == 0, "trials run in pairs");
// The grid-shaped kernels' stack: 128-lane groups, 32 LDS entries.
constexpr int kGridBlock = 128;
constexpr int kGridLds = 32;
// Spill bytes a slot needs for `rays` paths with `cap` stack entries, for
// either launch shape.
size_t slot_spill_bytes(uint32_t rays, uint32_t cap);

// What the kernels of a captured frame batch were given: slot 0's argument
// block and buffers, the stack bound and the work counters.
struct GraphKey {
    ShadeArgs args;
    const void* spill;
    const void* flags;
    const void* accum;
    const void* stats;
    uint32_t cap;
    uint64_t world_gen;
};

struct PathTracer {
    uint32_t width = 0, height = 0;
    int32_t bounces = 0;
    int32_t light_count = 0;
    const atlas_rt_world* world = nullptr;
    // The world's device arrays as of its update number world_gen; every
    // call that launches work rebinds them when the world has been updated
    // since (bind_world).
    uint64_t world_gen = 0;
    uint64_t history_serial = 0;        // the world's object history last taken
    std::vector<uint32_t> world_order;  // the world's instance order at that update
    WorldDev world_dev{};
    uint32_t cap = 0;             // traversal stack entries (the world's stack_cap())
    bool cap_bound = false;       // the slot spills are sized for cap
    // device buffers
    FrameCtl* ctl = nullptr;      // Random's PCG state, samples and frames so far
    FrameSlot slot[kMaxFrameSlots];
    uint32_t slots = 1;           // frames in flight during progressive rendering
    uint32_t next_slot = 0;       // the slot of the next progressive frame
    uint32_t last_slot = 0;       // the slot of the last frame rendered
    int32_t hist_prev = -1;       // slots of the last two progressive frames (-1: none yet):
    int32_t hist_prev2 = -1;      // the next call's first frames are ordered after them
    hipEvent_t ev_start = nullptr;
    float4* accum = nullptr;
    const float4* xform = nullptr;      // borrowed: the world's transform table
    size_t table_count = 0;             // entries of the three per-BLAS tables below
    const float4** shade_table = nullptr;
    const float4** uv_table = nullptr;
    const float4** tangent_table = nullptr;
    const float4** color_table = nullptr;
    Material* materials = nullptr;
    uint32_t material_count = 0;
    Light* lights = nullptr;
    uint32_t light_cap = 0;
    bool alpha = false;                     // some BLAS has non-opaque triangles
    std::vector<atlas_rt_light> punctual;   // the lights given at creation
    std::vector<Light> light_list;          // host copy of the device list
    std::vector<float> emitters;            // the last emissive table given (empty: none)
    TexDesc* tex = nullptr;
    uint8_t* texels = nullptr;
    uint32_t tex_count = 0;
    // ray binning (ATLAS_RT_PT_RAY_BINNING): each bounce's paths sorted into
    // the slot's `binned` by direction bin before the trace
    bool binning = false;
    // ATLAS_RT_PT_WAVEFRONT: bounce-by-bounce launches even where the
    // resident frame kernel could run
    bool wavefront = false;
    // progressive frames per resident launch (each slot holds this many
    // frames' control blocks and radiance buffers)
    uint32_t batch = 1;
    // traversal work counters of the resident frame kernel (null: off):
    // closest rays, shadow rays, hits, nodes, triangles, instances; partial
    // once a frame ran a launch shape that does not count
    unsigned long long* stats = nullptr;
    bool stats_partial = false;
    // a captured batch of progressive frames (atlas_rt_pathtracer_capture):
    // the graph and what its kernels were given, to refuse a replay after
    // the path tracer's buffers or bindings changed
    hipGraphExec_t graph = nullptr;
    uint32_t graph_frames = 0;
    GraphKey graph_key{};
    // the resident grid's groups per CU once both shapes have been timed
    // (0 before), the plain launches seen while it was open, the end of the
    // warm-up launch and each trial's start / end events, the trials' paths
    uint32_t occ = 0;
    uint32_t occ_launches = 0;
    // per pixel: whether the camera ray hit something, as the last resident
    // launch saw it; orders the next launch's paths (width x height bytes)
    uint8_t* hit0 = nullptr;
    hipEvent_t occ_ev[1 + 2 * kOccTrials] = {};
    uint64_t occ_paths[kOccTrials] = {};
    ShadeArgs shade_args{};       // every field but the slot's buffers
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
    uint32_t query_lag = 0;       // 1: frames query a snapshot of the network two frames old
    uint16_t* snap_table[2] = {nullptr, nullptr};  // FP16 table snapshots (query_lag 1)
    uint16_t* snap_wn[2] = {nullptr, nullptr};     // BF16 forward fragment snapshots
};

// Where a progressive frame sits: the slots of the two frames before it
// (-1: none yet) and its index in its render call.
struct FrameOrder {
    int32_t prev = -1;
    int32_t prev2 = -1;
    uint32_t index = 0;
};

// Progressive frames in slot `k` on that slot's stream: `frames` of them in
// one resident launch when resident_frame(P) and frames <= P.batch, else one.
atlas_rt_status launch_frame(PathTracer& P, const Camera& cam, uint32_t k, const FrameOrder& order, uint32_t frames);
// Whether frames run as the resident frame kernel (ray binning sorts the
// bounce-by-bounce queues; ATLAS_RT_PT_WAVEFRONT asks for them; with the
// radiance cache only one frame at a time).
bool resident_frame(const PathTracer& P);
// Progressive frames per launch: P.batch for plain resident frames, else 1
// (realtime mode, the radiance cache, bounce-by-bounce launches).
uint32_t frame_batch(const PathTracer& P);
// The kernels of `frames` progressive frames in slot 0 on `s`, with no event
// or wait (for graph capture): seeds, the resident frame kernel, the
// accumulation.
atlas_rt_status record_frames(PathTracer& P, const Camera& cam, uint32_t frames, hipStream_t s);
// Slot 0's argument block and buffers as a captured batch sees them.
GraphKey graph_key(const PathTracer& P);
// query_lag 1: both network snapshots from the live network, on `s`.
atlas_rt_status refresh_snapshots(PathTracer& P, hipStream_t s);
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
// The bounce loop shared by both modes: bounces + 1 x (trace, shade, shadow)
// in slot `k`.  With the cache, the inference waits for `train_ready` when it
// is not null and evaluates `net` (null: the live network).
atlas_rt_status launch_bounces(PathTracer& P, uint32_t k, hipStream_t s, hipEvent_t train_ready,
                               const nrc::NrcDev* net);
// Builds the light list (UpdateLights over the punctual lights and, with
// emitters, the triangle lights) into P.light_list.
void build_light_list(PathTracer& P, const float* emitters, uint32_t emitter_count);
void free_cache_buffers(PathTracer& P);
// Binds the world's current device arrays if an update replaced them since
// the last bind (waits for the device first): BLAS tables, transform table,
// stack bound and spill, the triangle lights, and in realtime mode the
// previous frame's transforms, carried over per caller instance index.
atlas_rt_status bind_world(PathTracer& P);
// Follows the world's stack bound (atlas_rt_world_set_stack_limit): the
// stack entries of the traces and the slot spills sized for them.
atlas_rt_status apply_stack_cap(PathTracer& P);

}  // namespace pt
}  // namespace atlas_rt

struct atlas_rt_pathtracer {
    atlas_rt::pt::PathTracer P;
};
