// Wavefront path tracer kernels, gfx950.
//
// Reference: PathTracingRenderer::Render (src/engine/renderer/PathTracingRenderer.cpp:29-292)
// drives rayGen.csh, then bounces+1 x (traceClosest.csh, rayHit.csh).  The
// engine traces each shadow ray inline in the hit shader (rayHit.csh:327-339,
// HitAnyTransparency).  Here the bounce is split MI355X-style into separate
// wavefront kernels so that traversal and shading do not share registers:
//   frame_begin  per-bounce seeds from the device-side PCG32 state, counters
//   raygen       camera rays (tile-ordered, rayGen.csh:25-90)
//   trace        closest hit of every live path (HitClosestTransparency,
//                engine order; HitClosest when every triangle is opaque)
//   shade        EvaluateBounce (rayHit.csh:160-325): env on miss, surface
//                reconstruction, emission, one light's contribution (queued
//                as a shadow ray), BSDF sample, Russian roulette, stream
//                compaction of survivors
//   shadow       HitAnyTransparency (HitAny with the world ray restored when
//                every triangle is opaque); adds radiance * transparency,
//                clamped at bounces > 0 as EvaluateBounce does
//   accumulate   per-pixel running sum (rayHit.csh:119-132, linear HDR)
// Every count lives on the device, so a whole frame can be graph-captured.
#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>

#include <cstring>

#include "alpha.h"
#include "binning.h"
#include "nrc.h"
#include "pathtracer.h"
#include "pt_device.h"
#include "trace_world_core.h"

namespace atlas_rt {
namespace pt {

namespace {

constexpr int kBlock = 128;
constexpr int kShadeBlock = 256;  // raygen and shade: one compaction atomic per counter per block
constexpr int kLdsDepth = 32;

__device__ __forceinline__ uint32_t pcg32_next(uint64_t& state) {
    uint64_t x = state;
    const uint32_t count = uint32_t(x >> 59u);
    state = x * 6364136223846793005ull + 1442695040888963407ull;
    x ^= x >> 18;
    const uint32_t v = uint32_t(x >> 27u);
    return (v >> count) | (v << ((32u - count) & 31u));
}

// A frame's seeds and counters: Random's PCG stream is global, so the
// seeds come from the master block in frame order; the slot's block holds
// this frame's copy of the sample and frame counts.  Progressive frames
// (advance = 1) count themselves; realtime frames do not.
// Frames f = 0 .. frames - 1 of one launch take ctl[f], in frame order.
__global__ void k_frame_begin(FrameCtl* master, FrameCtl* ctl, int32_t bounces, uint32_t pixels, uint32_t advance,
                              uint32_t frames) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    uint64_t state = master->pcg_state;
    for (uint32_t f = 0; f < frames; f++) {
        FrameCtl* c = ctl + f;
        for (int32_t b = 0; b <= bounces && b < 16; b++)
            c->seeds[b] = float(pcg32_next(state)) / float(0xFFFFFFFFu) * 0.999999f;
        c->sample_count = master->sample_count;
        c->frame = master->frame;
        if (advance) {
            master->sample_count += 1;
            master->frame += 1;
        }
        c->paths[0] = pixels;
        for (int i = 1; i < 17; i++) c->paths[i] = 0;
        for (int i = 0; i < 16; i++) c->shadows[i] = 0;
        c->train_paths = 0;
        c->queries = 0;
        c->train_samples = 0;
    }
    // the resident kernel's range counters live in the first kPullRanges blocks
    for (uint32_t r = 0; r < kPullRanges; r++) ctl[r].pull = 0;
    master->pcg_state = state;
}

// Stream compaction for a whole block: one LDS atomic per wave and one
// global atomic per counter per block.  Every thread of the block must call
// it (threads with nothing to write pass want = false).
template <int N>
__device__ __forceinline__ void block_alloc(uint32_t* const (&ctr)[N], const bool (&want)[N], uint32_t (&slot)[N]) {
    __shared__ uint32_t s_cnt[N], s_base[N];
    const uint32_t tid = threadIdx.x, lane = tid & 63u;
    if (tid < uint32_t(N)) s_cnt[tid] = 0;
    __syncthreads();
    uint32_t woff[N], below[N];
#pragma unroll
    for (int k = 0; k < N; k++) {
        const uint64_t m = __ballot(want[k]);
        below[k] = uint32_t(__popcll(m & ((1ull << lane) - 1ull)));
        woff[k] = 0;
        if (m) {
            const int leader = __ffsll((unsigned long long)m) - 1;
            uint32_t off = 0;
            if (lane == uint32_t(leader)) off = atomicAdd(&s_cnt[k], uint32_t(__popcll(m)));
            woff[k] = __shfl(off, leader, 64);
        }
    }
    __syncthreads();
    if (tid < uint32_t(N) && s_cnt[tid]) s_base[tid] = atomicAdd(ctr[tid], s_cnt[tid]);
    __syncthreads();
#pragma unroll
    for (int k = 0; k < N; k++) slot[k] = want[k] ? s_base[k] + woff[k] + below[k] : 0u;
}

// Integer hash (murmur3 finaliser) choosing the frame's training pixels.
__device__ __forceinline__ uint32_t mix32(uint32_t h) {
    h ^= h >> 16;
    h *= 0x85ebca6bu;
    h ^= h >> 13;
    h *= 0xc2b2ae35u;
    h ^= h >> 16;
    return h;
}

// Resident frames with the cache: the frame's training pixels, picked as
// k_raygen picks them, listed by training index so that the resident kernel
// starts them first (they trace every bounce; started late, they would
// hold the frame long after the short render paths are done).  Every
// pixel's radiance starts at zero, with w = 1 marking a listed pixel.
__global__ __launch_bounds__(kShadeBlock) void k_train_list(uint32_t pixels, FrameCtl* __restrict__ ctl,
                                                            float4* __restrict__ Lpix, CacheArgs C, int32_t bounces) {
    const uint32_t g0 = blockIdx.x * kShadeBlock + threadIdx.x;
    const bool valid = g0 < pixels;
    uint32_t* const ctr[1] = {&ctl->train_paths};
    const bool pick[1] = {valid && mix32(g0 * 0x9e3779b1u + ctl->frame * 0x632be5abu) % C.train_stride == 0};
    uint32_t t[1];
    block_alloc<1>(ctr, pick, t);
    if (!valid) return;
    const bool listed = pick[0] && t[0] < C.train_cap;
    if (listed) {
        C.train_list[t[0]] = g0;
        for (int32_t k = 0; k <= bounces; k++) {
            const size_t r = size_t(t[0]) * (bounces + 1) + k;
            C.rec_beta[r] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            C.rec_add[r] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        }
    }
    Lpix[g0] = make_float4(0.0f, 0.0f, 0.0f, listed ? 1.0f : 0.0f);
}

// rayGen.csh:25-90 with one sample per frame (z = 0): jittered pinhole ray
// through the near-plane corner frame; the write index groups 8x8 tiles.
__global__ __launch_bounds__(kShadeBlock) void k_raygen(Camera cam, uint32_t width, uint32_t height,
                                                        FrameCtl* __restrict__ ctl, float4* __restrict__ paths,
                                                        float4* __restrict__ Lpix, CacheArgs C, int32_t bounces) {
    const uint32_t g0 = blockIdx.x * kShadeBlock + threadIdx.x;
    const bool valid = g0 < width * height;
    // radiance cache: about one pixel in train_stride traces a training path
    uint32_t* const ctr[1] = {&ctl->train_paths};
    const bool pick[1] = {valid && C.enabled &&
                          mix32(g0 * 0x9e3779b1u + ctl->frame * 0x632be5abu) % C.train_stride == 0};
    uint32_t t[1];
    if (C.enabled) block_alloc<1>(ctr, pick, t);
    if (!valid) return;
    const uint32_t g = g0;
    const uint32_t px = g % width, py = g / width;
    const float sc = float(ctl->sample_count);
    const float jx = rnd_pair(sc, 0.0f);
    const float jy = rnd_pair(sc, 1.0f);
    const float cx = (float(px) + jx) / float(width);
    const float cy = (float(py) + jy) / float(height);
    const V o = mk(cam.origin[0], cam.origin[1], cam.origin[2]);
    const V r = mk(cam.right[0], cam.right[1], cam.right[2]);
    const V b = mk(cam.bottom[0], cam.bottom[1], cam.bottom[2]);
    const V loc = mk(cam.loc[0], cam.loc[1], cam.loc[2]);
    const V dir = normalize(((o + r * cx) + b * cy) - loc);
    // 8x8 tiles, row-major tiles then row-major inside a tile
    const uint32_t tiles_x = (width + 7) / 8;
    const uint32_t tile = (py / 8) * tiles_x + (px / 8);
    const uint32_t slot_in = (py % 8) * 8 + (px % 8);
    uint32_t idx = tile * 64 + slot_in;
    // ragged right / bottom tiles are packed behind the full ones
    if ((width % 8) || (height % 8)) idx = g;
    const uint32_t pixel = py * width + px;
    uint32_t train = 0;
    if (C.enabled) {
        if (pick[0] && t[0] < C.train_cap) {
            train = t[0] + 1;
            for (int32_t k = 0; k <= bounces; k++) {
                const size_t r = size_t(t[0]) * (bounces + 1) + k;
                C.rec_beta[r] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
                C.rec_add[r] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            }
        }
    }
    paths[3 * size_t(idx)] = make_float4(loc.x, loc.y, loc.z, __uint_as_float(pixel));
    paths[3 * size_t(idx) + 1] = make_float4(dir.x, dir.y, dir.z, __uint_as_float(train));
    paths[3 * size_t(idx) + 2] = make_float4(1.0f, 1.0f, 1.0f, 0.0f);
    Lpix[pixel] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
}

// rayGen.csh with REALTIME: spf rays per pixel through the pixel centre
// shifted by a quarter of the camera jitter (rayGen.csh:39-40), ray ID =
// pixel * spf + sample; written in 8x8-tile order, the samples of a pixel
// next to each other.
__global__ __launch_bounds__(kShadeBlock) void k_raygen_rt(Camera cam, uint32_t width, uint32_t height, uint32_t spf,
                                                           float jx, float jy, float4* __restrict__ paths,
                                                           float4* __restrict__ Lpix) {
    const uint32_t g = blockIdx.x * kShadeBlock + threadIdx.x;
    if (g >= width * height * spf) return;
    const uint32_t pixel = g / spf, sample = g % spf;
    const uint32_t px = pixel % width, py = pixel / width;
    const float cx = jx * 0.25f + (float(px) + 0.5f) / float(width);
    const float cy = jy * 0.25f + (float(py) + 0.5f) / float(height);
    const V o = mk(cam.origin[0], cam.origin[1], cam.origin[2]);
    const V r = mk(cam.right[0], cam.right[1], cam.right[2]);
    const V b = mk(cam.bottom[0], cam.bottom[1], cam.bottom[2]);
    const V loc = mk(cam.loc[0], cam.loc[1], cam.loc[2]);
    const V dir = normalize(((o + r * cx) + b * cy) - loc);
    const uint32_t tiles_x = (width + 7) / 8;
    uint32_t slot = ((py / 8) * tiles_x + (px / 8)) * 64 + (py % 8) * 8 + (px % 8);
    if ((width % 8) || (height % 8)) slot = pixel;
    const size_t idx = size_t(slot) * spf + sample;
    paths[3 * idx] = make_float4(loc.x, loc.y, loc.z, __uint_as_float(g));
    paths[3 * idx + 1] = make_float4(dir.x, dir.y, dir.z, __uint_as_float(0u));
    paths[3 * idx + 2] = make_float4(1.0f, 1.0f, 1.0f, 0.0f);
    Lpix[g] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
}

__device__ __forceinline__ float glsl_max(float x, float y) { return (x < y) ? y : x; }
__device__ __forceinline__ float glsl_min(float x, float y) { return (y < x) ? y : x; }

// The frame's radiance per pixel as temporal.csh's FetchTexel reads it: one
// sample stores its radiance (rayHit.csh:133-137); more are quantised to
// 0xFFFFFFFF / spf steps of max_radiance, summed as integers (the engine's
// imageAtomicAdd, rayHit.csh:139-149) and decoded; then max(., 0).
__global__ __launch_bounds__(kBlock) void k_frame_rt(const float4* __restrict__ Lray, uint32_t pixels, uint32_t spf,
                                                     float max_radiance, float4* __restrict__ frame) {
    const uint32_t p = blockIdx.x * kBlock + threadIdx.x;
    if (p >= pixels) return;
    float c[3];
    if (spf == 1) {
        const float4 L = Lray[p];
        c[0] = L.x;
        c[1] = L.y;
        c[2] = L.z;
    } else {
        const uint32_t per = 0xFFFFFFFFu / spf;
        uint32_t sum[3] = {0u, 0u, 0u};
        for (uint32_t k = 0; k < spf; k++) {
            const float4 L = Lray[size_t(p) * spf + k];
            const float v[3] = {L.x, L.y, L.z};
            for (int a = 0; a < 3; a++) {
                const float n = glsl_min(glsl_max(v[a] / max_radiance, 0.0f), 1.0f);
                const float q = n * float(per);  // in [0, 2^31]; NaN radiance counts as 0
                const uint32_t u = q >= 0.0f ? uint32_t(q) : 0u;
                sum[a] += u < per ? u : per;
            }
        }
        for (int a = 0; a < 3; a++) c[a] = float(sum[a]) / float(per) / float(spf) * max_radiance;
    }
    frame[p] = make_float4(glsl_max(c[0], 0.0f), glsl_max(c[1], 0.0f), glsl_max(c[2], 0.0f), 0.0f);
}

// Closest hit of live paths: hit = (engine float32 t, slot, instance slot).
// ALPHA: HitClosestTransparency (candidates need opacity > 0).
template <int BLOCK, int LDS_DEPTH, bool ALPHA>
__global__ __launch_bounds__(BLOCK) void k_trace(WorldDev W, const uint32_t* __restrict__ count_ptr,
                                                 const float4* __restrict__ paths, int4* __restrict__ hits,
                                                 int32_t* __restrict__ spill, uint32_t cap,
                                                 uint32_t* __restrict__ flags, AlphaArgs A) {
    __shared__ int32_t lds[LDS_DEPTH * BLOCK];
    const uint32_t gid = blockIdx.x * BLOCK + threadIdx.x;
    if (gid >= *count_ptr) return;
    const float4 p0 = paths[3 * size_t(gid)];
    const float4 p1 = paths[3 * size_t(gid) + 1];
    WorldLane L;
    world_begin(L, make_float4(p0.x, p0.y, p0.z, 0.0f), make_float4(p1.x, p1.y, p1.z, kInf));
    int32_t slot = -1, inst = -1;
    if (!any_nan3(p1.x, p1.y, p1.z)) {
        const LaneStack<BLOCK, LDS_DEPTH> S{lds, spill, threadIdx.x, gid, gridDim.x * BLOCK};
        int r;
        if (ALPHA) {
            const AcceptOpaque acc{TriOpacity{W, A}};
            do {
                r = world_step_closest<BLOCK, LDS_DEPTH>(L, W, ATLAS_RT_MASK_ALL, S, cap, acc);
            } while (r == kContinue);
        } else {
            do {
                r = world_step_closest<BLOCK, LDS_DEPTH>(L, W, ATLAS_RT_MASK_ALL, S, cap);
            } while (r == kContinue);
        }
        if (r == kOverflow)
            atomicOr(&flags[2], 1u);
        else if (L.slot >= 0) {
            slot = L.slot;
            inst = L.inst;
        }
    }
    hits[gid] = make_int4(__float_as_int(L.hitT), slot, inst, 0);
}

__device__ __forceinline__ V row_xform(const float4* m, V p, float w) {
    // (m0 x + m1 y) + (m2 z + m3 w) per row
    return mk((m[0].x * p.x + m[0].y * p.y) + (m[0].z * p.z + m[0].w * w),
              (m[1].x * p.x + m[1].y * p.y) + (m[1].z * p.z + m[1].w * w),
              (m[2].x * p.x + m[2].y * p.y) + (m[2].z * p.z + m[2].w * w));
}

// glm mat4 (column-major, m[4 * column + row]) times (x, y, z, w), each row
// as ((m0 x + m1 y) + (m2 z + m3 w)).
__device__ __forceinline__ float4 mat4_mul(const float* m, float x, float y, float z, float w) {
    float r[4];
    for (int k = 0; k < 4; k++) r[k] = (m[k] * x + m[4 + k] * y) + (m[8 + k] * z + m[12 + k] * w);
    return make_float4(r[0], r[1], r[2], r[3]);
}

__device__ __forceinline__ uint32_t pack_half2(float a, float b) {
    return uint32_t(__half_as_ushort(__float2half_rn(a))) | (uint32_t(__half_as_ushort(__float2half_rn(b))) << 16);
}

// The realtime G-buffer of a pixel (rayHit.csh:76-117): velocity from the
// surface point's object-space position under this and last frame's
// matrices, view-space depth, EncodeNormal of the geometry normal
// (common/normalencode.hsh:3-15), the material ID.  inv / last: the
// instance's inverse rows and last object-to-world rows (identity on a miss).
__device__ __forceinline__ void write_gbuffer(const RealtimeArgs& R, uint32_t pixel, bool hit, V P, V Ng,
                                              uint32_t material, const float4* inv, const float4* last) {
    const float ox = dot4(P.x, P.y, P.z, 1.0f, inv[0]);
    const float oy = dot4(P.x, P.y, P.z, 1.0f, inv[1]);
    const float oz = dot4(P.x, P.y, P.z, 1.0f, inv[2]);
    const V lp = row_xform(last, mk(ox, oy, oz), 1.0f);
    const float4 vs = mat4_mul(R.view, P.x, P.y, P.z, 1.0f);
    const float4 pc = mat4_mul(R.proj, vs.x, vs.y, vs.z, vs.w);
    const float4 pl = mat4_mul(R.last_vp, lp.x, lp.y, lp.z, 1.0f);
    const float vx = (pl.x / pl.w - pc.x / pc.w) * 0.5f;
    const float vy = (pl.y / pl.w - pc.y / pc.w) * 0.5f;
    R.velocity[pixel] = pack_half2(vx, vy);
    R.depth[pixel] = hit ? vs.z : 1.0e12f;
    R.material[pixel] = uint16_t(material);
    const float l1 = (fabsf(Ng.x) + fabsf(Ng.y)) + fabsf(Ng.z);
    float nx = Ng.x / l1, nz = Ng.z / l1;
    if (Ng.y < 0.0f) {
        const float a = nx, b = nz;
        nx = (a >= 0.0f ? 1.0f : -1.0f) * (1.0f - fabsf(b));
        nz = (b >= 0.0f ? 1.0f : -1.0f) * (1.0f - fabsf(a));
    }
    R.normal[pixel] = pack_half2(nx, nz);
}

__constant__ float4 kIdentityRows[3] = {{1.0f, 0.0f, 0.0f, 0.0f}, {0.0f, 1.0f, 0.0f, 0.0f}, {0.0f, 0.0f, 1.0f, 0.0f}};

// Records a shaded path hands to the stream compaction of k_shade.
struct ShadeOut {
    bool query = false, shadow = false, cont = false;
    float4 q[4];   // cache query (3 float4) + (throughput, pixel)
    float4 sh[3];  // shadow ray
    float4 pa[3];  // continued path
};

// EvaluateBounce of one path: p0 = (origin, ray ID bits), p1 = (direction,
// training index bits), p2 = (throughput, 0); h = (t bits, hit slot or -1,
// instance slot, 0).
__device__ __forceinline__ void shade_path(const ShadeArgs& A, const FrameCtl* ctl, float4* Lpix, int32_t bounce,
                                           float4 p0, float4 p1, float4 p2, int4 h, ShadeOut& out) {
    // ray ID: the pixel in progressive mode, pixel * samples per frame +
    // sample in realtime mode (ray.ID, rayGen.csh:45)
    const uint32_t id = __float_as_uint(p0.w);
    // radiance cache: 0 = render path, else training path index + 1
    const uint32_t train = A.cache.enabled ? __float_as_uint(p1.w) : 0u;
    const size_t rec = train ? size_t(train - 1) * (A.bounces + 1) + bounce : 0;
    const V thr_in = mk(p2.x, p2.y, p2.z);
    V thr = thr_in;
    const V ro = mk(p0.x, p0.y, p0.z), rd = mk(p1.x, p1.y, p1.z);
    float4 acc = Lpix[id];
    V add = mk(0.0f, 0.0f, 0.0f);
    V no = ro, nd = rd;
    const bool gbuffer = A.rt.enabled && bounce == 0 && id % A.rt.spf == 0;
    if (h.y < 0) {
        if (gbuffer)
            write_gbuffer(A.rt, id / A.rt.spf, false, ro + rd * 100.0f, mk(1.0f, 0.0f, 0.0f), 65535u, kIdentityRows,
                          kIdentityRows);
        // miss: min(env * throughput, 10) (rayHit.csh:166-172)
        const V e = mk(A.env[0], A.env[1], A.env[2]) * thr;
        add = mk(fminf(e.x, 10.0f), fminf(e.y, 10.0f), fminf(e.z, 10.0f));
        thr = mk(0.0f, 0.0f, 0.0f);
    } else {
        const int32_t slot = h.y, inst = h.z;
        const float4 meta = A.world.instances[4 * size_t(inst) + 3];
        const int32_t mesh = __float_as_int(meta.x);
        const int32_t mat_off = __float_as_int(meta.y);
        const float4* M = A.xform + 3 * size_t(inst);
        const float4* vp = ld_blas(A.world.blases + mesh).vtx + 3 * size_t(slot);
        const float4* sp = A.shade[mesh] + 3 * size_t(slot);
        const float4 a0 = ldg4(vp), a1 = ldg4(vp + 1), a2 = ldg4(vp + 2);
        const float4 s0 = ldg4(sp), s1 = ldg4(sp + 1), s2 = ldg4(sp + 2);
        // TransformTriangle (raytracer/surface.hsh:46-61)
        const V v0 = row_xform(M, mk(a0.x, a0.y, a0.z), 1.0f);
        const V v1 = row_xform(M, mk(a1.x, a1.y, a1.z), 1.0f);
        const V v2 = row_xform(M, mk(a2.x, a2.y, a2.z), 1.0f);
        const V n0 = normalize(row_xform(M, mk(s0.x, s0.y, s0.z), 0.0f));
        const V n1 = normalize(row_xform(M, mk(s1.x, s1.y, s1.z), 0.0f));
        const V n2 = normalize(row_xform(M, mk(s2.x, s2.y, s2.z), 0.0f));
        Material m = A.materials[__float_as_int(s0.w) + mat_off];
        // re-intersect for the barycentrics (surface.hsh:72-78)
        const V e0 = v1 - v0, e1 = v2 - v0, s = ro - v0;
        const V p = cross(s, e0), q = cross(rd, e1);
        const float den = dot(q, e0);
        const float dist = dot(p, e1) / den;
        const float bs = dot(q, s) / den;
        const float bt = dot(p, rd) / den;
        const float br = 1.0f - bs - bt;
        // vertex colours (surface.hsh:84, 98): r c0 + s c1 + t c2, rgb only
        const float4* cp = A.color ? A.color[mesh] : nullptr;
        if (m.use_vertex_colors > 0 && cp) {
            const float4 c0 = cp[3 * size_t(slot)], c1 = cp[3 * size_t(slot) + 1], c2 = cp[3 * size_t(slot) + 2];
            m.base[0] *= (br * c0.x + bs * c1.x) + bt * c2.x;
            m.base[1] *= (br * c0.y + bs * c1.y) + bt * c2.y;
            m.base[2] *= (br * c0.z + bs * c1.z) + bt * c2.z;
        }
        // material maps (surface.hsh:86-124) at the texture coordinate of
        // these barycentrics
        const AlphaArgs& TX = A.alpha;
        float2 tc = make_float2(0.0f, 0.0f);
        if (has_texture(TX, m.base_color_texture) || has_texture(TX, m.opacity_texture) ||
            has_texture(TX, m.roughness_texture) || has_texture(TX, m.metalness_texture) ||
            has_texture(TX, m.ao_texture) || has_texture(TX, m.normal_texture)) {
            tc = uv_at(TX.uv[mesh], slot, bs, bt);
            if (m.invert_uvs > 0) tc.y = 1.0f - tc.y;
            if (has_texture(TX, m.base_color_texture)) {
                const float4 c = sample_texture(TX, m.base_color_texture, tc.x, tc.y);
                m.base[0] *= c.x;
                m.base[1] *= c.y;
                m.base[2] *= c.z;
            }
            if (has_texture(TX, m.opacity_texture)) m.opacity *= sample_texture(TX, m.opacity_texture, tc.x, tc.y).x;
            if (has_texture(TX, m.roughness_texture))
                m.roughness *= sample_texture(TX, m.roughness_texture, tc.x, tc.y).x;
            if (has_texture(TX, m.metalness_texture))
                m.metalness *= sample_texture(TX, m.metalness_texture, tc.x, tc.y).x;
            if (has_texture(TX, m.ao_texture)) m.ao *= sample_texture(TX, m.ao_texture, tc.x, tc.y).x;
        }
        Surface S;
        S.m = m;
        S.P = ro + rd * dist;
        V normal = normalize((n0 * br + n1 * bs) + n2 * bt);
        V Ng = normalize(cross(v0 - v1, v0 - v2));
        const bool flip = dot(Ng, rd) > 0.0f;
        Ng = Ng * ((flip && m.two_sided > 0) ? -1.0f : 1.0f);
        if (has_texture(TX, m.normal_texture)) {
            // TBN = (tangent, +-bitangent, interpolated normal); without a
            // tangent frame both are zero
            V tg = mk(0.0f, 0.0f, 0.0f), bg = mk(0.0f, 0.0f, 0.0f);
            const float4* tp = A.tangent ? A.tangent[mesh] : nullptr;
            if (tp) {
                const float4 ta = tp[2 * size_t(slot)], tb = tp[2 * size_t(slot) + 1];
                tg = normalize(row_xform(M, mk(ta.x, ta.y, ta.z), 0.0f));
                bg = normalize(row_xform(M, mk(tb.x, tb.y, tb.z), 0.0f));
            }
            const V bit = m.invert_uvs > 0 ? bg : -bg;
            const float4 c = sample_texture(TX, m.normal_texture, tc.x, tc.y);
            const float tx = 2.0f * c.x - 1.0f, ty = 2.0f * c.y - 1.0f, tz = 2.0f * c.z - 1.0f;
            V tn = (tg * tx + bit * ty) + normal * tz;
            tn = tn / glsl_max(sqrtf(dot(tn, tn)), 0.00001f);
            const float w = glsl_min(glsl_max(m.normal_scale, 0.0f), 1.0f);
            normal = normal * (1.0f - w) + tn * w;
        }
        if (dot(normal, Ng) < 0.0f && m.two_sided > 0) normal = -normal;
        S.V_ = -rd;
        S.N = normalize(normal);
        S.Ng = Ng;
        if (gbuffer)
            write_gbuffer(A.rt, id / A.rt.spf, true, S.P, Ng, uint32_t(m.id), A.world.instances + 4 * size_t(inst),
                          A.rt.last_xform + 3 * size_t(inst));
        if (A.cache.enabled) {
            const float4 q0 = make_float4(S.P.x, S.P.y, S.P.z, m.roughness);
            const float4 q1 = make_float4(-rd.x, -rd.y, -rd.z, 0.0f);
            const float4 q2 = make_float4(S.N.x, S.N.y, S.N.z, 0.0f);
            if (train) {
                A.cache.rec_q[3 * rec] = q0;
                A.cache.rec_q[3 * rec + 1] = q1;
                A.cache.rec_q[3 * rec + 2] = q2;
                A.cache.rec_beta[rec] = make_float4(thr_in.x, thr_in.y, thr_in.z, 1.0f);
            } else if (bounce == 1) {
                // the render path ends here with the cache's radiance estimate
                out.query = true;
                out.q[0] = q0;
                out.q[1] = q1;
                out.q[2] = q2;
                out.q[3] = make_float4(thr_in.x, thr_in.y, thr_in.z, __uint_as_float(id));
                return;
            }
        }
        const V base = mk(m.base[0], m.base[1], m.base[2]);
        S.F0 = mixv(mk(0.04f, 0.04f, 0.04f), base, m.metalness);
        S.F90 = 1.0f;
        S.L = mk(0.0f, 0.0f, 1.0f);
        S.NdotL = 0.0f;
        if ((m.emissive[0] + m.emissive[1]) + m.emissive[2] > 0.0f && bounce == 0)
            add = add + mk(m.emissive[0], m.emissive[1], m.emissive[2]);
        // EvaluateDirectLight (rayHit.csh:209-235): GetLight / SampleLight
        // (raytracer/direct.hsh:12-86).  The light's seed is the ray ID (the
        // engine uses the invocation index, which depends on the order of
        // the compaction atomics).
        if (A.light_count > 0) {
            float seed = ctl->seeds[bounce];
            const float ray_seed = float(id);
            float total = 0.0f, lpdf = 1.0f;
            int32_t sel = 0;
            for (int32_t l = 0; l < A.light_count; l++) {
                const Light& Lt = A.lights[l];
                const float brightness = dot(mk(Lt.color[0], Lt.color[1], Lt.color[2]),
                                             mk(0.33333f, 0.33333f, 0.33333f));
                float w;
                if ((__float_as_uint(Lt.data[0]) >> 28) == ATLAS_RT_LIGHT_TRIANGLE) {
                    const V pl = mk(Lt.P[0], Lt.P[1], Lt.P[2]) - S.P;
                    const float sq = dot(pl, pl);
                    const V Ld = normalize(pl);
                    const float ndl = fmaxf(dot(mk(Lt.N[0], Lt.N[1], Lt.N[2]), -Ld), 0.0f);
                    w = ((brightness * Lt.data[2]) * ndl) / sq;
                } else {
                    w = brightness;
                }
                w = clampf(w, 0.0000001f, 1.0f);
                total += w;
                if (rnd2(ray_seed, seed) < w / total) {
                    lpdf = w;
                    sel = l;
                }
            }
            const Light& Lt = A.lights[sel];
            lpdf = lpdf / total * Lt.data[1];
            float solid = 1.0f, ldist = kInf;
            if ((__float_as_uint(Lt.data[0]) >> 28) == ATLAS_RT_LIGHT_TRIANGLE) {
                const float r0 = rnd2(ray_seed, seed), r1 = rnd2(ray_seed, seed);
                const int32_t li = __float_as_int(Lt.data[3]);
                const int32_t lslot = int32_t(__float_as_uint(Lt.data[0]) & 0x0FFFFFFFu);
                const int32_t lmesh = __float_as_int(A.world.instances[4 * size_t(li) + 3].x);
                const float4* LM = A.xform + 3 * size_t(li);
                const float4* lv = ld_blas(A.world.blases + lmesh).vtx + 3 * size_t(lslot);
                const float4* ls = A.shade[lmesh] + 3 * size_t(lslot);
                const float4 b0 = ldg4(lv), b1 = ldg4(lv + 1), b2 = ldg4(lv + 2);
                const float4 c0 = ldg4(ls), c1 = ldg4(ls + 1), c2 = ldg4(ls + 2);
                // TransformTriangle + SampleTriangleLight (raytracer/lights.hsh:11-31)
                const V a0 = row_xform(LM, mk(b0.x, b0.y, b0.z), 1.0f);
                const V a1 = row_xform(LM, mk(b1.x, b1.y, b1.z), 1.0f);
                const V a2 = row_xform(LM, mk(b2.x, b2.y, b2.z), 1.0f);
                const V m0 = normalize(row_xform(LM, mk(c0.x, c0.y, c0.z), 0.0f));
                const V m1 = normalize(row_xform(LM, mk(c1.x, c1.y, c1.z), 0.0f));
                const V m2 = normalize(row_xform(LM, mk(c2.x, c2.y, c2.z), 0.0f));
                const float sr = sqrtf(r0);
                const float bu = 1.0f - sr, bv = r1 * sr, bw = (1.0f - bu) - bv;
                const V LP = (a0 * bu + a1 * bv) + a2 * bw;
                const V LN = (m0 * bu + m1 * bv) + m2 * bw;
                const float spdf = 1.0f / Lt.data[2];
                const V pl = LP - S.P;
                const float sq = dot(pl, pl);
                ldist = sqrtf(sq);
                S.L = mk(pl.x / ldist, pl.y / ldist, pl.z / ldist);
                update_surface(S);
                const float ndl = fmaxf(dot(LN, -S.L), 0.0f);
                solid = ndl / (sq * spdf);
            } else {
                S.L = normalize(-mk(Lt.N[0], Lt.N[1], Lt.N[2]));
                update_surface(S);
            }
            V refl = (eval_diffuse(S) + eval_specular(S)) * m.opacity;
            const V rad = mk(Lt.color[0], Lt.color[1], Lt.color[2]) * solid;
            if (S.NdotL > 0.0f) {
                // visibility and the bounce > 0 clamp are applied by k_shadow
                const V direct = (refl * rad) * S.NdotL / lpdf;
                const V c = thr * m.opacity * direct;
                const V so = S.P + S.N * kEpsilon;
                out.shadow = true;
                out.sh[0] = make_float4(so.x, so.y, so.z, __uint_as_float(id));
                out.sh[1] = make_float4(S.L.x, S.L.y, S.L.z, ldist - 2.0f * kEpsilon);
                out.sh[2] = make_float4(c.x, c.y, c.z, __uint_as_float(train ? uint32_t(rec) + 1u : 0u));
            }
        }
        // EvaluateIndirectLight (rayHit.csh:237-325)
        no = S.P;
        float seed = ctl->seeds[bounce];
        const float ray_seed = float(id);
        float refract = clampf(1.0f - m.opacity, 0.1f, 0.9f);
        refract = m.opacity == 1.0f ? 0.0f : refract;
        float rnd = saturate(rnd2(ray_seed, seed));
        BrdfSample B;
        bool refracted = false;
        if (rnd >= refract) {
            rnd = rnd2(ray_seed, seed);
            const V F = fresnel_schlick(S.F0, S.F90, S.NdotV);
            const float spec = clampf(dot(F, mk(0.33333f, 0.33333f, 0.33333f)), 0.1f, 0.9f);
            const float u0 = rnd2(ray_seed, seed);
            const float u1 = rnd2(ray_seed, seed);
            if (rnd < spec) {
                B = sample_specular(S, u0, u1);
                B.reflectance = B.reflectance * m.opacity;
                B.pdf *= spec;
            } else {
                B = sample_diffuse(S, u0, u1);
                B.reflectance = B.reflectance * ((1.0f - m.metalness) * m.opacity);
                B.pdf *= (1.0f - spec);
            }
            B.pdf *= (1.0f - refract);
            no = no + S.V_ * kEpsilon;
        } else {
            B.L = rd;
            B.pdf = refract;
            B.reflectance = mk(1.0f - m.opacity, 1.0f - m.opacity, 1.0f - m.opacity);
            S.NdotL = 1.0f;
            no = no - S.N * kEpsilon;
            refracted = true;
        }
        if (B.pdf > 0.0f && (B.reflectance.x + B.reflectance.y) + B.reflectance.z > 0.0f)
            thr = thr * ((B.reflectance * S.NdotL) / B.pdf);
        else
            thr = mk(0.0f, 0.0f, 0.0f);
        nd = normalize(B.L);
        thr = thr * m.ao;
        float prob = clampf(fmaxf(thr.x, fmaxf(thr.y, thr.z)), 0.01f, 0.99f);
        prob = bounce < 3 ? fminf(3.0f * prob, 1.0f) : prob;
        if (rnd2(ray_seed, seed) > prob)
            thr = mk(0.0f, 0.0f, 0.0f);
        else if (dot(nd, S.Ng) <= 0.0f && !refracted)
            thr = mk(0.0f, 0.0f, 0.0f);
        else
            thr = thr / prob;
    }
    acc.x += add.x;
    acc.y += add.y;
    acc.z += add.z;
    Lpix[id] = acc;
    if (train) A.cache.rec_add[rec] = make_float4(add.x, add.y, add.z, 0.0f);
    const float energy = (thr.x + thr.y) + thr.z;
    if (energy == 0.0f || bounce == A.bounces) return;
    out.cont = true;
    out.pa[0] = make_float4(no.x, no.y, no.z, p0.w);
    out.pa[1] = make_float4(nd.x, nd.y, nd.z, p1.w);
    out.pa[2] = make_float4(thr.x, thr.y, thr.z, 0.0f);
}

__device__ __forceinline__ void shade_one(const ShadeArgs& A, int32_t bounce, uint32_t i, ShadeOut& out) {
    shade_path(A, A.ctl, A.Lpix, bounce, A.paths_in[3 * size_t(i)], A.paths_in[3 * size_t(i) + 1],
               A.paths_in[3 * size_t(i) + 2], A.hits[i], out);
}

__global__ __launch_bounds__(kShadeBlock) void k_shade(ShadeArgs A, int32_t bounce) {
    const uint32_t i = blockIdx.x * kShadeBlock + threadIdx.x;
    FrameCtl* ctl = A.ctl;
    ShadeOut o;
    if (i < ctl->paths[bounce]) shade_one(A, bounce, i, o);
    uint32_t* const ctr[3] = {&ctl->queries, &ctl->shadows[bounce], &ctl->paths[bounce + 1]};
    const bool want[3] = {o.query, o.shadow, o.cont};
    uint32_t k[3];
    block_alloc<3>(ctr, want, k);
    if (o.query) {
        A.cache.q_buf[3 * size_t(k[0])] = o.q[0];
        A.cache.q_buf[3 * size_t(k[0]) + 1] = o.q[1];
        A.cache.q_buf[3 * size_t(k[0]) + 2] = o.q[2];
        A.cache.q_meta[k[0]] = o.q[3];
    }
    if (o.shadow) {
        A.shadow[3 * size_t(k[1])] = o.sh[0];
        A.shadow[3 * size_t(k[1]) + 1] = o.sh[1];
        A.shadow[3 * size_t(k[1]) + 2] = o.sh[2];
    }
    if (o.cont) {
        A.paths_out[3 * size_t(k[2])] = o.pa[0];
        A.paths_out[3 * size_t(k[2]) + 1] = o.pa[1];
        A.paths_out[3 * size_t(k[2]) + 2] = o.pa[2];
    }
}

// Shadow rays: HitAnyTransparency's traversal (restore every TLAS step),
// INSTANCE_MASK_SHADOW, tmin 0 (rayHit.csh:327-339).
template <int BLOCK, int LDS_DEPTH, bool ALPHA>
__global__ __launch_bounds__(BLOCK) void k_shadow(WorldDev W, const FrameCtl* __restrict__ ctl, int32_t bounce,
                                                  const float4* __restrict__ shadow, float4* __restrict__ Lpix,
                                                  int32_t* __restrict__ spill, uint32_t cap,
                                                  uint32_t* __restrict__ flags, float4* __restrict__ rec_add,
                                                  AlphaArgs A) {
    __shared__ int32_t lds[LDS_DEPTH * BLOCK];
    const uint32_t gid = blockIdx.x * BLOCK + threadIdx.x;
    if (gid >= ctl->shadows[bounce]) return;
    const float4 s0 = shadow[3 * size_t(gid)];
    const float4 s1 = shadow[3 * size_t(gid) + 1];
    const float4 c = shadow[3 * size_t(gid) + 2];
    WorldLane L;
    world_begin(L, make_float4(s0.x, s0.y, s0.z, 0.0f), make_float4(s1.x, s1.y, s1.z, s1.w));
    const LaneStack<BLOCK, LDS_DEPTH> S{lds, spill, threadIdx.x, gid, gridDim.x * BLOCK};
    int r;
    float vis;
    if (ALPHA) {
        const TriOpacity op{W, A};
        do {
            r = world_step_transparency<BLOCK, LDS_DEPTH>(L, W, ATLAS_RT_MASK_SHADOW, s1.w, S, cap, op);
        } while (r == kContinue);
        vis = L.trans;
    } else {
        do {
            r = world_step_any<BLOCK, LDS_DEPTH, true>(L, W, ATLAS_RT_MASK_SHADOW, s1.w, S, cap);
        } while (r == kContinue);
        vis = L.slot >= 0 ? 0.0f : 1.0f;
    }
    if (r == kOverflow) {
        atomicOr(&flags[2], 1u);
        return;
    }
    if (vis == 0.0f) return;  // occluded
    V cc = mk(c.x, c.y, c.z) * vis;
    if (bounce > 0) {
        // EvaluateBounce's clamp of the direct radiance (rayHit.csh:196-201)
        const float rm = fmaxf(fmaxf(cc.x, fmaxf(cc.y, cc.z)), 10.0f);
        cc = cc * (10.0f / rm);
    }
    const uint32_t pixel = __float_as_uint(s0.w);
    float4 acc = Lpix[pixel];
    acc.x += cc.x;
    acc.y += cc.y;
    acc.z += cc.z;
    Lpix[pixel] = acc;
    const uint32_t rec = __float_as_uint(c.w);
    if (rec) {
        float4 a = rec_add[rec - 1];
        a.x += cc.x;
        a.y += cc.y;
        a.z += cc.z;
        rec_add[rec - 1] = a;
    }
}

// ---------------------------------------------------------------------------
// Resident frame kernel: one path per lane from camera ray to its last
// bounce.  A resident grid of waves takes paths in the raygen kernels' tile
// order, from eight ranges of that order (one counter per XCD, RangePull).
// Each lane walks its path as
// trace (closest hit) -> shade -> shadow ray (any hit) -> next bounce, so no
// launch waits for the slowest path of a bounce before the next bounce
// starts: the frame's only tail is its slowest whole path.  Lanes that trace
// closest hits and lanes that trace shadow rays share one step body over
// opaque worlds (world_step_mixed).  Shading is batched per wave: the wave
// keeps stepping until its busy lanes have all finished a walk (kShadeDiv
// = 1; 1 / kShadeDiv of them in general), then those lanes shade or add
// their shadow ray's light together.  Paths refill idle lanes between
// batches (one atomic per wave).
//
// Every path runs the same arithmetic in the same order as in the
// bounce-by-bounce kernels (raygen, k_trace, shade_path, k_shadow), and
// everything a path writes (Lpix, the G-buffer) is its own, so the image is
// the same bit for bit.
// ---------------------------------------------------------------------------
constexpr int kFrameBlock = kResidentBlock;
constexpr int kFrameLds = kResidentLds;
constexpr uint32_t kShadeDiv = 1;

// The pixel of the k-th path slot of the raygen kernels' 8x8-tile order.
__device__ __forceinline__ uint32_t tile_pixel(uint32_t k, uint32_t width, uint32_t height) {
    if ((width % 8) || (height % 8)) return k;
    const uint32_t tiles_x = width / 8;
    const uint32_t tile = k / 64, in = k % 64;
    const uint32_t px = (tile % tiles_x) * 8 + in % 8;
    const uint32_t py = (tile / tiles_x) * 8 + in / 8;
    return py * width + px;
}

struct FrameGen {
    Camera cam;
    uint32_t width, height;
    uint32_t rays;        // paths per frame
    uint32_t frames;      // frames of this launch: frame f's paths follow frame f - 1's
    uint32_t spf;         // realtime: samples per pixel (ray ID = pixel * spf + sample); 0: progressive
    float jx, jy;         // realtime camera jitter
    const uint32_t* order;  // path numbers in pull order (null: pull order), see k_order_scatter
    uint8_t* hit0;          // per pixel: did the camera ray hit (written for the next launch's order)
};

// Sum of a lane value over the wave.
__device__ __forceinline__ uint32_t wave_sum(uint32_t v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Path order of a resident launch (longest processing time first).  A
// launch ends with its slowest paths, each walked start to end by one lane,
// so paths that will run every bounce should start early.  Whether a pixel's
// camera ray hit something in the previous launch predicts it: paths that
// miss end at the sky after one walk of the TLAS.  RangePull hands each XCD a
// contiguous range of the launch's path numbers (frame-major, tiles within a
// frame); the order puts, within every range, the paths of last launch's
// hit pixels first, each class in its own order (a stable partition).  The
// image does not depend on the order: every path writes only its own ray ID.
constexpr uint32_t kOrderBlock = 256;
static_assert(kOrderSpan % kOrderBlock == 0, "a counting block is whole rounds of the group");

struct OrderArgs {
    const uint8_t* hit0;  // per pixel
    uint8_t* cls;         // per path number: hit0 of its pixel as k_order_count read it
    uint32_t* order;      // total entries
    uint32_t* cnt;        // kPullRanges x blocks: hit paths per block, scanned in place; then kPullRanges totals
    uint32_t total, span, rays, width, height, spf, blocks;
};

__device__ __forceinline__ bool order_hit(const OrderArgs& O, uint32_t kk) {
    const uint32_t k = kk % O.rays;
    const uint32_t pixel = O.spf ? tile_pixel(k / O.spf, O.width, O.height) : tile_pixel(k, O.width, O.height);
    return O.hit0[pixel] != 0;
}

// Range r's counting block j: path numbers [lo + j * kOrderSpan, ...) of the
// range [lo, hi).
__device__ __forceinline__ void order_block(const OrderArgs& O, uint32_t r, uint32_t j, uint32_t& lo, uint32_t& a,
                                            uint32_t& z) {
    lo = r * O.span;
    const uint32_t hi = min(lo + O.span, O.total);
    a = min(lo + j * kOrderSpan, hi);
    z = min(a + kOrderSpan, hi);
}

__global__ __launch_bounds__(kOrderBlock) void k_order_count(OrderArgs O) {
    __shared__ uint32_t part[kOrderBlock / 64];
    const uint32_t r = blockIdx.y, j = blockIdx.x;
    uint32_t lo, a, z;
    order_block(O, r, j, lo, a, z);
    // the frames of other slots may write hit0 meanwhile: the scatter reads
    // the classes this pass counted
    uint32_t c = 0;
    for (uint32_t i = a + threadIdx.x; i < z; i += kOrderBlock) {
        const bool hit = order_hit(O, i);
        O.cls[i] = hit ? 1 : 0;
        c += hit ? 1u : 0u;
    }
    c = wave_sum(c);
    if ((threadIdx.x & 63u) == 0) part[threadIdx.x / 64] = c;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t t = 0;
        for (uint32_t w = 0; w < kOrderBlock / 64; w++) t += part[w];
        O.cnt[size_t(r) * O.blocks + j] = t;
    }
}

// One group per range: exclusive scan of its blocks' counts, and the total.
__global__ __launch_bounds__(kOrderBlock) void k_order_scan(OrderArgs O) {
    __shared__ uint32_t part[kOrderBlock / 64];
    __shared__ uint32_t carry;
    const uint32_t r = blockIdx.x, lane = threadIdx.x & 63u, wave = threadIdx.x / 64;
    uint32_t* c = O.cnt + size_t(r) * O.blocks;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    for (uint32_t b0 = 0; b0 < O.blocks; b0 += kOrderBlock) {
        const uint32_t j = b0 + threadIdx.x;
        const uint32_t v = j < O.blocks ? c[j] : 0u;
        uint32_t x = v;  // inclusive scan within the wave
        for (uint32_t o = 1; o < 64; o <<= 1) {
            const uint32_t y = __shfl_up(x, o, 64);
            if (lane >= o) x += y;
        }
        if (lane == 63) part[wave] = x;
        __syncthreads();
        uint32_t before = carry;
        for (uint32_t w = 0; w < wave; w++) before += part[w];
        if (j < O.blocks) c[j] = before + x - v;
        __syncthreads();
        if (threadIdx.x == kOrderBlock - 1) carry = before + x;
        __syncthreads();
    }
    if (threadIdx.x == 0) O.cnt[size_t(kPullRanges) * O.blocks + r] = carry;
}

// Stable partition of each counting block's path numbers: hit paths to the
// front of the range, the others after the range's last hit path.
__global__ __launch_bounds__(kOrderBlock) void k_order_scatter(OrderArgs O) {
    __shared__ uint32_t part[2][kOrderBlock / 64];
    const uint32_t r = blockIdx.y, j = blockIdx.x, lane = threadIdx.x & 63u, wave = threadIdx.x / 64;
    uint32_t lo, a, z;
    order_block(O, r, j, lo, a, z);
    if (a >= z) return;
    const uint32_t hits_before = O.cnt[size_t(r) * O.blocks + j];
    const uint32_t hits_range = O.cnt[size_t(kPullRanges) * O.blocks + r];
    uint32_t done_hit = hits_before, done_miss = (a - lo) - hits_before;
    const uint64_t below = (1ull << lane) - 1ull;
    for (uint32_t b = a; b < z; b += kOrderBlock) {
        const uint32_t i = b + threadIdx.x;
        const bool valid = i < z;
        const bool hit = valid && O.cls[i] != 0;
        const uint64_t mh = __ballot(hit), mm = __ballot(valid && !hit);
        if (lane == 0) {
            part[0][wave] = uint32_t(__popcll(mh));
            part[1][wave] = uint32_t(__popcll(mm));
        }
        __syncthreads();
        uint32_t bh = 0, bm = 0, th = 0, tm = 0;
        for (uint32_t w = 0; w < kOrderBlock / 64; w++) {
            if (w < wave) {
                bh += part[0][w];
                bm += part[1][w];
            }
            th += part[0][w];
            tm += part[1][w];
        }
        if (valid) {
            const uint32_t pos = hit ? lo + done_hit + bh + uint32_t(__popcll(mh & below))
                                     : lo + hits_range + done_miss + bm + uint32_t(__popcll(mm & below));
            O.order[pos] = i;
        }
        done_hit += th;
        done_miss += tm;
        __syncthreads();
    }
}

// With STATS the kernel also counts its traversal work into stats[0..5]:
// closest-hit rays, shadow rays, closest hits, inner nodes fetched (both
// levels), triangles tested and instances entered; the walks themselves are
// unchanged.  With CACHE (radiance cache, one frame per launch) the
// training paths listed by k_train_list come first, recording every vertex,
// then every other pixel in tile order, ending at its second vertex with a
// cache query as k_shade writes them.  OCC: resident groups per CU (waves
// per SIMD) the registers are allocated for.
template <bool ALPHA, bool STATS, bool CACHE, int OCC>
__global__ __launch_bounds__(kFrameBlock, OCC) void k_frame_persistent(ShadeArgs A, FrameGen G, int32_t* __restrict__ spill,
                                                                  uint32_t cap, uint32_t* __restrict__ flags,
                                                                  unsigned long long* __restrict__ stats) {
    __shared__ int32_t lds[kFrameLds * kFrameBlock];
    const uint32_t tid = threadIdx.x, lane = tid & 63u;
    const LaneStack<kFrameBlock, kFrameLds> S{lds, spill, tid, blockIdx.x * kFrameBlock + tid,
                                              gridDim.x * kFrameBlock};
    FrameCtl* const ctl0 = A.ctl;  // frame f's seeds and counts in ctl0[f]; ctl0[0].pull counts paths
    const WorldDev W = A.world;
    const AlphaArgs& AA = A.alpha;
    // with the cache, the listed training paths are pulled before the pixels
    const uint32_t listed = CACHE ? min(ctl0->train_paths, A.cache.train_cap) : 0u;
    const uint32_t total = G.rays * G.frames + listed;
    const V o = mk(G.cam.origin[0], G.cam.origin[1], G.cam.origin[2]);
    const V rgt = mk(G.cam.right[0], G.cam.right[1], G.cam.right[2]);
    const V btm = mk(G.cam.bottom[0], G.cam.bottom[1], G.cam.bottom[2]);
    const V loc = mk(G.cam.loc[0], G.cam.loc[1], G.cam.loc[2]);

    WorldLane L;
    world_begin(L, make_float4(0.0f, 0.0f, 0.0f, 0.0f), make_float4(1.0f, 0.0f, 0.0f, 0.0f));
    int phase = 0;       // 0 idle, 1 closest hit, 2 shadow ray
    bool ready = false;  // the walk of this phase has ended (result in L / rstat)
    int rstat = kDone;
    int32_t bounce = 0;
    float4 p0 = make_float4(0.0f, 0.0f, 0.0f, 0.0f), p1 = p0, p2 = p0;
    float4 shc = p0;     // the shadow ray's light before visibility
    bool pending = false;  // the path continues after its shadow ray (p0..p2 hold its next ray)
    bool exhausted = false;
    FrameCtl* ctl = ctl0;  // the frame of the lane's path
    float4* Lpix = A.Lpix;       // that frame's radiance per ray ID
    const CacheArgs& C = A.cache;
    uint32_t c_closest = 0, c_shadow = 0, c_hits = 0, c_nodes = 0, c_tris = 0, c_insts = 0;
    // paths from kPullRanges ranges of the path order, counters ctl0[r].pull
    RangePull rp;
    rp.init(total);
    static_assert(sizeof(FrameCtl) % sizeof(uint32_t) == 0, "counters are word-strided");

    for (;;) {
        if (!exhausted) {
            const uint64_t idle = __ballot(phase == 0);
            if (idle) {
                const uint32_t n = uint32_t(__popcll(idle));
                uint32_t base = 0, end = 0;
                if (lane == 0)
                    rp.take(&ctl0[0].pull, uint32_t(sizeof(FrameCtl) / sizeof(uint32_t)), total, n, base, end);
                base = __shfl(base, 0, 64);
                end = __shfl(end, 0, 64);
                rp.broadcast(0);
                if (rp.dry >= kPullRanges) exhausted = true;
                if (phase == 0) {
                    const uint32_t pull = base + uint32_t(__popcll(idle & ((1ull << lane) - 1ull)));
                    if (pull < end) {
                        const uint32_t kk = (!CACHE && G.order) ? G.order[pull] : pull;
                        // with the cache (one frame): listed training paths, then every pixel
                        const uint32_t f = CACHE ? 0u : kk / G.rays;
                        const uint32_t k = CACHE ? (kk < listed ? 0u : kk - listed) : kk - f * G.rays;
                        ctl = ctl0 + f;
                        Lpix = A.Lpix + size_t(f) * G.rays;
                        // rayGen.csh:25-90, as k_raygen / k_raygen_rt write slot k
                        uint32_t id;
                        uint32_t train = 0;
                        bool skip = false;
                        V dir;
                        if (G.spf == 0) {
                            uint32_t pixel;
                            if (CACHE && kk < listed) {
                                pixel = C.train_list[kk];
                                train = kk + 1;
                            } else {
                                pixel = tile_pixel(k, G.width, G.height);
                                // a listed pixel was traced first, as a training path
                                skip = CACHE && Lpix[pixel].w != 0.0f;
                            }
                            const uint32_t px = pixel % G.width, py = pixel / G.width;
                            // the frame's jitter, the same for every pixel (rayGen.csh:33-34)
                            const float sc = float(ctl->sample_count);
                            const float cx = (float(px) + rnd_pair(sc, 0.0f)) / float(G.width);
                            const float cy = (float(py) + rnd_pair(sc, 1.0f)) / float(G.height);
                            dir = normalize(((o + rgt * cx) + btm * cy) - loc);
                            id = pixel;
                        } else {
                            const uint32_t pixel = tile_pixel(k / G.spf, G.width, G.height);
                            const uint32_t px = pixel % G.width, py = pixel / G.width;
                            const float cx = G.jx * 0.25f + (float(px) + 0.5f) / float(G.width);
                            const float cy = G.jy * 0.25f + (float(py) + 0.5f) / float(G.height);
                            dir = normalize(((o + rgt * cx) + btm * cy) - loc);
                            id = pixel * G.spf + k % G.spf;
                        }
                        if (!skip) {
                            // k_train_list zeroed the cache mode's radiance (and its marks)
                            if (!CACHE) Lpix[id] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
                            p0 = make_float4(loc.x, loc.y, loc.z, __uint_as_float(id));
                            p1 = make_float4(dir.x, dir.y, dir.z, __uint_as_float(train));
                            p2 = make_float4(1.0f, 1.0f, 1.0f, 0.0f);
                            bounce = 0;
                            phase = 1;
                            world_begin(L, make_float4(p0.x, p0.y, p0.z, 0.0f),
                                        make_float4(p1.x, p1.y, p1.z, kInf));
                            ready = any_nan3(p1.x, p1.y, p1.z);
                            rstat = kDone;
                            if (STATS) c_closest++;
                        }
                    }
                }
            }
        }
        if (!__any(phase != 0)) {
            if (exhausted) break;
            continue;
        }
        // walk until enough lanes have ended their walk
        for (;;) {
            if (phase != 0 && !ready) {
                int r;
                if (ALPHA) {
                    if (phase == 1) {
                        const AcceptOpaque acc{TriOpacity{W, AA}};
                        r = world_step_closest<kFrameBlock, kFrameLds>(L, W, ATLAS_RT_MASK_ALL, S, cap, acc);
                    } else {
                        const TriOpacity op{W, AA};
                        r = world_step_transparency<kFrameBlock, kFrameLds>(L, W, ATLAS_RT_MASK_SHADOW, L.hitT, S,
                                                                            cap, op);
                    }
                } else {
                    if (STATS) {
                        // the step's kind, as world_step_mixed decides it
                        const bool top = L.sp + 1 < L.tlas_index;
                        if (top && L.node < 0)
                            c_insts++;
                        else if (top || L.node >= 0)
                            c_nodes++;
                        else
                            c_tris++;
                    }
                    r = world_step_mixed<kFrameBlock, kFrameLds>(L, W, phase == 1 ? ATLAS_RT_MASK_ALL
                                                                                   : ATLAS_RT_MASK_SHADOW,
                                                                 phase == 2, S, cap);
                }
                if (r != kContinue) {
                    ready = true;
                    rstat = r;
                }
            }
            const uint32_t busy = uint32_t(__popcll(__ballot(phase != 0)));
            const uint32_t done = uint32_t(__popcll(__ballot(ready)));
            if (done == busy || done * kShadeDiv >= busy) break;
        }
        if (!ready) continue;
        if (phase == 1) {
            // k_trace's hit record, then EvaluateBounce
            int32_t slot = -1, inst = -1;
            if (rstat == kOverflow)
                atomicOr(&flags[2], 1u);
            else if (L.slot >= 0) {
                slot = L.slot;
                inst = L.inst;
            }
            if (STATS) c_hits += slot >= 0;
            if (bounce == 0 && G.hit0) {
                const uint32_t id = __float_as_uint(p0.w);
                G.hit0[G.spf ? id / G.spf : id] = slot >= 0 ? 1 : 0;
            }
            ShadeOut so;
            shade_path(A, ctl, Lpix, bounce, p0, p1, p2, make_int4(__float_as_int(L.hitT), slot, inst, 0), so);
            if (CACHE) {
                // the path ends with a cache query: one counter update per wave
                const uint64_t qm = __ballot(so.query);
                if (qm) {
                    const int leader = __ffsll((unsigned long long)qm) - 1;
                    uint32_t qb = 0;
                    if (int(lane) == leader) qb = atomicAdd(&ctl->queries, uint32_t(__popcll(qm)));
                    qb = __shfl(qb, leader, 64);
                    if (so.query) {
                        const uint32_t qk = qb + uint32_t(__popcll(qm & ((1ull << lane) - 1ull)));
                        C.q_buf[3 * size_t(qk)] = so.q[0];
                        C.q_buf[3 * size_t(qk) + 1] = so.q[1];
                        C.q_buf[3 * size_t(qk) + 2] = so.q[2];
                        C.q_meta[qk] = so.q[3];
                    }
                }
            }
            pending = so.cont;
            if (so.cont) {
                p0 = so.pa[0];
                p1 = so.pa[1];
                p2 = so.pa[2];
            }
            if (so.shadow) {
                shc = so.sh[2];
                world_begin(L, make_float4(so.sh[0].x, so.sh[0].y, so.sh[0].z, 0.0f),
                            make_float4(so.sh[1].x, so.sh[1].y, so.sh[1].z, so.sh[1].w));
                phase = 2;
                ready = false;
                if (STATS) c_shadow++;
                continue;
            }
        } else {
            // k_shadow: the light of the shadow ray times its visibility
            if (rstat == kOverflow) {
                atomicOr(&flags[2], 1u);
            } else {
                const float vis = ALPHA ? L.trans : (L.slot >= 0 ? 0.0f : 1.0f);
                if (vis != 0.0f) {
                    V cc = mk(shc.x, shc.y, shc.z) * vis;
                    if (bounce > 0) {
                        const float rm = fmaxf(fmaxf(cc.x, fmaxf(cc.y, cc.z)), 10.0f);
                        cc = cc * (10.0f / rm);
                    }
                    const uint32_t id = __float_as_uint(p0.w);
                    float4 acc = Lpix[id];
                    acc.x += cc.x;
                    acc.y += cc.y;
                    acc.z += cc.z;
                    Lpix[id] = acc;
                    const uint32_t rec = CACHE ? __float_as_uint(shc.w) : 0u;
                    if (rec) {
                        // a training path's vertex takes the light too
                        float4 a = C.rec_add[rec - 1];
                        a.x += cc.x;
                        a.y += cc.y;
                        a.z += cc.z;
                        C.rec_add[rec - 1] = a;
                    }
                }
            }
        }
        if (pending) {
            bounce++;
            world_begin(L, make_float4(p0.x, p0.y, p0.z, 0.0f), make_float4(p1.x, p1.y, p1.z, kInf));
            phase = 1;
            ready = any_nan3(p1.x, p1.y, p1.z);
            rstat = kDone;
            if (STATS) c_closest++;
        } else {
            phase = 0;
            ready = false;
        }
    }
    if (STATS) {
        // every lane of the wave leaves the loop together
        const uint32_t s[6] = {wave_sum(c_closest), wave_sum(c_shadow), wave_sum(c_hits),
                               wave_sum(c_nodes),   wave_sum(c_tris),   wave_sum(c_insts)};
        if (lane == 0)
            for (int i = 0; i < 6; i++) atomicAdd(&stats[i], (unsigned long long)s[i]);
    }
}

// Render paths ended at their second vertex: add throughput x cache radiance.
__global__ __launch_bounds__(kBlock) void k_cache_apply(const FrameCtl* __restrict__ ctl,
                                                        const float4* __restrict__ meta,
                                                        const float4* __restrict__ pred, float4* __restrict__ Lpix) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= ctl->queries) return;
    const float4 m = meta[i];
    const float4 y = pred[i];
    const uint32_t pixel = __float_as_uint(m.w);
    float4 acc = Lpix[pixel];
    acc.x += m.x * fmaxf(y.x, 0.0f);
    acc.y += m.y * fmaxf(y.y, 0.0f);
    acc.z += m.z * fmaxf(y.z, 0.0f);
    Lpix[pixel] = acc;
}

// One training sample per surface vertex of every training path: the
// radiance the path added from that bounce on, over the throughput that
// reached the vertex.
__global__ __launch_bounds__(kBlock) void k_build_targets(FrameCtl* __restrict__ ctl, CacheArgs C, int32_t bounces,
                                                          float4* __restrict__ s_q, float4* __restrict__ s_t) {
    const uint32_t t = blockIdx.x * kBlock + threadIdx.x;
    if (t >= min(ctl->train_paths, C.train_cap)) return;
    uint32_t n = 0;
    for (int32_t k = 0; k <= bounces; k++) n += C.rec_beta[size_t(t) * (bounces + 1) + k].w != 0.0f;
    if (n == 0) return;
    uint32_t si = atomicAdd(&ctl->train_samples, n);  // one reservation per training path
    float sx = 0.0f, sy = 0.0f, sz = 0.0f;
    for (int32_t k = bounces; k >= 0; k--) {
        const size_t r = size_t(t) * (bounces + 1) + k;
        const float4 a = C.rec_add[r];
        sx += a.x;
        sy += a.y;
        sz += a.z;
        const float4 b = C.rec_beta[r];
        if (b.w == 0.0f) continue;
        s_q[3 * size_t(si)] = C.rec_q[3 * r];
        s_q[3 * size_t(si) + 1] = C.rec_q[3 * r + 1];
        s_q[3 * size_t(si) + 2] = C.rec_q[3 * r + 2];
        s_t[si] = make_float4(b.x > 0.0f ? sx / b.x : 0.0f, b.y > 0.0f ? sy / b.y : 0.0f,
                              b.z > 0.0f ? sz / b.z : 0.0f, 0.0f);
        si++;
    }
}

// Adds `frames` frames' radiance to the accumulation image, frame after
// frame per pixel (the order of one frame per launch).
__global__ __launch_bounds__(kBlock) void k_accumulate(const float4* __restrict__ Lpix, float4* __restrict__ accum,
                                                       uint32_t pixels, uint32_t frames) {
    const uint32_t p = blockIdx.x * kBlock + threadIdx.x;
    if (p < pixels) {
        float4 a = accum[p];
        for (uint32_t f = 0; f < frames; f++) {
            const float4 l = Lpix[size_t(f) * pixels + p];
            a.x += l.x;
            a.y += l.y;
            a.z += l.z;
            a.w += 1.0f;
        }
        accum[p] = a;
    }
}

// The shade kernel's arguments with slot k's buffers.
ShadeArgs slot_args(const PathTracer& P, uint32_t k) {
    const FrameSlot& S = P.slot[k];
    ShadeArgs A = P.shade_args;
    A.ctl = S.ctl;
    A.hits = S.hits;
    A.shadow = S.shadow;
    A.Lpix = S.Lpix;
    if (A.cache.enabled) {
        const size_t recs = size_t(P.train_cap) * size_t(P.bounces + 1);
        A.cache.rec_q = S.rec;
        A.cache.rec_beta = S.rec + 3 * recs;
        A.cache.rec_add = S.rec + 4 * recs;
        A.cache.q_buf = S.q_buf;
        A.cache.q_meta = S.q_meta;
        A.cache.train_list = S.train_list;
    }
    return A;
}

// Groups per CU for a plain resident launch.  While the shape is open, the
// first launch runs four groups untimed (it warms the caches and code); the
// next kOccTrials launches are trials in pairs, four groups then three, each
// starting after the previous launch has ended (frames in flight would
// otherwise overlap them) and timed by events t0 / t1 around the kernel.
// The launch after the trials waits for the last one to end and keeps three
// groups only if they took less time over the pairs whose two launches
// traced the same number of paths (launch tails differ with launch size).
// Launches inside a stream capture are never trials and take four groups
// while the shape is open.
uint32_t resident_shape(PathTracer& P, const FrameGen& G, hipStream_t s, hipEvent_t& t0, hipEvent_t& t1) {
    t0 = t1 = nullptr;
    if (P.occ) return P.occ;
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(s, &cs) != hipSuccess || cs != hipStreamCaptureStatusNone) return kResidentGroupsPerCu;
    const uint32_t n = P.occ_launches;
    if (n == 1 + kOccTrials) {
        P.occ = kResidentGroupsPerCu;
        if (hipEventSynchronize(P.occ_ev[2 * kOccTrials]) != hipSuccess) return P.occ;
        double t4 = 0.0, t3 = 0.0;
        for (uint32_t p = 0; p < kOccTrials / 2; p++) {
            const uint32_t a = 2 * p, b = 2 * p + 1;  // four groups, three groups
            float ma = 0.0f, mb = 0.0f;
            if (hipEventElapsedTime(&ma, P.occ_ev[1 + 2 * a], P.occ_ev[2 + 2 * a]) != hipSuccess ||
                hipEventElapsedTime(&mb, P.occ_ev[1 + 2 * b], P.occ_ev[2 + 2 * b]) != hipSuccess)
                return P.occ;
            if (P.occ_paths[a] != P.occ_paths[b]) continue;
            t4 += ma;
            t3 += mb;
        }
        if (t3 > 0.0 && t3 < t4) P.occ = 3;
        return P.occ;
    }
    for (hipEvent_t& e : P.occ_ev)
        if (!e && hipEventCreate(&e) != hipSuccess) {
            P.occ = kResidentGroupsPerCu;
            return P.occ;
        }
    P.occ_launches++;
    if (n == 0) {  // the warm-up launch: its end starts the first trial
        t1 = P.occ_ev[0];
        return kResidentGroupsPerCu;
    }
    const uint32_t i = n - 1;
    if (hipStreamWaitEvent(s, P.occ_ev[2 * i], 0) != hipSuccess) {  // the previous launch's end
        P.occ = kResidentGroupsPerCu;
        return P.occ;
    }
    t0 = P.occ_ev[1 + 2 * i];
    t1 = P.occ_ev[2 + 2 * i];
    P.occ_paths[i] = uint64_t(G.rays) * G.frames;
    return (i & 1) ? 3u : 4u;
}

// The path order of a resident launch of G (k_order_*), written to the
// slot's order buffer; false (pull order) when it does not fit the slot's
// buffers.
bool launch_path_order(PathTracer& P, FrameSlot& S, const FrameGen& G, hipStream_t s) {
    OrderArgs O;
    O.hit0 = P.hit0;
    O.cls = S.order_cls;
    O.order = S.order;
    O.cnt = S.order_cnt;
    O.total = G.rays * G.frames;
    O.span = (O.total + kPullRanges - 1) / kPullRanges;
    O.rays = G.rays;
    O.width = G.width;
    O.height = G.height;
    O.spf = G.spf;
    O.blocks = (O.span + kOrderSpan - 1) / kOrderSpan;
    if (!O.hit0 || !O.cls || !O.order || !O.cnt || O.total == 0 || size_t(G.rays) * G.frames > S.order_cap) return false;
    hipLaunchKernelGGL(k_order_count, dim3(O.blocks, kPullRanges), dim3(kOrderBlock), 0, s, O);
    hipLaunchKernelGGL(k_order_scan, dim3(kPullRanges), dim3(kOrderBlock), 0, s, O);
    hipLaunchKernelGGL(k_order_scatter, dim3(O.blocks, kPullRanges), dim3(kOrderBlock), 0, s, O);
    return true;
}

atlas_rt_status launch_resident(PathTracer& P, uint32_t k, const ShadeArgs& A, const FrameGen& G0, hipStream_t s) {
    FrameSlot& S = P.slot[k];
    // One frame's launch ends with its longest path whatever the order (C2:
    // about 0.46 ms, against 0.09 ms of work per frame), so only batches of
    // frames are ordered; every plain launch records the camera-ray hits.
    FrameGen G = G0;
    if (!A.cache.enabled) {
        G.hit0 = P.hit0;
        if (G.frames > 1 && launch_path_order(P, S, G, s)) G.order = S.order;
    }
    const uint32_t cus = uint32_t(device_cu_count());
    const uint32_t groups = cus * kResidentGroupsPerCu;
    int32_t* spill = reinterpret_cast<int32_t*>(S.spill.ptr);
    uint32_t* flags = reinterpret_cast<uint32_t*>(S.flags.ptr);
    const bool cache = A.cache.enabled;
    // the transparency walks test a whole leaf per step, and the cache mode
    // is not counted
    if (P.stats && (P.alpha || cache)) P.stats_partial = true;
    if (cache && P.alpha)
        hipLaunchKernelGGL((k_frame_persistent<true, false, true, 4>), dim3(groups), dim3(kFrameBlock), 0, s, A, G,
                           spill, P.cap, flags, nullptr);
    else if (cache)
        hipLaunchKernelGGL((k_frame_persistent<false, false, true, 4>), dim3(groups), dim3(kFrameBlock), 0, s, A, G,
                           spill, P.cap, flags, nullptr);
    else if (P.stats && !P.alpha)
        hipLaunchKernelGGL((k_frame_persistent<false, true, false, 4>), dim3(groups), dim3(kFrameBlock), 0, s, A, G,
                           spill, P.cap, flags, P.stats);
    else {
        hipEvent_t t0, t1;
        const uint32_t occ = resident_shape(P, G, s, t0, t1);
        if (t0) ATLAS_HIP_TRY(hipEventRecord(t0, s));
        const dim3 grid(cus * occ);
        if (P.alpha && occ == 3)
            hipLaunchKernelGGL((k_frame_persistent<true, false, false, 3>), grid, dim3(kFrameBlock), 0, s, A, G, spill,
                               P.cap, flags, nullptr);
        else if (P.alpha)
            hipLaunchKernelGGL((k_frame_persistent<true, false, false, 4>), grid, dim3(kFrameBlock), 0, s, A, G, spill,
                               P.cap, flags, nullptr);
        else if (occ == 3)
            hipLaunchKernelGGL((k_frame_persistent<false, false, false, 3>), grid, dim3(kFrameBlock), 0, s, A, G,
                               spill, P.cap, flags, nullptr);
        else
            hipLaunchKernelGGL((k_frame_persistent<false, false, false, 4>), grid, dim3(kFrameBlock), 0, s, A, G,
                               spill, P.cap, flags, nullptr);
        if (t1) ATLAS_HIP_TRY(hipEventRecord(t1, s));
    }
    ATLAS_HIP_TRY(hipGetLastError());
    return ATLAS_RT_OK;
}

}  // namespace

bool resident_frame(const PathTracer& P) {
    // With the cache, frames run resident one at a time; with several in
    // flight the bounce-by-bounce launches overlap their training-path tails
    return !P.binning && !P.wavefront && (!P.shade_args.cache.enabled || P.slots == 1);
}

uint32_t frame_batch(const PathTracer& P) {
    return resident_frame(P) && !P.realtime && !P.shade_args.cache.enabled ? P.batch : 1u;
}

GraphKey graph_key(const PathTracer& P) {
    GraphKey key;
    std::memset(&key, 0, sizeof(key));  // padding too: keys are compared bytewise
    key.args = slot_args(P, 0);
    key.spill = P.slot[0].spill.ptr;
    key.flags = P.slot[0].flags.ptr;
    key.accum = P.accum;
    key.stats = P.stats;
    key.cap = P.cap;
    key.world_gen = P.world_gen;
    return key;
}

atlas_rt_status record_frames(PathTracer& P, const Camera& cam, uint32_t frames, hipStream_t s) {
    FrameSlot& S = P.slot[0];
    const uint32_t pixels = P.width * P.height;
    hipLaunchKernelGGL(k_frame_begin, dim3(1), dim3(64), 0, s, P.ctl, S.ctl, P.bounces, pixels, 1u, frames);
    const FrameGen G{cam, P.width, P.height, pixels, frames, 0u, 0.0f, 0.0f};
    const atlas_rt_status rs = launch_resident(P, 0, slot_args(P, 0), G, s);
    if (rs != ATLAS_RT_OK) return rs;
    hipLaunchKernelGGL(k_accumulate, dim3((pixels + kBlock - 1) / kBlock), dim3(kBlock), 0, s, S.Lpix, P.accum,
                       pixels, frames);
    ATLAS_HIP_TRY(hipGetLastError());
    return ATLAS_RT_OK;
}

atlas_rt_status launch_bounces(PathTracer& P, uint32_t k, hipStream_t s, hipEvent_t train_ready,
                               const nrc::NrcDev* net) {
    FrameSlot& S = P.slot[k];
    const uint32_t rays = P.rays;
    if (P.stats) P.stats_partial = true;  // only the resident frame kernel counts
    const uint32_t grid_all = (rays + kBlock - 1) / kBlock;
    uint32_t* flags = reinterpret_cast<uint32_t*>(S.flags.ptr);
    const ShadeArgs base = slot_args(P, k);
    const CacheArgs& C = base.cache;
    for (int32_t b = 0; b <= P.bounces; b++) {
        float4* in = S.paths[b & 1];
        float4* out = S.paths[(b + 1) & 1];
        if (P.binning) {
            // coherence sort of this bounce's live paths (the engine sorts
            // before each trace dispatch, RayTracingHelper.cpp:304-344)
            ATLAS_HIP_TRY(launch_bin_sort(in, S.binned, 3, &S.ctl->paths[b], 0, rays, S.bin_ws.ptr, nullptr,
                                          nullptr, s));
            in = S.binned;
        }
        // with the cache only training paths outlive bounce 1
        const uint32_t grid = (C.enabled && b >= 2) ? (P.train_cap + kBlock - 1) / kBlock : grid_all;
        int32_t* spill = reinterpret_cast<int32_t*>(S.spill.ptr);
        const AlphaArgs& AA = base.alpha;
        if (P.alpha)
            hipLaunchKernelGGL((k_trace<kBlock, kLdsDepth, true>), dim3(grid), dim3(kBlock), 0, s, P.world_dev,
                               &S.ctl->paths[b], in, S.hits, spill, P.cap, flags, AA);
        else
            hipLaunchKernelGGL((k_trace<kBlock, kLdsDepth, false>), dim3(grid), dim3(kBlock), 0, s, P.world_dev,
                               &S.ctl->paths[b], in, S.hits, spill, P.cap, flags, AA);
        ShadeArgs A = base;
        A.paths_in = in;
        A.paths_out = out;
        const uint32_t shade_grid = (grid * kBlock + kShadeBlock - 1) / kShadeBlock;
        hipLaunchKernelGGL(k_shade, dim3(shade_grid), dim3(kShadeBlock), 0, s, A, b);
        if (P.light_count > 0 && P.alpha)
            hipLaunchKernelGGL((k_shadow<kBlock, kLdsDepth, true>), dim3(grid), dim3(kBlock), 0, s, P.world_dev,
                               S.ctl, b, S.shadow, S.Lpix, spill, P.cap, flags, C.rec_add, AA);
        else if (P.light_count > 0)
            hipLaunchKernelGGL((k_shadow<kBlock, kLdsDepth, false>), dim3(grid), dim3(kBlock), 0, s, P.world_dev,
                               S.ctl, b, S.shadow, S.Lpix, spill, P.cap, flags, C.rec_add, AA);
        if (C.enabled && b == 1) {
            if (train_ready) ATLAS_HIP_TRY(hipStreamWaitEvent(s, train_ready, 0));
            ATLAS_HIP_TRY(nrc::launch_infer(net ? *net : P.nrc->dev, S.q_buf, S.q_out, &S.ctl->queries, rays, s));
            hipLaunchKernelGGL(k_cache_apply, dim3(grid_all), dim3(kBlock), 0, s, S.ctl, S.q_meta, S.q_out, S.Lpix);
        }
    }
    ATLAS_HIP_TRY(hipGetLastError());
    return ATLAS_RT_OK;
}

atlas_rt_status launch_frame_realtime(PathTracer& P, const Camera& cam, const atlas_rt_frame_matrices& M,
                                      hipStream_t s) {
    const uint32_t pixels = P.width * P.height;
    const uint32_t spf = uint32_t(P.rt.samples_per_frame);
    // PathTracerRenderTarget::Swap: last frame's buffers become the history
    P.rt_cur ^= 1u;
    const uint32_t cur = P.rt_cur, hist = cur ^ 1u;
    ShadeArgs& A = P.shade_args;
    A.rt.depth = P.rt_depth[cur];
    A.rt.velocity = P.rt_velocity[cur];
    A.rt.normal = P.rt_normal[cur];
    A.rt.material = P.rt_material[cur];
    A.rt.last_xform = P.rt_last_xform;
    std::memcpy(A.rt.view, M.view, sizeof(A.rt.view));
    std::memcpy(A.rt.proj, M.projection, sizeof(A.rt.proj));
    std::memcpy(A.rt.last_vp, M.last_view_projection, sizeof(A.rt.last_vp));
    A.rt.spf = spf;
    A.rt.enabled = 1;
    // realtime frames run one at a time in slot 0 on the caller's stream
    FrameSlot& S = P.slot[0];
    hipLaunchKernelGGL(k_frame_begin, dim3(1), dim3(64), 0, s, P.ctl, S.ctl, P.bounces, P.rays, 0u, 1u);
    atlas_rt_status st;
    if (resident_frame(P)) {
        const FrameGen G{cam, P.width, P.height, P.rays, 1u, spf, M.jitter[0], M.jitter[1]};
        st = launch_resident(P, 0, slot_args(P, 0), G, s);
    } else {
        hipLaunchKernelGGL(k_raygen_rt, dim3((P.rays + kShadeBlock - 1) / kShadeBlock), dim3(kShadeBlock), 0, s, cam,
                           P.width, P.height, spf, M.jitter[0], M.jitter[1], S.paths[0], S.Lpix);
        st = launch_bounces(P, 0, s, nullptr, nullptr);
    }
    A.rt.enabled = 0;
    if (st != ATLAS_RT_OK) return st;
    hipLaunchKernelGGL(k_frame_rt, dim3((pixels + kBlock - 1) / kBlock), dim3(kBlock), 0, s, S.Lpix, pixels, spf,
                       P.rt.max_radiance, P.rt_frame);
    TemporalArgs T;
    T.width = P.width;
    T.height = P.height;
    T.frame = P.rt_frame;
    T.depth = P.rt_depth[cur];
    T.velocity = P.rt_velocity[cur];
    T.normal = P.rt_normal[cur];
    T.material = P.rt_material[cur];
    T.hdepth = P.rt_depth[hist];
    T.hnormal = P.rt_normal[hist];
    T.hmaterial = P.rt_material[hist];
    T.hradiance = P.rt_radiance[hist];
    T.out = P.rt_radiance[cur];
    T.history_clip_max = P.rt.history_clip_max;
    T.current_clip_factor = P.rt.current_clip_factor;
    T.max_history_length = float(P.rt.history_length_max);
    const atlas_rt_status ts = launch_temporal(T, s);
    if (ts != ATLAS_RT_OK) return ts;
    // this frame's transforms are the next frame's last ones
    const size_t n = 3 * P.world->instances.size();
    ATLAS_HIP_TRY(hipMemcpyAsync(P.rt_last_xform, P.world->d_xform, n * sizeof(float4), hipMemcpyDeviceToDevice, s));
    P.rt_frames++;
    ATLAS_HIP_TRY(hipGetLastError());
    return ATLAS_RT_OK;
}

// Frames in flight give the results of frames run one after another
// because everything they share is ordered by events (also across render
// calls: the first frames of a call wait for the last frames of the one
// before):
// - the master block's PCG state and counters: frame f draws its seeds after
//   frame f - 1 did (ev_begin);
// - the accumulation image: frame f adds its radiance after frame f - 1
//   (ev_acc), so every pixel sums its samples in frame order;
// - the radiance cache: frame f trains after frame f - 1 trained, and
//   queries the network after frame f - 1's training (query_lag 0) or
//   frame f - 2's, through the snapshot that training left (query_lag 1)
//   (ev_train).
// Everything else a frame touches is its slot's own.
atlas_rt_status launch_frame(PathTracer& P, const Camera& cam, uint32_t k, const FrameOrder& order, uint32_t frames) {
    FrameSlot& S = P.slot[k];
    hipStream_t s = S.stream;
    const uint32_t pixels = P.width * P.height;
    const uint32_t grid_all = (pixels + kBlock - 1) / kBlock;
    const ShadeArgs A = slot_args(P, k);
    const CacheArgs& C = A.cache;
    const FrameSlot* before = order.prev >= 0 ? &P.slot[order.prev] : nullptr;
    const FrameSlot* before2 = order.prev2 >= 0 ? &P.slot[order.prev2] : nullptr;
    const bool resident = resident_frame(P);
    if (frames > frame_batch(P)) frames = 1;
    if (before) ATLAS_HIP_TRY(hipStreamWaitEvent(s, before->ev_begin, 0));
    hipLaunchKernelGGL(k_frame_begin, dim3(1), dim3(64), 0, s, P.ctl, S.ctl, P.bounces, pixels, 1u, frames);
    ATLAS_HIP_TRY(hipEventRecord(S.ev_begin, s));
    const bool lag = C.enabled && P.query_lag;
    nrc::NrcDev snap{};
    if (lag) {
        snap = P.nrc->dev;
        snap.table = P.snap_table[order.index & 1u];
        snap.wn = P.snap_wn[order.index & 1u];
    }
    const FrameSlot* ready = lag ? before2 : before;
    if (resident) {
        // With the cache, the resident grid (every CU) starts after the frame
        // before has trained: launched beside that frame's network kernels it
        // would hold the CUs they wait for.
        if (C.enabled && before) ATLAS_HIP_TRY(hipStreamWaitEvent(s, before->ev_train, 0));
        if (C.enabled)
            hipLaunchKernelGGL(k_train_list, dim3((pixels + kShadeBlock - 1) / kShadeBlock), dim3(kShadeBlock), 0, s,
                               pixels, S.ctl, S.Lpix, C, P.bounces);
        const FrameGen G{cam, P.width, P.height, pixels, frames, 0u, 0.0f, 0.0f};
        const atlas_rt_status rs = launch_resident(P, k, A, G, s);
        if (rs != ATLAS_RT_OK) return rs;
        if (C.enabled) {
            // the render paths' cache queries, once the network they read is trained
            if (ready) ATLAS_HIP_TRY(hipStreamWaitEvent(s, ready->ev_train, 0));
            ATLAS_HIP_TRY(nrc::launch_infer(lag ? snap : P.nrc->dev, S.q_buf, S.q_out, &S.ctl->queries, P.rays, s));
            hipLaunchKernelGGL(k_cache_apply, dim3(grid_all), dim3(kBlock), 0, s, S.ctl, S.q_meta, S.q_out, S.Lpix);
        }
    } else {
        hipLaunchKernelGGL(k_raygen, dim3((pixels + kShadeBlock - 1) / kShadeBlock), dim3(kShadeBlock), 0, s, cam,
                           P.width, P.height, S.ctl, S.paths[0], S.Lpix, C, P.bounces);
        const atlas_rt_status bs = launch_bounces(P, k, s, ready ? ready->ev_train : nullptr, lag ? &snap : nullptr);
        if (bs != ATLAS_RT_OK) return bs;
    }
    if (C.enabled) {
        hipLaunchKernelGGL(k_build_targets, dim3((P.train_cap + kBlock - 1) / kBlock), dim3(kBlock), 0, s, S.ctl, C,
                           P.bounces, S.s_q, S.s_t);
        if (before) ATLAS_HIP_TRY(hipStreamWaitEvent(s, before->ev_train, 0));
        const uint32_t max_samples = P.train_cap * uint32_t(P.bounces + 1);
        for (uint32_t t = 0; t < P.train_steps; t++) {
            atlas_rt_status st = nrc::train_step(*P.nrc, S.s_q, S.s_t, &S.ctl->train_samples, max_samples, t,
                                                 P.train_steps, &S.ctl->train_loss, s);
            if (st != ATLAS_RT_OK) return st;
        }
        if (lag) {
            // the network this training left, for the frame two after this one
            const atlas_rt_nrc& N = *P.nrc;
            const uint32_t j = order.index & 1u;
            ATLAS_HIP_TRY(hipMemcpyAsync(P.snap_table[j], N.table16, size_t(N.table_params) * sizeof(uint16_t),
                                         hipMemcpyDeviceToDevice, s));
            ATLAS_HIP_TRY(hipMemcpyAsync(P.snap_wn[j], N.wn, size_t(nrc::kWnSize) * sizeof(uint16_t),
                                         hipMemcpyDeviceToDevice, s));
        }
    }
    ATLAS_HIP_TRY(hipEventRecord(S.ev_train, s));
    if (before) ATLAS_HIP_TRY(hipStreamWaitEvent(s, before->ev_acc, 0));
    hipLaunchKernelGGL(k_accumulate, dim3(grid_all), dim3(kBlock), 0, s, S.Lpix, P.accum, pixels, frames);
    ATLAS_HIP_TRY(hipEventRecord(S.ev_acc, s));
    ATLAS_HIP_TRY(hipGetLastError());
    return ATLAS_RT_OK;
}

atlas_rt_status refresh_snapshots(PathTracer& P, hipStream_t s) {
    const atlas_rt_nrc& N = *P.nrc;
    for (int j = 0; j < 2; j++) {
        ATLAS_HIP_TRY(hipMemcpyAsync(P.snap_table[j], N.table16, size_t(N.table_params) * sizeof(uint16_t),
                                     hipMemcpyDeviceToDevice, s));
        ATLAS_HIP_TRY(hipMemcpyAsync(P.snap_wn[j], N.wn, size_t(nrc::kWnSize) * sizeof(uint16_t),
                                     hipMemcpyDeviceToDevice, s));
    }
    return ATLAS_RT_OK;
}

}  // namespace pt
}  // namespace atlas_rt
