// Plain-data structs shared by the path tracer's host code and kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace atlas_rt {
namespace pt {

// atlas_rt_material (96 B): the RaytraceMaterial fields the bounce shader
// reads (raytracer/structures.hsh:107-135, surface.hsh:63-140).
struct Material {
    float base[3];
    float opacity;
    float emissive[3];
    float roughness;
    float metalness;
    float ao;
    float reflectance;
    int32_t two_sided;
    int32_t opacity_texture;  // index into the texture table; outside it = none
    int32_t invert_uvs;
    int32_t id;  // RaytraceMaterial.ID: the realtime G-buffer's material index
    float normal_scale;
    int32_t base_color_texture;
    int32_t roughness_texture;
    int32_t metalness_texture;
    int32_t ao_texture;
    int32_t normal_texture;
    int32_t use_vertex_colors;  // base colour times the interpolated vertex colour
    int32_t reserved[2];
};

// atlas_rt_gpu_light / GPULight (64 B), read as UnpackLight does
// (raytracer/common.hsh:100-121): data[0] = bits(type << 28 | slot),
// data[1] = normalised selection weight, data[2] = area, data[3] = bits(instance).
struct Light {
    float P[4];
    float N[4];
    float color[4];
    float data[4];
};

// One texture of the path tracer's table: UNORM8 texels at `offset`,
// `channels` per texel.
struct TexDesc {
    uint32_t offset;
    uint32_t width;
    uint32_t height;
    uint32_t channels;
};

// What a kernel needs to evaluate a triangle's opacity.  materials = null:
// opacity-map slots read as 1 (the world-level traces hold no material table).
struct AlphaArgs {
    const float4* const* shade;  // per BLAS: 3 float4 per slot (n0 + material, n1 + opacity, n2)
    const float4* const* uv;     // per BLAS: 2 float4 per slot, or null
    const Material* materials;
    const TexDesc* tex;
    const uint8_t* texels;
    uint32_t tex_count;
    uint32_t pad;
};

struct Camera {
    float loc[3];
    float origin[3];
    float right[3];
    float bottom[3];
};

// Per-frame device control block, advanced on the device so a whole frame
// can be captured in a graph and replayed.
struct FrameCtl {
    uint64_t pcg_state;      // Random's PCG32 state (RandomHelper.cpp:8-26)
    int32_t sample_count;    // samples accumulated so far
    uint32_t pad;
    float seeds[16];         // per-bounce seeds of the current frame
    uint32_t paths[17];      // live paths per bounce (paths[0] = W*H)
    uint32_t shadows[16];    // shadow rays per bounce
    // radiance-cache mode
    uint32_t frame;          // frames rendered (selects the training pixels)
    uint32_t train_paths;    // training paths started this frame
    uint32_t queries;        // cache queries issued this frame
    uint32_t train_samples;  // training samples built this frame
    float train_loss;        // loss of the frame's last training step
    uint32_t pull;           // the resident frame kernel's counter of path range r (ctl[r], r < kPullRanges)
    uint32_t pad2[2];
};

// Realtime mode's G-buffer, written by the first bounce of each pixel's
// first sample (rayHit.csh:76-117); enabled = 0 outside realtime mode.
// Matrices are glm column-major mat4 (m[4 * column + row]).
struct RealtimeArgs {
    float* depth;             // view-space z, the shaders' INF (1e12) on a miss
    uint32_t* velocity;       // 2 x binary16: (ndc last - ndc now) / 2
    uint32_t* normal;         // 2 x binary16: EncodeNormal(geometry normal)
    uint16_t* material;       // material ID, 65535 on a miss
    const float4* last_xform; // last frame's object-to-world rows, 3 float4 per instance slot
    float view[16];
    float proj[16];
    float last_vp[16];
    uint32_t spf;             // samples per frame: ray ID = pixel * spf + sample
    uint32_t enabled;
};

// Radiance-cache buffers the bounce kernels write (all null in plain mode).
// Training record r = (training path, bounce): rec_q 3 float4 (the cache
// query of the vertex), rec_beta (throughput reaching the vertex, w = 1 if
// the vertex is a surface hit), rec_add (radiance the bounce added to the
// pixel, already multiplied by that throughput).
struct CacheArgs {
    float4* rec_q;
    float4* rec_beta;
    float4* rec_add;
    float4* q_buf;           // 3 float4 per render-path query
    float4* q_meta;          // (throughput xyz, pixel bits)
    uint32_t* train_list;    // resident frames: the frame's training pixels, by training index
    uint32_t train_cap;      // maximum training paths per frame
    uint32_t train_stride;   // about one pixel in train_stride trains
    uint32_t enabled;
    uint32_t pad;
};

}  // namespace pt
}  // namespace atlas_rt
