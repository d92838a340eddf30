// Plain-data structs shared by the path tracer's host code and kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace atlas_rt {
namespace pt {

// atlas_rt_material (48 B): the RaytraceMaterial fields the bounce shader
// reads (raytracer/surface.hsh:22-44).
struct Material {
    float base[3];
    float opacity;
    float emissive[3];
    float roughness;
    float metalness;
    float ao;
    float reflectance;
    int32_t two_sided;
};

// Unpacked GPULight (raytracer/common.hsh:UnpackLight): N = direction the
// light travels, pdf = normalised selection weight (UpdateLights).
struct Light {
    float N[3];
    uint32_t type;
    float radiance[3];
    float pdf;
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
    uint32_t pad2[3];
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
    uint32_t train_cap;      // maximum training paths per frame
    uint32_t train_stride;   // about one pixel in train_stride trains
    uint32_t enabled;
    uint32_t pad;
};

}  // namespace pt
}  // namespace atlas_rt
