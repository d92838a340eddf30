// Plain-data structs shared by the path tracer's host code and kernels.
#pragma once

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
};

}  // namespace pt
}  // namespace atlas_rt
