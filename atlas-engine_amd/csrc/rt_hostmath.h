// Host math for the builder with the reference's scalar semantics.
//   glm 0.9.8 (the engine's pinned math library, vcpkg.json:46-47):
//     min(a,b) = (b < a) ? b : a      max(a,b) = (a < b) ? b : a
//     clamp(x,lo,hi) = min(max(x,lo),hi)      mix(x,y,t) = x + t*(y - x)
//   AABB::Grow / Intersect / GetSurfaceArea (src/engine/volume/AABB.cpp:88-115).
// Compiled without FMA contraction so each operation rounds as written.
#pragma once
#include <cfloat>
#include <cstdint>

namespace atlas_rt {

struct HBox {
    float lo[3];
    float hi[3];
};

inline float gl_min(float a, float b) { return (b < a) ? b : a; }
inline float gl_max(float a, float b) { return (a < b) ? b : a; }
inline float gl_clamp(float x, float lo, float hi) { return gl_min(gl_max(x, lo), hi); }

inline HBox empty_box() {
    HBox b;
    for (int k = 0; k < 3; k++) {
        b.lo[k] = FLT_MAX;
        b.hi[k] = -FLT_MAX;
    }
    return b;
}

// AABB::Grow(AABB): max = max(max, o.max); min = min(min, o.min).
inline void box_grow(HBox& b, const HBox& o) {
    for (int k = 0; k < 3; k++) {
        b.hi[k] = gl_max(b.hi[k], o.hi[k]);
        b.lo[k] = gl_min(b.lo[k], o.lo[k]);
    }
}

// AABB::Grow(vec3): max = max(p, max); min = min(p, min).
inline void box_grow_point(HBox& b, const float* p) {
    for (int k = 0; k < 3; k++) {
        b.hi[k] = gl_max(p[k], b.hi[k]);
        b.lo[k] = gl_min(p[k], b.lo[k]);
    }
}

// AABB::Intersect.
inline void box_clip(HBox& b, const HBox& o) {
    for (int k = 0; k < 3; k++) {
        b.lo[k] = gl_max(b.lo[k], o.lo[k]);
        b.hi[k] = gl_min(b.hi[k], o.hi[k]);
    }
}

// AABB::GetSurfaceArea: 2 * ((x*y + y*z) + z*x).
inline float box_area(const HBox& b) {
    float dx = b.hi[0] - b.lo[0];
    float dy = b.hi[1] - b.lo[1];
    float dz = b.hi[2] - b.lo[2];
    return 2.0f * (dx * dy + dy * dz + dz * dx);
}

}  // namespace atlas_rt
