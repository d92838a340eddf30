// TEST INFRASTRUCTURE - the oracle.  Only tests/, __graft_entry__.smoke() and
// bench.py's cpu_baseline leg may load it, and only as the checker.
//
// Scalar vector math with the exact semantics the reference relies on.
//  - glm 0.9.8.0 (vcpkg.json:46-47, not installed here) scalar min/max/clamp
//    written out: min(a,b) = (b < a) ? b : a, max(a,b) = (a < b) ? b : a,
//    clamp(x,lo,hi) = min(max(x,lo),hi), mix(x,y,a) = x + a*(y - x).
//    Parity with glm itself is unpinned (glm is absent): these choices only
//    differ from other readings for NaN inputs and the sign of zero.
//  - GLSL min/max in the shaders (intersections.hsh:8-12) are undefined for NaN;
//    the restatement uses IEEE-754 minNum/maxNum (fminf/fmaxf), which is what
//    AMD's v_min_f32/v_max_f32 implement.
//  - dot(a,b) = (a.x*b.x + a.y*b.y) + a.z*b.z, cross as glm writes it; no FMA
//    contraction anywhere (the oracle is compiled with -ffp-contract=off).
#pragma once
#include <cmath>
#include <cstdint>
#include <cfloat>

namespace oracle {

struct V3 {
    float x, y, z;
    float& operator[](int i) { return i == 0 ? x : (i == 1 ? y : z); }
    float operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
};

inline V3 v3(float x, float y, float z) { return V3{x, y, z}; }
inline V3 operator+(V3 a, V3 b) { return V3{a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 operator-(V3 a, V3 b) { return V3{a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 operator*(V3 a, V3 b) { return V3{a.x * b.x, a.y * b.y, a.z * b.z}; }
inline V3 operator/(V3 a, V3 b) { return V3{a.x / b.x, a.y / b.y, a.z / b.z}; }
inline V3 operator*(float s, V3 a) { return V3{s * a.x, s * a.y, s * a.z}; }
inline V3 operator/(V3 a, float s) { return V3{a.x / s, a.y / s, a.z / s}; }

// glm 0.9.8 scalar forms.
inline float gmin(float a, float b) { return (b < a) ? b : a; }
inline float gmax(float a, float b) { return (a < b) ? b : a; }
inline float gclamp(float x, float lo, float hi) { return gmin(gmax(x, lo), hi); }
inline V3 gmin(V3 a, V3 b) { return V3{gmin(a.x, b.x), gmin(a.y, b.y), gmin(a.z, b.z)}; }
inline V3 gmax(V3 a, V3 b) { return V3{gmax(a.x, b.x), gmax(a.y, b.y), gmax(a.z, b.z)}; }
inline V3 gmix(V3 x, V3 y, float a) { return x + a * (y - x); }

inline float dot(V3 a, V3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
inline V3 cross(V3 a, V3 b) {
    return V3{a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y};
}

// IEEE minNum / maxNum for the GLSL side.
inline float smin(float a, float b) { return std::fmin(a, b); }
inline float smax(float a, float b) { return std::fmax(a, b); }

struct Box {
    V3 min, max;
};

// BVHBuilder::InitialAABB (BVH.cpp:863-869).
inline Box initial_box() { return Box{V3{FLT_MAX, FLT_MAX, FLT_MAX}, V3{-FLT_MAX, -FLT_MAX, -FLT_MAX}}; }
// AABB::Grow(AABB) (AABB.cpp:88-93).
inline void grow(Box& b, const Box& o) {
    b.max = gmax(b.max, o.max);
    b.min = gmin(b.min, o.min);
}
// AABB::Grow(vec3) (AABB.cpp:95-100): note argument order (vector first).
inline void grow(Box& b, V3 p) {
    b.max = gmax(p, b.max);
    b.min = gmin(p, b.min);
}
// AABB::Intersect (AABB.cpp:102-107).
inline void intersect(Box& b, const Box& o) {
    b.min = gmax(b.min, o.min);
    b.max = gmin(b.max, o.max);
}
// AABB::GetSurfaceArea (AABB.cpp:109-115).
inline float surface_area(const Box& b) {
    V3 d = b.max - b.min;
    return 2.0f * (d.x * d.y + d.y * d.z + d.z * d.x);
}

}  // namespace oracle
