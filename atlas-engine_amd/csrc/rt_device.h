// Device-side intersection math for gfx950, bit-exact with the engine's shaders.
//
// Slab test: IntersectAABB (data/shader/raytracer/intersections.hsh:18-34)
//   t0 = (bmin - o) / d, t1 = (bmax - o) / d, tsmall = min, tbig = max,
//   tminf = max(max(tmin, ts.x), max(ts.y, ts.z)),
//   tmaxf = min(min(tmax, tb.x), min(tb.y, tb.z)), hit iff tminf <= tmaxf,
//   dist = hit ? tminf : tmax.
// Triangle: IntersectTriangle (intersections.hsh:36-58)
//   e0 = v1-v0, e1 = v2-v0, s = o-v0, p = s x e0, q = d x e1,
//   (t,u,v) = (p.e1, q.s, p.d) / (q.e0); accept t,u,v >= 0 and u+v <= 1.
//
// Exactness:
//  * Every float operation is written out in a fixed order, and FMA
//    contraction is off (#pragma below + -ffp-contract=off), so products and
//    sums round exactly as in the CPU oracle.  dot = (x*x' + y*y') + z*z'.
//  * Division: a correctly rounded float quotient a/b equals
//    (float)((double)a * (1.0 / (double)b)).  The double reciprocal and
//    product each carry at most 2^-53 relative error, so the double quotient is
//    within 2^-52 of a/b, while a/b of two floats lies at least 2^-50 (relative)
//    from any float rounding midpoint; the final conversion therefore rounds
//    to the same float as IEEE division, including infinities, NaN and
//    subnormal results.  It costs cvt + v_mul_f64 + cvt (3 VALU ops) per
//    quotient, against 11 for the IEEE f32 division sequence, and the
//    reciprocals of the ray direction are computed once per ray.
//  * min/max: GLSL leaves NaN behaviour undefined; this code uses IEEE
//    minNum/maxNum (fminf/fmaxf -> v_min_f32/v_max_f32), the same rule the
//    CPU oracle uses.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#pragma clang fp contract(off)

namespace atlas_rt {

// Loads of tree, triangle and instance records.  Their pointers are often
// read from tables in memory (a world's BLAS table, a batch's BLAS list), so
// the compiler cannot tell their address space and emits flat loads, whose
// waits also cover every LDS access in flight (the traversal stack).  These
// records always live in global memory: load them as such.
__device__ __forceinline__ float4 ldg4(const float4* p) {
    using V4 = float __attribute__((ext_vector_type(4)));
    const V4 v = *(const __attribute__((address_space(1))) V4*)p;
    return make_float4(v.x, v.y, v.z, v.w);
}

struct DevRay {
    float ox, oy, oz;
    float dx, dy, dz;
    double rx, ry, rz;  // 1 / d, correctly rounded double
    float fx, fy, fz;   // 1 / d rounded to float (the widened float slab test)
    float tmin;
    bool fast;          // every |d| is 0 or in [2^-126, 2^126]: fx, fy, fz are accurate
};

__device__ __forceinline__ float qdiv(float a, double r) { return (float)((double)a * r); }

__device__ __forceinline__ bool fast_component(float d) {
    const float a = fabsf(d);
    return d == 0.0f || (a >= 1.1754943508222875e-38f && a <= 8.507059173023462e+37f);
}

__device__ __forceinline__ void make_dev_ray(DevRay& R, float ox, float oy, float oz, float dx, float dy, float dz,
                                             float tmin) {
    R.ox = ox;
    R.oy = oy;
    R.oz = oz;
    R.dx = dx;
    R.dy = dy;
    R.dz = dz;
    R.rx = 1.0 / (double)dx;
    R.ry = 1.0 / (double)dy;
    R.rz = 1.0 / (double)dz;
    R.fx = (float)R.rx;
    R.fy = (float)R.ry;
    R.fz = (float)R.rz;
    R.tmin = tmin;
    R.fast = fast_component(dx) && fast_component(dy) && fast_component(dz);
}

// Slab test against one box; returns hit and writes the engine's `dist`.
__device__ __forceinline__ bool slab_dist(const DevRay& R, float lx, float ly, float lz, float hx, float hy, float hz,
                                          float tmax, float& dist) {
    const float t0x = qdiv(lx - R.ox, R.rx);
    const float t0y = qdiv(ly - R.oy, R.ry);
    const float t0z = qdiv(lz - R.oz, R.rz);
    const float t1x = qdiv(hx - R.ox, R.rx);
    const float t1y = qdiv(hy - R.oy, R.ry);
    const float t1z = qdiv(hz - R.oz, R.rz);
    const float sx = fminf(t0x, t1x), sy = fminf(t0y, t1y), sz = fminf(t0z, t1z);
    const float bx = fmaxf(t0x, t1x), by = fmaxf(t0y, t1y), bz = fmaxf(t0z, t1z);
    const float tminf = fmaxf(fmaxf(R.tmin, sx), fmaxf(sy, sz));
    const float tmaxf = fminf(fminf(tmax, bx), fminf(by, bz));
    const bool hit = tminf <= tmaxf;
    dist = hit ? tminf : tmax;
    return hit;
}

// Widened float slab test of one inner node (the SLAB_FLOAT kernel variants).
// A slab distance computed as (b - o) * RN(1/d) is within 2^-22 of its
// magnitude of the correctly rounded (b - o) / d when 1/d is a normal float;
// min / max keep that bound for the entry and exit distances.  A box counts
// as hit when entry - exit <= 2^-20 (|entry| + |exit|), so every box the
// exact test enters is entered here too (a few more may be); infinities are
// the exact ones and never widen.  Child order uses the float distances.
// Triangles stay exact, so the nearest hit is the exact traversal's; only
// which of several triangles at exactly the same t wins may differ, and a
// triangle whose box the exact test culls at a face may be found.
__device__ __forceinline__ void node_test_float(const DevRay& R, const float4& a, const float4& b, const float4& c,
                                                float tmax, bool& hl, bool& hr, bool& left_first) {
    if (!R.fast) {
        float dl, dr;
        hl = slab_dist(R, a.x, a.y, a.z, a.w, b.x, b.y, tmax, dl);
        hr = slab_dist(R, b.z, b.w, c.x, c.y, c.z, c.w, tmax, dr);
        left_first = dl <= dr;
        return;
    }
    const float l0x = (a.x - R.ox) * R.fx, l0y = (a.y - R.oy) * R.fy, l0z = (a.z - R.oz) * R.fz;
    const float l1x = (a.w - R.ox) * R.fx, l1y = (b.x - R.oy) * R.fy, l1z = (b.y - R.oz) * R.fz;
    const float r0x = (b.z - R.ox) * R.fx, r0y = (b.w - R.oy) * R.fy, r0z = (c.x - R.oz) * R.fz;
    const float r1x = (c.y - R.ox) * R.fx, r1y = (c.z - R.oy) * R.fy, r1z = (c.w - R.oz) * R.fz;
    const float lo_l = fmaxf(fmaxf(R.tmin, fminf(l0x, l1x)), fmaxf(fminf(l0y, l1y), fminf(l0z, l1z)));
    const float hi_l = fminf(fminf(tmax, fmaxf(l0x, l1x)), fminf(fmaxf(l0y, l1y), fmaxf(l0z, l1z)));
    const float lo_r = fmaxf(fmaxf(R.tmin, fminf(r0x, r1x)), fmaxf(fminf(r0y, r1y), fminf(r0z, r1z)));
    const float hi_r = fminf(fminf(tmax, fmaxf(r0x, r1x)), fminf(fmaxf(r0y, r1y), fmaxf(r0z, r1z)));
    constexpr float kWiden = 9.5367431640625e-07f;   // 2^-20
    constexpr float kFloor = 7.52316384526264e-37f;   // 2^-120, for subnormal distances
    constexpr float kBig = 3.4028234663852886e+38f;   // FLT_MAX: keeps infinities out of the margin
    hl = lo_l - hi_l <= kWiden * (fminf(fabsf(lo_l), kBig) + fminf(fabsf(hi_l), kBig)) + kFloor;
    hr = lo_r - hi_r <= kWiden * (fminf(fabsf(lo_r), kBig) + fminf(fabsf(hi_r), kBig)) + kFloor;
    left_first = (hl ? lo_l : tmax) <= (hr ? lo_r : tmax);
}

// The same widened float test for one box (the wide-node SLAB_FLOAT
// variants): hit and entry distance as node_test_float decides them for a
// child; the exact test where the direction has a component outside the
// float reciprocal's accurate range.
__device__ __forceinline__ bool slab_dist_float(const DevRay& R, float lx, float ly, float lz, float hx, float hy,
                                                float hz, float tmax, float& dist) {
    if (!R.fast) return slab_dist(R, lx, ly, lz, hx, hy, hz, tmax, dist);
    const float t0x = (lx - R.ox) * R.fx, t0y = (ly - R.oy) * R.fy, t0z = (lz - R.oz) * R.fz;
    const float t1x = (hx - R.ox) * R.fx, t1y = (hy - R.oy) * R.fy, t1z = (hz - R.oz) * R.fz;
    const float lo = fmaxf(fmaxf(R.tmin, fminf(t0x, t1x)), fmaxf(fminf(t0y, t1y), fminf(t0z, t1z)));
    const float hi = fminf(fminf(tmax, fmaxf(t0x, t1x)), fminf(fmaxf(t0y, t1y), fmaxf(t0z, t1z)));
    constexpr float kWiden = 9.5367431640625e-07f;   // 2^-20
    constexpr float kFloor = 7.52316384526264e-37f;   // 2^-120
    constexpr float kBig = 3.4028234663852886e+38f;   // FLT_MAX
    const bool hit = lo - hi <= kWiden * (fminf(fabsf(lo), kBig) + fminf(fabsf(hi), kBig)) + kFloor;
    dist = hit ? lo : tmax;
    return hit;
}

// Triangle test on the precomputed layout (v0, e0 = v1 - v0, e1 = v2 - v0; the
// subtractions are the shader's own first two operations, done once on the
// host with the same rounding).  Writes (t,u,v); returns the engine's
// geometric accept (t,u,v >= 0, u+v <= 1).
__device__ __forceinline__ bool tri_uvt(const DevRay& R, float v0x, float v0y, float v0z, float e0x, float e0y,
                                        float e0z, float e1x, float e1y, float e1z, float& t, float& u, float& v) {
    const float sx = R.ox - v0x, sy = R.oy - v0y, sz = R.oz - v0z;
    // p = cross(s, e0), q = cross(d, e1) in glm's component order.
    const float px = sy * e0z - e0y * sz;
    const float py = sz * e0x - e0z * sx;
    const float pz = sx * e0y - e0x * sy;
    const float qx = R.dy * e1z - e1y * R.dz;
    const float qy = R.dz * e1x - e1z * R.dx;
    const float qz = R.dx * e1y - e1x * R.dy;
    const float den = (qx * e0x + qy * e0y) + qz * e0z;
    const float a = (px * e1x + py * e1y) + pz * e1z;
    const float b = (qx * sx + qy * sy) + qz * sz;
    const float c = (px * R.dx + py * R.dy) + pz * R.dz;
    const double rd = 1.0 / (double)den;
    t = qdiv(a, rd);
    u = qdiv(b, rd);
    v = qdiv(c, rd);
    return t >= 0.0f && u >= 0.0f && v >= 0.0f && u + v <= 1.0f;
}

// Reported (t, u, v) of the final hit: the same formula evaluated in double
// from the float32 vertices and ray, then rounded once to float.  Traversal
// decisions never use it (they use tri_uvt), so hit ids stay the engine's;
// the reported values are within ~1 ulp of an exact evaluation, whereas the
// float32 formula loses up to ~1e-4 in u, v when the origin is far from a
// small triangle.  Operation order is fixed (no contraction) so the CPU
// oracle reproduces it bit for bit.
__device__ __forceinline__ void refine_tuv(float ox, float oy, float oz, float dx, float dy, float dz,
                                           float4 v0, float4 v1, float4 v2, float& t, float& u, float& v) {
    const double ax = v0.x, ay = v0.y, az = v0.z;
    const double e0x = (double)v1.x - ax, e0y = (double)v1.y - ay, e0z = (double)v1.z - az;
    const double e1x = (double)v2.x - ax, e1y = (double)v2.y - ay, e1z = (double)v2.z - az;
    const double sx = (double)ox - ax, sy = (double)oy - ay, sz = (double)oz - az;
    const double Dx = dx, Dy = dy, Dz = dz;
    const double px = sy * e0z - e0y * sz, py = sz * e0x - e0z * sx, pz = sx * e0y - e0x * sy;
    const double qx = Dy * e1z - e1y * Dz, qy = Dz * e1x - e1z * Dx, qz = Dx * e1y - e1x * Dy;
    const double den = (qx * e0x + qy * e0y) + qz * e0z;
    t = (float)(((px * e1x + py * e1y) + pz * e1z) / den);
    u = (float)(((qx * sx + qy * sy) + qz * sz) / den);
    v = (float)(((px * Dx + py * Dy) + pz * Dz) / den);
}

__device__ __forceinline__ bool any_nan3(float x, float y, float z) {
    return __builtin_isnan(x) || __builtin_isnan(y) || __builtin_isnan(z);
}

}  // namespace atlas_rt
