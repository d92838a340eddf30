// Device-side shading for the wavefront path tracer, restating the engine's
// hit shader and BRDF code in FP32:
//   RNG      data/shader/common/random.hsh:5-54 (Jenkins one-at-a-time hash of
//            float bits, mantissa -> [0,1); seed += 1 per draw)
//   BRDF     data/shader/brdf/brdf.hsh:9-52 (Schlick Fresnel, renormalized
//            Disney diffuse, Smith-GGX correlated/separable visibility, GGX D)
//   eval     brdf/brdfEval.hsh:8-32, surface update brdf/surface.hsh:25-42
//   sampling brdf/brdfSample.hsh:8-90 (cosine hemisphere, GGX VNDF)
//   bounce   data/shader/pathtracer/rayHit.csh:160-325 (EvaluateBounce,
//            EvaluateDirectLight, EvaluateIndirectLight)
// GLSL leaves the precision of sin, cos and pow to the implementation, so
// the restatement fixes them: sin / cos of the sample angles by the Cephes
// single-precision method (S. L. Moshier, Cephes Math Library: Cody-Waite
// reduction by pi/4 in three parts, degree-7 / degree-8 polynomials on
// [-pi/4, pi/4]) and pow(x, 5) as (x^2)^2 * x.  The CPU restatement uses the
// same operations, so both round identically.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "pt_types.h"

#pragma clang fp contract(off)

namespace atlas_rt {
namespace pt {

constexpr float kPI = 3.1415926535897932384626433832795f;
constexpr float kInvPI = 0.3183098861837906715377675267450f;
constexpr float kEpsilon = 0.1f;   // EPSILON, raytracer/common.hsh:9
constexpr float kInf = 1.0e12f;    // INF, raytracer/common.hsh:8

struct V {
    float x, y, z;
};
__device__ __forceinline__ V mk(float x, float y, float z) { return V{x, y, z}; }
__device__ __forceinline__ V operator+(V a, V b) { return V{a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ V operator-(V a, V b) { return V{a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ V operator*(V a, V b) { return V{a.x * b.x, a.y * b.y, a.z * b.z}; }
__device__ __forceinline__ V operator*(V a, float s) { return V{a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ V operator*(float s, V a) { return V{s * a.x, s * a.y, s * a.z}; }
__device__ __forceinline__ V operator/(V a, float s) { return V{a.x / s, a.y / s, a.z / s}; }
__device__ __forceinline__ V operator-(V a) { return V{-a.x, -a.y, -a.z}; }
__device__ __forceinline__ float dot(V a, V b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
__device__ __forceinline__ V cross(V a, V b) {
    return V{a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y};
}
__device__ __forceinline__ V normalize(V a) { return a / sqrtf(dot(a, a)); }
__device__ __forceinline__ float saturate(float x) { return fminf(fmaxf(x, 0.0f), 1.0f); }
__device__ __forceinline__ float clampf(float x, float lo, float hi) { return fminf(fmaxf(x, lo), hi); }
__device__ __forceinline__ float mixf(float a, float b, float t) { return a + t * (b - a); }
__device__ __forceinline__ V mixv(V a, V b, float t) { return V{mixf(a.x, b.x, t), mixf(a.y, b.y, t), mixf(a.z, b.z, t)}; }

// ---- RNG (random.hsh) -----------------------------------------------------
__device__ __forceinline__ uint32_t jhash(uint32_t x) {
    x += (x << 10u);
    x ^= (x >> 6u);
    x += (x << 3u);
    x ^= (x >> 11u);
    x += (x << 15u);
    return x;
}
__device__ __forceinline__ float float_construct(uint32_t m) {
    m &= 0x007FFFFFu;
    m |= 0x3F800000u;
    return __uint_as_float(m) - 1.0f;
}
// random(vec2(x, seed)); seed += 1.0
__device__ __forceinline__ float rnd2(float x, float& seed) {
    const uint32_t h = jhash(__float_as_uint(x) ^ jhash(__float_as_uint(seed)));
    seed += 1.0f;
    return float_construct(h);
}
// random(vec2(a, b)) without a seed update (rayGen.csh:33-34).
__device__ __forceinline__ float rnd_pair(float a, float b) {
    return float_construct(jhash(__float_as_uint(a) ^ jhash(__float_as_uint(b))));
}

// ---- sin / cos / pow(x, 5) (see the header) --------------------------------
// x >= 0 and moderate (the sample angles 2 pi u, u in [0, 1)).
__device__ __forceinline__ void sin_cos(float x, float& sn, float& cs) {
    int j = int(x * 1.27323954473516f);  // 4 / pi
    float y = float(j);
    if (j & 1) {
        j += 1;
        y += 1.0f;
    }
    j &= 7;
    const bool swap = j == 2 || j == 6;
    const bool sneg = j > 3;
    const bool cneg = j == 2 || j == 4;
    const float z = ((x - y * 0.78515625f) - y * 2.4187564849853515625e-4f) - y * 3.77489497744594108e-8f;
    const float zz = z * z;
    const float sp = ((-1.9515295891e-4f * zz + 8.3321608736e-3f) * zz - 1.6666654611e-1f) * zz * z + z;
    const float cp = ((2.443315711809948e-5f * zz - 1.388731625493765e-3f) * zz + 4.166664568298827e-2f) * zz * zz -
                     0.5f * zz + 1.0f;
    const float s0 = swap ? cp : sp, c0 = swap ? sp : cp;
    sn = sneg ? -s0 : s0;
    cs = cneg ? -c0 : c0;
}
__device__ __forceinline__ float pow5(float x) {
    const float x2 = x * x;
    return (x2 * x2) * x;
}

// ---- BRDF (brdf.hsh) ------------------------------------------------------
__device__ __forceinline__ V fresnel_schlick(V F0, float F90, float c) {
    const float p = pow5(1.0f - c);
    return F0 + (V{F90, F90, F90} - F0) * p;
}
__device__ __forceinline__ float disney_diffuse(float NdotV, float NdotL, float LdotH, float lr) {
    const float energyBias = mixf(0.0f, 0.5f, lr);
    const float energyFactor = mixf(1.0f, 1.0f / 1.51f, lr);
    const float FD90 = energyBias + 2.0f * LdotH * LdotH * lr;
    const float lightScatter = 1.0f + (FD90 - 1.0f) * pow5(1.0f - NdotL);
    const float viewScatter = 1.0f + (FD90 - 1.0f) * pow5(1.0f - NdotV);
    return lightScatter * viewScatter * energyFactor;
}
__device__ __forceinline__ float vis_smith_separable(float c, float alpha) {
    const float a2 = alpha * alpha;
    return 2.0f * c / (c + sqrtf(a2 + (1.0f - a2) * c * c));
}
__device__ __forceinline__ float vis_smith_correlated(float NdotL, float NdotV, float alpha) {
    const float a2 = alpha * alpha;
    const float GGXL = NdotV * sqrtf((-NdotL * a2 + NdotL) * NdotL + a2);
    const float GGXV = NdotL * sqrtf((-NdotV * a2 + NdotV) * NdotV + a2);
    return 0.5f / (GGXL + GGXV + 0.0000001f);
}
__device__ __forceinline__ float ggx_d(float NdotH, float alpha) {
    const float a2 = alpha * alpha;
    const float f = (NdotH * a2 - NdotH) * NdotH + 1.0f;
    return a2 / (f * f + 0.0000001f) * kInvPI;
}

struct Surface {
    V P, V_, N, L, H, Ng;
    float NdotL, LdotH, NdotH, NdotV;
    V F0;
    float F90;
    Material m;
};

// UpdateSurface (brdf/surface.hsh:25-42).
__device__ __forceinline__ void update_surface(Surface& s) {
    s.L = normalize(s.L);
    s.V_ = normalize(s.V_);
    s.N = normalize(s.N);
    s.H = normalize(s.L + s.V_);
    s.NdotL = saturate(dot(s.N, s.L));
    s.LdotH = saturate(dot(s.L, s.H));
    s.NdotH = saturate(dot(s.N, s.H));
    s.NdotV = saturate(dot(s.N, s.V_));
    const float f = 0.16f * (s.m.reflectance * s.m.reflectance);
    const V base = mk(s.m.base[0], s.m.base[1], s.m.base[2]);
    s.F0 = mixv(V{f, f, f}, base, s.m.metalness);
    s.F90 = saturate(50.0f * dot(s.F0, V{0.333f, 0.333f, 0.333f}));
}

// EvaluateDiffuseBRDF / EvaluateSpecularBRDF (brdfEval.hsh:8-32).
__device__ __forceinline__ V eval_diffuse(const Surface& s) {
    const float r = fmaxf(s.m.roughness * s.m.roughness, 0.00001f);
    const float d = disney_diffuse(s.NdotV, s.NdotL, s.LdotH, r);
    const V base = mk(s.m.base[0], s.m.base[1], s.m.base[2]);
    return (1.0f - s.m.metalness) * base * d * kInvPI;
}
__device__ __forceinline__ V eval_specular(const Surface& s) {
    const float r = fmaxf(s.m.roughness * s.m.roughness, 0.00001f);
    const V F = fresnel_schlick(s.F0, s.F90, s.LdotH);
    const float G = vis_smith_correlated(s.NdotV, s.NdotL, r);
    const float D = ggx_d(s.NdotH, r);
    return (F * D) * G;
}

struct BrdfSample {
    V reflectance;
    V L;
    float pdf;
};

__device__ __forceinline__ void tangent_frame(V N, V& t, V& b, bool normalize_b) {
    const V up = fabsf(N.z) < 0.999f ? V{0.0f, 0.0f, 1.0f} : V{1.0f, 0.0f, 0.0f};
    t = normalize(cross(up, N));
    b = cross(N, t);
    if (normalize_b) b = normalize(b);
}

// SampleDiffuseBRDF (brdfSample.hsh:8-32).
__device__ __forceinline__ BrdfSample sample_diffuse(Surface& s, float u0, float u1) {
    const float theta = sqrtf(u0);
    const float phi = 2.0f * kPI * u1;
    float sp, cp;
    sin_cos(phi, sp, cp);
    const V Ll = V{theta * cp, theta * sp, sqrtf(1.0f - u0)};
    V t, b;
    tangent_frame(s.N, t, b, false);
    s.L = normalize(t * Ll.x + b * Ll.y + s.N * Ll.z);
    update_surface(s);
    BrdfSample o;
    o.L = s.L;
    o.pdf = s.NdotL / kPI;
    o.reflectance = eval_diffuse(s);
    return o;
}

// SampleGGXVNDF (brdfSample.hsh:34-55).
__device__ __forceinline__ V ggx_vndf(float x0, float x1, V Vt, float alpha) {
    const V Vh = normalize(V{Vt.x * alpha, Vt.y * alpha, Vt.z});
    const float phi = kPI * 2.0f * x0;
    float y, x;
    sin_cos(phi, y, x);
    const float z = (1.0f - x1) * (1.0f + Vh.z) + -Vh.z;
    const float r = sqrtf(clampf(1.0f - z * z, 0.0f, 1.0f));
    const V c = V{x * r, y * r, z};
    const V H = c + Vh;
    return V{H.x * alpha, H.y * alpha, fmaxf(H.z, 0.0f)};
}

// SampleSpecularBRDF (brdfSample.hsh:57-90).
__device__ __forceinline__ BrdfSample sample_specular(Surface& s, float u0, float u1) {
    const float alpha = s.m.roughness * s.m.roughness;
    s.V_ = normalize(s.V_);
    const V N = normalize(s.N);
    V t, b;
    tangent_frame(N, t, b, true);
    const V Vt = normalize(V{dot(s.V_, t), dot(s.V_, b), dot(s.V_, N)});
    const V h = ggx_vndf(u0, u1, Vt, alpha);
    const V M = normalize(t * h.x + b * h.y + N * h.z);
    s.L = 2.0f * dot(s.V_, M) * M - s.V_;
    update_surface(s);
    BrdfSample o;
    o.pdf = 1.0f;
    o.reflectance = V{0.0f, 0.0f, 0.0f};
    o.L = s.L;
    if (s.NdotL > 0.0f && s.LdotH > 0.0f) {
        const V F = fresnel_schlick(s.F0, s.F90, s.LdotH);
        const float Vis = vis_smith_correlated(s.NdotV, s.NdotL, alpha);
        const float G1 = vis_smith_separable(s.NdotV, alpha);
        o.pdf = G1 / (4.0f * fabsf(dot(s.V_, s.N)));
        o.reflectance = F * Vis;
    }
    return o;
}

}  // namespace pt
}  // namespace atlas_rt
