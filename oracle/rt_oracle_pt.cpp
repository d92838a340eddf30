// TEST INFRASTRUCTURE - the oracle.  Only tests/, __graft_entry__.smoke() and
// bench.py's cpu_baseline leg may load it, and only as the checker.
//
// CPU restatement of the engine's progressive path tracer, one path per pixel
// per frame:
//   PathTracingRenderer::Render (src/engine/renderer/PathTracingRenderer.cpp:29-292):
//     per-bounce seeds from Random::SampleFastUniformFloat (PCG32,
//     src/engine/common/RandomHelper.cpp:8-51), bounces + 1 passes;
//   rayGen.csh:25-90 (jittered pinhole ray through the near-plane frame);
//   rayHit.csh:55-339 (EvaluateBounce / EvaluateDirectLight / EvaluateIndirectLight /
//     CheckVisibility), raytracer/surface.hsh:46-140, raytracer/direct.hsh:12-86;
//   brdf/brdf.hsh, brdfEval.hsh, brdfSample.hsh, brdf/surface.hsh;
//   common/random.hsh (hash RNG).
// Traversal: HitClosestTransparency and HitAnyTransparency
// (raytracer/bvh.hsh:275-357, 443-524) from rt_oracle_trace.cpp; a world whose
// triangles are all opaque takes HitClosest / HitAny, which give the same
// results there.  The shadow ray is traced inline, as the engine does.  Light
// selection and light sampling draw their seed from the pixel index (the
// engine uses the invocation index of the compacted ray buffer).  Payload
// radiance/throughput stay float32 (the engine packs them to half floats
// between bounces).
// Triangle lights: RayTracingWorld::BuildTriangleLightsForMesh /
// UpdateTriangleLights (src/engine/raytracing/RayTracingWorld.cpp:322-407) and
// RayTracingHelper::UpdateLights (src/engine/renderer/helper/RayTracingHelper.cpp:436-500).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <thread>
#include <mutex>
#include <vector>

#include "rt_oracle.h"

// Optional surface and light inputs of the oracle path tracer (NULL: opaque
// triangles, no textures, directional lights only).
struct OracleSurface {
    const float* const* opacity;  // per BLAS: baked opacity per slot, or NULL
    const float* const* uvs;      // per BLAS: 6 floats per slot, or NULL
    const atlas_rt_texture* textures;
    uint32_t texture_count;
    const float* emitters;        // per material: sRGB emissive rgb + intensity; NULL = no triangle lights
    uint32_t emitter_count;
    atlas_rt_gpu_light* lights_out;  // optional copy of the light list
    uint32_t lights_cap;
    uint32_t* lights_count;
    const float* const* tangents;  // per BLAS: 6 floats per slot (tangent, bitangent), or NULL
    const float* const* colors;    // per BLAS: 12 floats per slot (vertex colours RGBA), or NULL
};

extern "C" float oracle_h2f(uint16_t h);

namespace oracle {
namespace {

const float kPI = 3.1415926535897932384626433832795f;
const float kInvPI = 0.3183098861837906715377675267450f;
const float kEps = 0.1f;
const float kInf = 1.0e12f;

inline V3 mul(V3 a, float s) { return V3{a.x * s, a.y * s, a.z * s}; }
inline V3 neg(V3 a) { return V3{-a.x, -a.y, -a.z}; }
inline V3 nrm(V3 a) {
    float l = std::sqrt(dot(a, a));
    return V3{a.x / l, a.y / l, a.z / l};
}
inline float sat(float x) { return std::fmin(std::fmax(x, 0.0f), 1.0f); }
inline float clampf(float x, float lo, float hi) { return std::fmin(std::fmax(x, lo), hi); }
inline float mixf(float a, float b, float t) { return a + t * (b - a); }

// random.hsh:5-54
inline uint32_t ohash(uint32_t x) {
    x += (x << 10u);
    x ^= (x >> 6u);
    x += (x << 3u);
    x ^= (x >> 11u);
    x += (x << 15u);
    return x;
}
inline uint32_t fbits(float f) {
    uint32_t u;
    std::memcpy(&u, &f, 4);
    return u;
}
inline float to_unit(uint32_t m) {
    m = (m & 0x007FFFFFu) | 0x3F800000u;
    float f;
    std::memcpy(&f, &m, 4);
    return f - 1.0f;
}
inline float rand2(float a, float b) { return to_unit(ohash(fbits(a) ^ ohash(fbits(b)))); }
inline float rand_seq(float x, float& seed) {
    float r = rand2(x, seed);
    seed += 1.0f;
    return r;
}

// RandomHelper.cpp:8-51
struct Pcg {
    uint64_t state;
    uint32_t next() {
        uint64_t x = state;
        unsigned count = unsigned(x >> 59u);
        state = x * 6364136223846793005ull + 1442695040888963407ull;
        x ^= x >> 18;
        uint32_t v = uint32_t(x >> 27u);
        return (v >> count) | (v << ((0u - count) & 31u));
    }
    float fast_uniform() { return float(next()) / float(0xFFFFFFFFu) * 0.999999f; }
};

struct Mat {
    V3 base, emissive;
    float opacity, roughness, metalness, ao, reflectance;
    int32_t two_sided, opacity_texture, invert_uvs;
    int32_t id;  // RaytraceMaterial.ID
    float normal_scale;
    int32_t base_color_texture, roughness_texture, metalness_texture, ao_texture, normal_texture;
    int32_t use_vertex_colors;  // GPUMaterial.useVertexColors
};

struct Tex {
    uint32_t w, h;
    const uint8_t* texels;  // UNORM8, row-major, `ch` channels per texel
    uint32_t ch = 1;
};

// textureLod with the engine's global sampler (linear filter, repeat
// addressing, MainRenderer.cpp:922-929) at level 0, as the Sample*Bilinear
// functions of raytracer/texture.hsh read it: missing channels read g = b = 0,
// a = 1.  The caller checks the index.
void sample_tex(const Tex& T, float x, float y, float out[4]) {
    const float xw = x - std::floor(x), yw = y - std::floor(y);
    const float fx = xw * float(T.w) - 0.5f, fy = yw * float(T.h) - 0.5f;
    const float x0 = std::floor(fx), y0 = std::floor(fy);
    const float a = fx - x0, b = fy - y0;
    const int32_t ix = int32_t(x0), iy = int32_t(y0);
    const uint32_t i0 = ix < 0 ? T.w - 1 : std::min(uint32_t(ix), T.w - 1);
    const uint32_t j0 = iy < 0 ? T.h - 1 : std::min(uint32_t(iy), T.h - 1);
    const uint32_t i1 = i0 + 1 >= T.w ? 0u : i0 + 1, j1 = j0 + 1 >= T.h ? 0u : j0 + 1;
    out[0] = out[1] = out[2] = 0.0f;
    out[3] = 1.0f;
    for (uint32_t c = 0; c < T.ch; c++) {
        auto at = [&](uint32_t i, uint32_t j) { return float(T.texels[(size_t(j) * T.w + i) * T.ch + c]) / 255.0f; };
        const float r0 = at(i0, j0) * (1.0f - a) + at(i1, j0) * a;
        const float r1 = at(i0, j1) * (1.0f - a) + at(i1, j1) * a;
        out[c] = r0 * (1.0f - b) + r1 * b;
    }
}

bool has_tex(const std::vector<Tex>& tex, int32_t id) { return id >= 0 && size_t(id) < tex.size(); }

// SampleOpacityBilinear (raytracer/texture.hsh:18-25): .r; an index outside
// the table reads as no texture (1.0).
float sample_opacity(const std::vector<Tex>& tex, int32_t id, float x, float y) {
    if (!has_tex(tex, id)) return 1.0f;
    float v[4];
    sample_tex(tex[size_t(id)], x, y, v);
    return v[0];
}

struct Surf {
    V3 P, Vv, N, L, H, Ng;
    float NdotL, LdotH, NdotH, NdotV;
    V3 F0;
    float F90;
    Mat m;
};

// GLSL's sin, cos and pow(x, 5) have implementation-defined precision; the
// restatement fixes them (and the kernels use the same operations): the
// sample angles' sin / cos by Cephes' single-precision method (Moshier,
// Cephes Math Library sinf.c / cosf.c: x * 4/pi to an even octant, pi/4
// subtracted in three parts, sine and cosine polynomials on [-pi/4, pi/4]),
// pow(x, 5) as (x^2)^2 * x.
void sincos_ph(float x, float& sn, float& cs) {
    int oct = int(x * 1.27323954473516f);
    float k = float(oct);
    if (oct & 1) {
        oct += 1;
        k += 1.0f;
    }
    oct &= 7;
    const float r = ((x - k * 0.78515625f) - k * 2.4187564849853515625e-4f) - k * 3.77489497744594108e-8f;
    const float r2 = r * r;
    const float ps = ((-1.9515295891e-4f * r2 + 8.3321608736e-3f) * r2 - 1.6666654611e-1f) * r2 * r + r;
    const float pc = ((2.443315711809948e-5f * r2 - 1.388731625493765e-3f) * r2 + 4.166664568298827e-2f) * r2 * r2 -
                     0.5f * r2 + 1.0f;
    switch (oct) {  // octants 0, 2, 4, 6: (sin, cos) = (s, c), (c, -s), (-s, -c), (-c, s)
        case 0: sn = ps; cs = pc; break;
        case 2: sn = pc; cs = -ps; break;
        case 4: sn = -ps; cs = -pc; break;
        default: sn = -pc; cs = ps; break;
    }
}
inline float pow5f(float x) {
    float sq = x * x;
    return (sq * sq) * x;
}

inline V3 schlick(V3 F0, float F90, float c) {
    float p = pow5f(1.0f - c);
    return F0 + mul(V3{F90, F90, F90} - F0, p);
}
void update(Surf& s) {
    s.L = nrm(s.L);
    s.Vv = nrm(s.Vv);
    s.N = nrm(s.N);
    s.H = nrm(s.L + s.Vv);
    s.NdotL = sat(dot(s.N, s.L));
    s.LdotH = sat(dot(s.L, s.H));
    s.NdotH = sat(dot(s.N, s.H));
    s.NdotV = sat(dot(s.N, s.Vv));
    float f = 0.16f * (s.m.reflectance * s.m.reflectance);
    s.F0 = V3{mixf(f, s.m.base.x, s.m.metalness), mixf(f, s.m.base.y, s.m.metalness), mixf(f, s.m.base.z, s.m.metalness)};
    s.F90 = sat(50.0f * dot(s.F0, V3{0.333f, 0.333f, 0.333f}));
}
V3 diffuse(const Surf& s) {
    float lr = std::fmax(s.m.roughness * s.m.roughness, 0.00001f);
    float bias = mixf(0.0f, 0.5f, lr), factor = mixf(1.0f, 1.0f / 1.51f, lr);
    float fd90 = bias + 2.0f * s.LdotH * s.LdotH * lr;
    float ls = 1.0f + (fd90 - 1.0f) * pow5f(1.0f - s.NdotL);
    float vs = 1.0f + (fd90 - 1.0f) * pow5f(1.0f - s.NdotV);
    float d = ls * vs * factor;
    return mul(mul(mul(s.m.base, 1.0f - s.m.metalness), d), kInvPI);
}
float vis_corr(float NdotL, float NdotV, float a) {
    float a2 = a * a;
    float gl = NdotV * std::sqrt((-NdotL * a2 + NdotL) * NdotL + a2);
    float gv = NdotL * std::sqrt((-NdotV * a2 + NdotV) * NdotV + a2);
    return 0.5f / (gl + gv + 0.0000001f);
}
float vis_sep(float c, float a) {
    float a2 = a * a;
    return 2.0f * c / (c + std::sqrt(a2 + (1.0f - a2) * c * c));
}
float ggx(float NdotH, float a) {
    float a2 = a * a;
    float f = (NdotH * a2 - NdotH) * NdotH + 1.0f;
    return a2 / (f * f + 0.0000001f) * kInvPI;
}
V3 specular(const Surf& s) {
    float r = std::fmax(s.m.roughness * s.m.roughness, 0.00001f);
    V3 F = schlick(s.F0, s.F90, s.LdotH);
    float G = vis_corr(s.NdotV, s.NdotL, r);
    float D = ggx(s.NdotH, r);
    return mul(mul(F, D), G);
}
void frame_of(V3 N, V3& t, V3& b, bool nb) {
    V3 up = std::fabs(N.z) < 0.999f ? V3{0, 0, 1} : V3{1, 0, 0};
    t = nrm(cross(up, N));
    b = cross(N, t);
    if (nb) b = nrm(b);
}
struct Bs {
    V3 refl, L;
    float pdf;
};
Bs sample_diffuse(Surf& s, float u0, float u1) {
    float th = std::sqrt(u0), ph = 2.0f * kPI * u1;
    float sph, cph;
    sincos_ph(ph, sph, cph);
    V3 l{th * cph, th * sph, std::sqrt(1.0f - u0)};
    V3 t, b;
    frame_of(s.N, t, b, false);
    s.L = nrm((mul(t, l.x) + mul(b, l.y)) + mul(s.N, l.z));
    update(s);
    return Bs{diffuse(s), s.L, s.NdotL / kPI};
}
Bs sample_specular(Surf& s, float u0, float u1) {
    float alpha = s.m.roughness * s.m.roughness;
    s.Vv = nrm(s.Vv);
    V3 N = nrm(s.N), t, b;
    frame_of(N, t, b, true);
    V3 vt = nrm(V3{dot(s.Vv, t), dot(s.Vv, b), dot(s.Vv, N)});
    V3 vh = nrm(V3{vt.x * alpha, vt.y * alpha, vt.z});
    float phi = kPI * 2.0f * u0;
    float z = (1.0f - u1) * (1.0f + vh.z) + -vh.z;
    float r = std::sqrt(clampf(1.0f - z * z, 0.0f, 1.0f));
    float sph, cph;
    sincos_ph(phi, sph, cph);
    V3 c{cph * r, sph * r, z};
    V3 Hh = c + vh;
    V3 h{Hh.x * alpha, Hh.y * alpha, std::fmax(Hh.z, 0.0f)};
    V3 M = nrm((mul(t, h.x) + mul(b, h.y)) + mul(N, h.z));
    s.L = mul(M, 2.0f * dot(s.Vv, M)) - s.Vv;
    update(s);
    Bs o{V3{0, 0, 0}, s.L, 1.0f};
    if (s.NdotL > 0.0f && s.LdotH > 0.0f) {
        V3 F = schlick(s.F0, s.F90, s.LdotH);
        float Vis = vis_corr(s.NdotV, s.NdotL, alpha);
        float G1 = vis_sep(s.NdotV, alpha);
        o.pdf = G1 / (4.0f * std::fabs(dot(s.Vv, s.N)));
        o.refl = mul(F, Vis);
    }
    return o;
}

struct Scene {
    World w;
    std::vector<std::vector<float>> shade;    // per BLAS, slot order: n0, n1, n2 (9 floats)
    std::vector<std::vector<int32_t>> mat;    // per BLAS, slot order
    std::vector<std::vector<float>> opacity;  // per BLAS, slot order (MeshData.cpp:141), empty = opaque
    std::vector<std::vector<float>> uv;       // per BLAS, slot order: uv0, uv1, uv2 (6 floats), may be empty
    std::vector<std::vector<float>> tangent;  // per BLAS, slot order: tangent, bitangent (6 floats), may be empty
    std::vector<std::vector<float>> color;    // per BLAS, slot order: three RGBA vertex colours, may be empty
    bool alpha = false;                       // some triangle is not opaque
    const float* xform;                       // 12 floats per packed instance
    std::vector<Mat> mats;
    std::vector<Tex> tex;
    std::vector<atlas_rt_gpu_light> lights;   // the engine's GPULight list
    V3 env;
    int32_t bounces;
};

inline V3 xf(const float* M, V3 p, float w) {
    return V3{(M[0] * p.x + M[1] * p.y) + (M[2] * p.z + M[3] * w),
              (M[4] * p.x + M[5] * p.y) + (M[6] * p.z + M[7] * w),
              (M[8] * p.x + M[9] * p.y) + (M[10] * p.z + M[11] * w)};
}

// glm's mat4 * vec4 (columns summed in order) over the object-to-world rows.
inline V3 xf_glm(const float* M, V3 p, float w) {
    return V3{((M[0] * p.x + M[1] * p.y) + M[2] * p.z) + M[3] * w,
              ((M[4] * p.x + M[5] * p.y) + M[6] * p.z) + M[7] * w,
              ((M[8] * p.x + M[9] * p.y) + M[10] * p.z) + M[11] * w};
}

inline V3 uv_at(const std::vector<float>& uv, int32_t slot, float s, float t) {
    if (uv.empty()) return V3{0, 0, 0};
    const float* q = &uv[6 * size_t(slot)];
    const float r = (1.0f - s) - t;
    return V3{(r * q[0] + s * q[2]) + t * q[4], (r * q[1] + s * q[3]) + t * q[5], 0.0f};
}

// The triangle's opacity as the traversal sees it: the baked value, or for
// -1 GetOpacity (raytracer/surface.hsh:147-160) at the object-space (u, v).
float tri_opacity(const void* ctx, int32_t instance, int32_t slot, float u, float v) {
    const Scene& S = *static_cast<const Scene*>(ctx);
    const auto& in = S.w.inst[instance];
    const int32_t mesh = in.mesh_offset;
    if (S.opacity[mesh].empty()) return 1.0f;
    const float op = S.opacity[mesh][slot];
    if (op >= 0.0f) return op;
    const Mat& m = S.mats[size_t(S.mat[mesh][slot] + in.material_offset)];
    V3 tc = uv_at(S.uv[mesh], slot, u, v);
    if (m.invert_uvs > 0) tc.y = 1.0f - tc.y;
    return sample_opacity(S.tex, m.opacity_texture, tc.x, tc.y) * m.opacity;
}

// Radiance-cache bookkeeping of one path (both pointers null in plain mode).
// query: a render path that reaches a surface at bounce 1 stops there and
//   writes the cache query (position, roughness, -direction, 0, normal, 0),
//   the throughput reaching the vertex and a flag (16 floats).
// rec: a training path writes, per bounce k, 20 floats: the vertex's query
//   (12), the throughput reaching it with w = 1 for a surface hit (4), and
//   the radiance bounce k added to the pixel (4).
struct CacheOut {
    float* query = nullptr;
    float* rec = nullptr;
};

// The first bounce's surface, for the realtime G-buffer (rayHit.csh:76-117).
struct Primary {
    bool hit = false;
    V3 P{0, 0, 0}, Ng{0, 0, 0};
    int32_t material_id = 0;
    int32_t instance = -1;
};

// Traversal work of a path tracer call: rays traced (closest hits and
// shadow rays), inner nodes fetched, triangles tested, instances entered.
struct PtCount {
    uint64_t rays = 0, nodes = 0, tris = 0, insts = 0;
    void add(const TraceStats& st) {
        rays++;
        nodes += st.node_visits;
        tris += st.tri_tests;
        insts += st.inst_visits;
    }
};
std::mutex g_count_mu;
PtCount g_count;  // of the last oracle_pathtrace call

V3 trace_path(const Scene& S, V3 o, V3 d, float pixel_seed, const float* seeds, std::vector<int32_t>& stack,
              const CacheOut& co = CacheOut{}, Primary* prim = nullptr, PtCount* cnt = nullptr) {
    V3 thr{1, 1, 1}, Lr{0, 0, 0};
    for (int32_t bounce = 0; bounce <= S.bounces; bounce++) {
        const V3 thr_in = thr;
        float* rec = co.rec ? co.rec + 20 * size_t(bounce) : nullptr;
        HitRec h;
        TraceStats st;
        const Opacity op{&S, tri_opacity};
        closest_rgpu_world(S.w, o, d, 0.0f, kInf, ATLAS_RT_MASK_ALL, h, st, stack, S.alpha ? &op : nullptr);
        if (cnt) cnt->add(st);
        if (h.slot < 0) {
            if (prim && bounce == 0) prim->hit = false;
            V3 e = S.env * thr;
            const V3 ec{std::fmin(e.x, 10.0f), std::fmin(e.y, 10.0f), std::fmin(e.z, 10.0f)};
            Lr = Lr + ec;
            if (rec) {
                rec[16] = ec.x;
                rec[17] = ec.y;
                rec[18] = ec.z;
            }
            break;
        }
        const auto& in = S.w.inst[h.instance];
        const int32_t mesh = in.mesh_offset;
        const float* M = S.xform + 12 * size_t(h.instance);
        const auto& tri = S.w.blas[mesh].tris[h.slot];
        const float* ns = &S.shade[mesh][9 * size_t(h.slot)];
        V3 v0 = xf(M, V3{tri.v0[0], tri.v0[1], tri.v0[2]}, 1.0f);
        V3 v1 = xf(M, V3{tri.v1[0], tri.v1[1], tri.v1[2]}, 1.0f);
        V3 v2 = xf(M, V3{tri.v2[0], tri.v2[1], tri.v2[2]}, 1.0f);
        V3 n0 = nrm(xf(M, V3{ns[0], ns[1], ns[2]}, 0.0f));
        V3 n1 = nrm(xf(M, V3{ns[3], ns[4], ns[5]}, 0.0f));
        V3 n2 = nrm(xf(M, V3{ns[6], ns[7], ns[8]}, 0.0f));
        Mat m = S.mats[size_t(S.mat[mesh][h.slot] + in.material_offset)];
        V3 e0 = v1 - v0, e1 = v2 - v0, s = o - v0;
        V3 p = cross(s, e0), q = cross(d, e1);
        float den = dot(q, e0);
        float dist = dot(p, e1) / den, bs = dot(q, s) / den, bt = dot(p, d) / den;
        float br = 1.0f - bs - bt;
        // vertex colours (surface.hsh:84, 98): r c0 + s c1 + t c2, rgb only
        if (m.use_vertex_colors > 0 && !S.color[mesh].empty()) {
            const float* c = &S.color[mesh][12 * size_t(h.slot)];
            m.base = V3{m.base.x * ((br * c[0] + bs * c[4]) + bt * c[8]),
                        m.base.y * ((br * c[1] + bs * c[5]) + bt * c[9]),
                        m.base.z * ((br * c[2] + bs * c[6]) + bt * c[10])};
        }
        // material maps (surface.hsh:86-124) at the texture coordinate of the
        // world-space barycentrics
        V3 tc{0, 0, 0};
        const int32_t maps[6] = {m.base_color_texture, m.opacity_texture, m.roughness_texture,
                                 m.metalness_texture, m.ao_texture, m.normal_texture};
        bool textured = false;
        for (int32_t id : maps) textured = textured || has_tex(S.tex, id);
        if (textured) {
            tc = uv_at(S.uv[mesh], h.slot, bs, bt);
            if (m.invert_uvs > 0) tc.y = 1.0f - tc.y;
            float v[4];
            if (has_tex(S.tex, m.base_color_texture)) {
                sample_tex(S.tex[size_t(m.base_color_texture)], tc.x, tc.y, v);
                m.base = V3{m.base.x * v[0], m.base.y * v[1], m.base.z * v[2]};
            }
            m.opacity *= sample_opacity(S.tex, m.opacity_texture, tc.x, tc.y);
            if (has_tex(S.tex, m.roughness_texture)) m.roughness *= sample_opacity(S.tex, m.roughness_texture, tc.x, tc.y);
            if (has_tex(S.tex, m.metalness_texture)) m.metalness *= sample_opacity(S.tex, m.metalness_texture, tc.x, tc.y);
            if (has_tex(S.tex, m.ao_texture)) m.ao *= sample_opacity(S.tex, m.ao_texture, tc.x, tc.y);
        }
        Surf sf;
        sf.m = m;
        sf.P = o + mul(d, dist);
        V3 normal = nrm((mul(n0, br) + mul(n1, bs)) + mul(n2, bt));
        V3 Ng = nrm(cross(v0 - v1, v0 - v2));
        bool flip = dot(Ng, d) > 0.0f;
        Ng = mul(Ng, (flip && m.two_sided > 0) ? -1.0f : 1.0f);
        if (has_tex(S.tex, m.normal_texture)) {
            // TBN = mat3(tangent, +-bitangent, normal); texNormal /= max(length, 1e-5);
            // mix by saturate(normalScale) (surface.hsh:111-124)
            V3 tg{0, 0, 0}, bg{0, 0, 0};
            if (!S.tangent[mesh].empty()) {
                const float* tp = &S.tangent[mesh][6 * size_t(h.slot)];
                tg = nrm(xf(M, V3{tp[0], tp[1], tp[2]}, 0.0f));
                bg = nrm(xf(M, V3{tp[3], tp[4], tp[5]}, 0.0f));
            }
            const V3 bit = m.invert_uvs > 0 ? bg : neg(bg);
            float v[4];
            sample_tex(S.tex[size_t(m.normal_texture)], tc.x, tc.y, v);
            const float tx = 2.0f * v[0] - 1.0f, ty = 2.0f * v[1] - 1.0f, tz = 2.0f * v[2] - 1.0f;
            V3 tn = (mul(tg, tx) + mul(bit, ty)) + mul(normal, tz);
            const float len = std::sqrt(dot(tn, tn));
            const float den = (len < 0.00001f) ? 0.00001f : len;
            tn = V3{tn.x / den, tn.y / den, tn.z / den};
            const float wgt = (m.normal_scale < 0.0f) ? 0.0f : ((1.0f < m.normal_scale) ? 1.0f : m.normal_scale);
            normal = mul(normal, 1.0f - wgt) + mul(tn, wgt);
        }
        if (dot(normal, Ng) < 0.0f && m.two_sided > 0) normal = neg(normal);
        sf.Vv = neg(d);
        sf.N = nrm(normal);
        sf.Ng = Ng;
        if (prim && bounce == 0) {
            prim->hit = true;
            prim->P = sf.P;
            prim->Ng = Ng;
            prim->material_id = m.id;
            prim->instance = h.instance;
        }
        const float q12[12] = {sf.P.x, sf.P.y, sf.P.z, m.roughness, -d.x, -d.y, -d.z, 0.0f,
                               sf.N.x, sf.N.y, sf.N.z, 0.0f};
        if (co.query && bounce == 1) {
            std::memcpy(co.query, q12, sizeof(q12));
            co.query[12] = thr_in.x;
            co.query[13] = thr_in.y;
            co.query[14] = thr_in.z;
            co.query[15] = 1.0f;
            return Lr;
        }
        if (rec) {
            std::memcpy(rec, q12, sizeof(q12));
            rec[12] = thr_in.x;
            rec[13] = thr_in.y;
            rec[14] = thr_in.z;
            rec[15] = 1.0f;
        }
        sf.F0 = V3{mixf(0.04f, m.base.x, m.metalness), mixf(0.04f, m.base.y, m.metalness),
                   mixf(0.04f, m.base.z, m.metalness)};
        sf.F90 = 1.0f;
        sf.NdotL = 0.0f;
        V3 emis{0, 0, 0};
        if ((m.emissive.x + m.emissive.y) + m.emissive.z > 0.0f && bounce == 0) {
            emis = m.emissive;
            Lr = Lr + m.emissive;
        }
        // EvaluateDirectLight (rayHit.csh:209-235), GetLight / SampleLight
        // (raytracer/direct.hsh:12-86)
        V3 radiance{0, 0, 0};
        if (!S.lights.empty()) {
            float seed = seeds[bounce];
            float total = 0.0f, lpdf = 1.0f;
            size_t sel = 0;
            for (size_t l = 0; l < S.lights.size(); l++) {
                const atlas_rt_gpu_light& Lt = S.lights[l];
                const V3 lr{Lt.color[0], Lt.color[1], Lt.color[2]};
                const float brightness = dot(lr, V3{0.33333f, 0.33333f, 0.33333f});
                float wgt = 1.0f;
                if ((fbits(Lt.data[0]) >> 28) == 1u) {
                    const V3 pl = V3{Lt.P[0], Lt.P[1], Lt.P[2]} - sf.P;
                    const float sq = dot(pl, pl);
                    const V3 Ld = nrm(pl);
                    const float ndl = std::fmax(dot(V3{Lt.N[0], Lt.N[1], Lt.N[2]}, neg(Ld)), 0.0f);
                    wgt = ((brightness * Lt.data[2]) * ndl) / sq;
                } else {
                    wgt = brightness;
                }
                wgt = clampf(wgt, 0.0000001f, 1.0f);
                total += wgt;
                if (rand_seq(pixel_seed, seed) < wgt / total) {
                    lpdf = wgt;
                    sel = l;
                }
            }
            const atlas_rt_gpu_light& Lt = S.lights[sel];
            lpdf = lpdf / total * Lt.data[1];
            float solid = 1.0f, ldist = kInf;
            if ((fbits(Lt.data[0]) >> 28) == 1u) {
                const float r0 = rand_seq(pixel_seed, seed), r1 = rand_seq(pixel_seed, seed);
                const int32_t li = int32_t(fbits(Lt.data[3]));
                const int32_t lslot = int32_t(fbits(Lt.data[0]) & 0x0FFFFFFFu);
                const auto& lin = S.w.inst[li];
                const float* LM = S.xform + 12 * size_t(li);
                const auto& ltri = S.w.blas[lin.mesh_offset].tris[lslot];
                const float* lns = &S.shade[lin.mesh_offset][9 * size_t(lslot)];
                const V3 a0 = xf(LM, V3{ltri.v0[0], ltri.v0[1], ltri.v0[2]}, 1.0f);
                const V3 a1 = xf(LM, V3{ltri.v1[0], ltri.v1[1], ltri.v1[2]}, 1.0f);
                const V3 a2 = xf(LM, V3{ltri.v2[0], ltri.v2[1], ltri.v2[2]}, 1.0f);
                const V3 m0 = nrm(xf(LM, V3{lns[0], lns[1], lns[2]}, 0.0f));
                const V3 m1 = nrm(xf(LM, V3{lns[3], lns[4], lns[5]}, 0.0f));
                const V3 m2 = nrm(xf(LM, V3{lns[6], lns[7], lns[8]}, 0.0f));
                // SampleTriangleLight (raytracer/lights.hsh:11-31)
                const float sr = std::sqrt(r0);
                const float u = 1.0f - sr, v = r1 * sr, w = (1.0f - u) - v;
                const V3 LP = (mul(a0, u) + mul(a1, v)) + mul(a2, w);
                const V3 LN = (mul(m0, u) + mul(m1, v)) + mul(m2, w);
                const float spdf = 1.0f / Lt.data[2];
                const V3 pl = LP - sf.P;
                const float sq = dot(pl, pl);
                ldist = std::sqrt(sq);
                sf.L = V3{pl.x / ldist, pl.y / ldist, pl.z / ldist};
                update(sf);
                const float ndl = std::fmax(dot(LN, neg(sf.L)), 0.0f);
                solid = ndl / (sq * spdf);
            } else {
                sf.L = nrm(neg(V3{Lt.N[0], Lt.N[1], Lt.N[2]}));
                update(sf);
            }
            V3 refl = mul(diffuse(sf) + specular(sf), m.opacity);
            V3 rad = mul(V3{Lt.color[0], Lt.color[1], Lt.color[2]}, solid);
            float vis = 0.0f;
            if (sf.NdotL > 0.0f) {
                TraceStats st2;
                const V3 so = sf.P + mul(sf.N, kEps);
                const float tmax = ldist - 2.0f * kEps;
                if (S.alpha) {
                    vis = transparency_rgpu_world(S.w, so, sf.L, 0.0f, tmax, ATLAS_RT_MASK_SHADOW, st2, stack,
                                                  Opacity{&S, tri_opacity});
                } else {
                    bool occluded = any_rgpu_world(S.w, so, sf.L, 0.0f, tmax, ATLAS_RT_MASK_SHADOW, st2, stack, true);
                    vis = occluded ? 0.0f : 1.0f;
                }
                if (cnt) cnt->add(st2);
            }
            V3 direct = mul(refl * mul(rad, vis), sf.NdotL);
            direct = V3{direct.x / lpdf, direct.y / lpdf, direct.z / lpdf};
            radiance = mul(thr, m.opacity) * direct;
            if (bounce > 0) {
                float rm = std::fmax(std::fmax(radiance.x, std::fmax(radiance.y, radiance.z)), 10.0f);
                radiance = mul(radiance, 10.0f / rm);
            }
        }
        Lr = Lr + radiance;
        if (rec) {
            // emission, then the shadow ray's radiance (the GPU's order)
            rec[16] = emis.x + radiance.x;
            rec[17] = emis.y + radiance.y;
            rec[18] = emis.z + radiance.z;
        }
        // EvaluateIndirectLight
        V3 no = sf.P;
        float seed = seeds[bounce];
        float refract = clampf(1.0f - m.opacity, 0.1f, 0.9f);
        refract = m.opacity == 1.0f ? 0.0f : refract;
        float rnd = sat(rand_seq(pixel_seed, seed));
        Bs B;
        bool refracted = false;
        if (rnd >= refract) {
            rnd = rand_seq(pixel_seed, seed);
            V3 F = schlick(sf.F0, sf.F90, sf.NdotV);
            float spec = clampf(dot(F, V3{0.33333f, 0.33333f, 0.33333f}), 0.1f, 0.9f);
            float u0 = rand_seq(pixel_seed, seed), u1 = rand_seq(pixel_seed, seed);
            if (rnd < spec) {
                B = sample_specular(sf, u0, u1);
                B.refl = mul(B.refl, m.opacity);
                B.pdf *= spec;
            } else {
                B = sample_diffuse(sf, u0, u1);
                B.refl = mul(B.refl, (1.0f - m.metalness) * m.opacity);
                B.pdf *= (1.0f - spec);
            }
            B.pdf *= (1.0f - refract);
            no = no + mul(sf.Vv, kEps);
        } else {
            B.L = d;
            B.pdf = refract;
            B.refl = V3{1.0f - m.opacity, 1.0f - m.opacity, 1.0f - m.opacity};
            sf.NdotL = 1.0f;
            no = no - mul(sf.N, kEps);
            refracted = true;
        }
        if (B.pdf > 0.0f && (B.refl.x + B.refl.y) + B.refl.z > 0.0f) {
            V3 k = mul(B.refl, sf.NdotL);
            thr = thr * V3{k.x / B.pdf, k.y / B.pdf, k.z / B.pdf};
        } else {
            thr = V3{0, 0, 0};
        }
        V3 nd = nrm(B.L);
        thr = mul(thr, m.ao);
        float prob = clampf(std::fmax(thr.x, std::fmax(thr.y, thr.z)), 0.01f, 0.99f);
        prob = bounce < 3 ? std::fmin(3.0f * prob, 1.0f) : prob;
        if (rand_seq(pixel_seed, seed) > prob)
            thr = V3{0, 0, 0};
        else if (dot(nd, sf.Ng) <= 0.0f && !refracted)
            thr = V3{0, 0, 0};
        else
            thr = V3{thr.x / prob, thr.y / prob, thr.z / prob};
        if ((thr.x + thr.y) + thr.z == 0.0f) break;
        o = no;
        d = nd;
    }
    return Lr;
}

}  // namespace
}  // namespace oracle

using namespace oracle;

namespace {

uint32_t ubits(float f) {
    uint32_t u;
    std::memcpy(&u, &f, 4);
    return u;
}
float bitsf(uint32_t u) {
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}

// Triangle lights of every instance: BuildTriangleLightsForMesh
// (RayTracingWorld.cpp:322-376) per BLAS slot, then UpdateTriangleLights
// (:378-407) per instance of that BLAS.  Emitter e = material index +
// instance material_offset holds (sRGB emissive colour, emissive intensity).
// Order: BLAS, then slot, then instance slot.
void triangle_lights(const Scene& S, const float* emitters, uint32_t n_emitters,
                     std::vector<atlas_rt_gpu_light>& out) {
    for (uint32_t k = 0; k < S.w.blas_count; k++) {
        const BlasView& b = S.w.blas[k];
        for (uint32_t i = 0; i < b.tri_count; i++) {
            const auto& tri = b.tris[i];
            const float* ns = &S.shade[k][9 * size_t(i)];
            for (uint32_t j = 0; j < S.w.inst_count; j++) {
                const auto& in = S.w.inst[j];
                if (in.mesh_offset != int32_t(k)) continue;
                const int64_t e = int64_t(S.mat[k][i]) + in.material_offset;
                if (e < 0 || e >= int64_t(n_emitters)) continue;
                const float* em = emitters + 4 * size_t(e);
                const V3 lin{std::pow(em[0], 2.2f), std::pow(em[1], 2.2f), std::pow(em[2], 2.2f)};
                const float brightness = dot(lin, V3{0.3333f, 0.3333f, 0.3333f});
                if (!(brightness > 0.0f)) continue;
                const V3 v0{tri.v0[0], tri.v0[1], tri.v0[2]}, v1{tri.v1[0], tri.v1[1], tri.v1[2]},
                    v2{tri.v2[0], tri.v2[1], tri.v2[2]};
                const V3 c = (v0 + v1) + v2, nsum = (V3{ns[0], ns[1], ns[2]} + V3{ns[3], ns[4], ns[5]}) +
                                                    V3{ns[6], ns[7], ns[8]};
                const V3 P{c.x / 3.0f, c.y / 3.0f, c.z / 3.0f}, N{nsum.x / 3.0f, nsum.y / 3.0f, nsum.z / 3.0f};
                const V3 ea = v1 - v0, eb = v2 - v0, ec = v1 - v2;
                const float a = std::sqrt(dot(ea, ea)), bb = std::sqrt(dot(eb, eb)), cc = std::sqrt(dot(ec, ec));
                const float p = 0.5f * ((a + bb) + cc);
                const float area = std::sqrt(((p * (p - a)) * (p - bb)) * (p - cc));
                const V3 col{std::pow(lin.x, 2.2f) * em[3], std::pow(lin.y, 2.2f) * em[3],
                             std::pow(lin.z, 2.2f) * em[3]};
                const float* M = S.xform + 12 * size_t(j);
                const V3 WP = xf_glm(M, P, 1.0f), WN = xf_glm(M, N, 0.0f);
                atlas_rt_gpu_light L{};
                L.P[0] = WP.x; L.P[1] = WP.y; L.P[2] = WP.z; L.P[3] = 1.0f;
                L.N[0] = WN.x; L.N[1] = WN.y; L.N[2] = WN.z; L.N[3] = 0.0f;
                L.color[0] = col.x; L.color[1] = col.y; L.color[2] = col.z; L.color[3] = 0.0f;
                L.data[0] = bitsf((1u << 28) | i);
                L.data[1] = area * brightness;
                L.data[2] = area;
                L.data[3] = bitsf(j);
                out.push_back(L);
            }
        }
    }
}

void make_scene(Scene& S, std::vector<BlasView>& views, const atlas_rt_gpu_bvh_node* tlas, uint32_t tlas_count,
                const atlas_rt_gpu_bvh_instance* inst, uint32_t inst_count, const atlas_rt_gpu_bvh_node* const* blas_nodes,
                const uint32_t* blas_node_counts, const atlas_rt_gpu_bvh_triangle* const* blas_tris,
                const uint32_t* blas_tri_counts, const float* const* shade_normals, const int32_t* const* shade_mat,
                uint32_t blas_count, const float* xform, const atlas_rt_material* mats, uint32_t n_mats,
                const atlas_rt_light* lights, uint32_t n_lights, const float* env, int32_t bounces,
                const OracleSurface* ext) {
    views.resize(blas_count);
    for (uint32_t k = 0; k < blas_count; k++)
        views[k] = BlasView{blas_nodes[k], blas_node_counts[k], blas_tris[k], blas_tri_counts[k]};
    S.w = World{tlas, tlas_count, inst, inst_count, views.data(), blas_count};
    S.xform = xform;
    S.shade.resize(blas_count);
    S.mat.resize(blas_count);
    S.opacity.resize(blas_count);
    S.uv.resize(blas_count);
    S.tangent.resize(blas_count);
    S.color.resize(blas_count);
    for (uint32_t k = 0; k < blas_count; k++) {
        S.shade[k].assign(shade_normals[k], shade_normals[k] + 9 * size_t(blas_tri_counts[k]));
        S.mat[k].assign(shade_mat[k], shade_mat[k] + blas_tri_counts[k]);
        if (ext && ext->opacity && ext->opacity[k]) {
            S.opacity[k].assign(ext->opacity[k], ext->opacity[k] + blas_tri_counts[k]);
            for (float o : S.opacity[k]) S.alpha = S.alpha || o != 1.0f;
        }
        if (ext && ext->uvs && ext->uvs[k]) S.uv[k].assign(ext->uvs[k], ext->uvs[k] + 6 * size_t(blas_tri_counts[k]));
        if (ext && ext->tangents && ext->tangents[k])
            S.tangent[k].assign(ext->tangents[k], ext->tangents[k] + 6 * size_t(blas_tri_counts[k]));
        if (ext && ext->colors && ext->colors[k])
            S.color[k].assign(ext->colors[k], ext->colors[k] + 12 * size_t(blas_tri_counts[k]));
    }
    for (uint32_t i = 0; i < n_mats; i++) {
        const auto& m = mats[i];
        S.mats.push_back(Mat{V3{m.base_color[0], m.base_color[1], m.base_color[2]},
                             V3{m.emissive[0], m.emissive[1], m.emissive[2]}, m.opacity, m.roughness, m.metalness,
                             m.ao, m.reflectance, m.two_sided, m.opacity_texture, m.invert_uvs, m.id,
                             m.normal_scale, m.base_color_texture, m.roughness_texture, m.metalness_texture,
                             m.ao_texture, m.normal_texture, m.use_vertex_colors});
    }
    if (ext)
        for (uint32_t t = 0; t < ext->texture_count; t++)
            S.tex.push_back(Tex{ext->textures[t].width, ext->textures[t].height, ext->textures[t].texels,
                                ext->textures[t].channels ? ext->textures[t].channels : 1u});
    // UpdateLights (RayTracingHelper.cpp:436-500): punctual lights first, then
    // the triangle lights; weights floored at 0.5% of the maximum, normalised.
    for (uint32_t l = 0; l < n_lights; l++) {
        const float* r = lights[l].radiance;
        atlas_rt_gpu_light L{};
        L.P[3] = 1.0f;
        std::memcpy(L.N, lights[l].direction, 12);
        std::memcpy(L.color, r, 12);
        L.data[0] = bitsf((0u << 28) | l);
        L.data[1] = (r[0] * 0.3333f + r[1] * 0.3333f) + r[2] * 0.3333f;
        S.lights.push_back(L);
    }
    if (ext && ext->emitters) triangle_lights(S, ext->emitters, ext->emitter_count, S.lights);
    float maxw = 0.0f, total = 0.0f;
    for (const auto& L : S.lights) maxw = std::fmax(maxw, L.data[1]);
    for (auto& L : S.lights) {
        L.data[1] = std::fmax(L.data[1], 0.005f * maxw);
        total += L.data[1];
    }
    for (auto& L : S.lights) L.data[1] /= total;
    if (ext && ext->lights_out) {
        const uint32_t n = uint32_t(std::min<size_t>(S.lights.size(), ext->lights_cap));
        std::memcpy(ext->lights_out, S.lights.data(), n * sizeof(atlas_rt_gpu_light));
        *ext->lights_count = uint32_t(S.lights.size());
    }
    S.env = V3{env[0], env[1], env[2]};
    S.bounces = bounces;
}

// Frame jitter and camera ray of pixel `pix` (rayGen.csh:25-90).
V3 camera_dir(const atlas_rt_camera* cam, uint32_t pix, uint32_t width, uint32_t height, float jx, float jy) {
    const V3 loc{cam->location[0], cam->location[1], cam->location[2]};
    const V3 org{cam->origin[0], cam->origin[1], cam->origin[2]};
    const V3 rgt{cam->right[0], cam->right[1], cam->right[2]};
    const V3 bot{cam->bottom[0], cam->bottom[1], cam->bottom[2]};
    const uint32_t px = pix % width, py = pix / width;
    float cx = (float(px) + jx) / float(width);
    float cy = (float(py) + jy) / float(height);
    return nrm(((org + mul(rgt, cx)) + mul(bot, cy)) - loc);
}

template <class F>
void parallel_pixels(uint32_t n, int32_t threads, F&& work) {
    if (threads <= 1) {
        work(0u, n);
        return;
    }
    std::vector<std::thread> pool;
    uint32_t chunk = (n + uint32_t(threads) - 1) / uint32_t(threads);
    for (int t = 0; t < threads; t++) {
        uint32_t lo = uint32_t(t) * chunk, hi = std::min(n, lo + chunk);
        if (lo < hi) pool.emplace_back(work, lo, hi);
    }
    for (auto& th : pool) th.join();
}

// Training-pixel choice of the cache mode: a murmur3-finalised hash of the
// pixel and frame index, one pixel in `stride`.
uint32_t mix32(uint32_t h) {
    h ^= h >> 16;
    h *= 0x85ebca6bu;
    h ^= h >> 13;
    h *= 0xc2b2ae35u;
    h ^= h >> 16;
    return h;
}

}  // namespace

extern "C" {

// Bilinear opacity-texture sample (texture.hsh:18-25 with the global sampler).
float oracle_sample_opacity(uint32_t width, uint32_t height, const uint8_t* texels, float x, float y) {
    std::vector<Tex> t{Tex{width, height, texels, 1u}};
    return sample_opacity(t, 0, x, y);
}

// Transparency along a world ray (HitAnyTransparency) and the closest hit of
// HitClosestTransparency, with per-slot baked opacity per BLAS (no maps).
struct OpacityTable {
    const World* w;
    const float* const* opacity;
};
static float table_opacity(const void* ctx, int32_t instance, int32_t slot, float, float) {
    const OpacityTable& T = *static_cast<const OpacityTable*>(ctx);
    return T.opacity[T.w->inst[instance].mesh_offset][slot];
}

void oracle_trace_world_alpha(const atlas_rt_gpu_bvh_node* tlas, uint32_t tlas_count,
                              const atlas_rt_gpu_bvh_instance* inst, uint32_t inst_count,
                              const atlas_rt_gpu_bvh_node* const* blas_nodes, const uint32_t* blas_node_counts,
                              const atlas_rt_gpu_bvh_triangle* const* blas_tris, const uint32_t* blas_tri_counts,
                              uint32_t blas_count, const float* const* opacity, const float* rays, uint32_t n,
                              uint32_t mask, float* transparency, int32_t* slot, int32_t* instance, float* t) {
    std::vector<BlasView> views(blas_count);
    for (uint32_t k = 0; k < blas_count; k++)
        views[k] = BlasView{blas_nodes[k], blas_node_counts[k], blas_tris[k], blas_tri_counts[k]};
    World w{tlas, tlas_count, inst, inst_count, views.data(), blas_count};
    OpacityTable T{&w, opacity};
    const Opacity op{&T, table_opacity};
    std::vector<int32_t> stack;
    for (uint32_t i = 0; i < n; i++) {
        const float* r = rays + 8 * size_t(i);
        TraceStats st;
        transparency[i] = transparency_rgpu_world(w, V3{r[0], r[1], r[2]}, V3{r[4], r[5], r[6]}, r[3], r[7], mask, st,
                                                  stack, op);
        HitRec h;
        closest_rgpu_world(w, V3{r[0], r[1], r[2]}, V3{r[4], r[5], r[6]}, r[3], r[7], mask, h, st, stack, &op);
        slot[i] = h.slot;
        instance[i] = h.instance;
        t[i] = h.t;
    }
}

// shade_normals[k]: 9 floats per slot of BLAS k (BVH slot order);
// shade_mat[k]: one material index per slot.  accum: width*height*4 floats
// (rgb sum, count), accumulated over `frames` frames starting at sample
// `first_sample`; pcg_draws: number of Random draws made before this call
// (the PCG stream position).
void oracle_pathtrace(const atlas_rt_gpu_bvh_node* tlas, uint32_t tlas_count, const atlas_rt_gpu_bvh_instance* inst,
                      uint32_t inst_count, const atlas_rt_gpu_bvh_node* const* blas_nodes,
                      const uint32_t* blas_node_counts, const atlas_rt_gpu_bvh_triangle* const* blas_tris,
                      const uint32_t* blas_tri_counts, const float* const* shade_normals,
                      const int32_t* const* shade_mat, uint32_t blas_count, const float* xform,
                      const atlas_rt_material* mats, uint32_t n_mats, const atlas_rt_light* lights, uint32_t n_lights,
                      const float* env, const atlas_rt_camera* cam, uint32_t width, uint32_t height, int32_t bounces,
                      uint32_t frames, uint32_t first_sample, uint64_t pcg_draws, float* accum, int32_t threads,
                      const OracleSurface* ext) {
    Scene S;
    std::vector<BlasView> views;
    make_scene(S, views, tlas, tlas_count, inst, inst_count, blas_nodes, blas_node_counts, blas_tris, blas_tri_counts,
               shade_normals, shade_mat, blas_count, xform, mats, n_mats, lights, n_lights, env, bounces, ext);
    Pcg pcg{0ull + 1442695040888963407ull};
    pcg.next();  // pcg32_init's discarded draw
    for (uint64_t i = 0; i < pcg_draws; i++) pcg.next();
    const V3 loc{cam->location[0], cam->location[1], cam->location[2]};
    g_count = PtCount{};
    for (uint32_t f = 0; f < frames; f++) {
        std::vector<float> seeds(size_t(bounces) + 1);
        for (int32_t b = 0; b <= bounces; b++) seeds[b] = pcg.fast_uniform();
        const float sc = float(first_sample + f);
        const float jx = rand2(sc, 0.0f), jy = rand2(sc, 1.0f);
        parallel_pixels(width * height, threads, [&](uint32_t lo, uint32_t hi) {
            std::vector<int32_t> stack;
            PtCount cnt;
            for (uint32_t pix = lo; pix < hi; pix++) {
                V3 L = trace_path(S, loc, camera_dir(cam, pix, width, height, jx, jy), float(pix), seeds.data(), stack,
                                  CacheOut{}, nullptr, &cnt);
                accum[4 * size_t(pix)] += L.x;
                accum[4 * size_t(pix) + 1] += L.y;
                accum[4 * size_t(pix) + 2] += L.z;
                accum[4 * size_t(pix) + 3] += 1.0f;
            }
            std::lock_guard<std::mutex> lock(g_count_mu);
            g_count.rays += cnt.rays;
            g_count.nodes += cnt.nodes;
            g_count.tris += cnt.tris;
            g_count.insts += cnt.insts;
        });
    }
}

// One frame of the radiance-cache mode (the first frame of a fresh path
// tracer when frame_index = 0 and pcg_draws = 0): per pixel, radiance[4] is
// what the path added before any cache term; a render path that reaches a
// surface at bounce 1 writes query[16] (see CacheOut) instead of going on;
// a training pixel (train[pix] = 1) traces the full path and writes
// rec[(bounces + 1) * 20].
void oracle_pathtrace_cache(const atlas_rt_gpu_bvh_node* tlas, uint32_t tlas_count,
                            const atlas_rt_gpu_bvh_instance* inst, uint32_t inst_count,
                            const atlas_rt_gpu_bvh_node* const* blas_nodes, const uint32_t* blas_node_counts,
                            const atlas_rt_gpu_bvh_triangle* const* blas_tris, const uint32_t* blas_tri_counts,
                            const float* const* shade_normals, const int32_t* const* shade_mat, uint32_t blas_count,
                            const float* xform, const atlas_rt_material* mats, uint32_t n_mats,
                            const atlas_rt_light* lights, uint32_t n_lights, const float* env,
                            const atlas_rt_camera* cam, uint32_t width, uint32_t height, int32_t bounces,
                            uint32_t frame_index, uint32_t sample_index, uint64_t pcg_draws, uint32_t train_stride,
                            float* radiance, float* query, int32_t* train, float* rec, int32_t threads,
                            const OracleSurface* ext) {
    Scene S;
    std::vector<BlasView> views;
    make_scene(S, views, tlas, tlas_count, inst, inst_count, blas_nodes, blas_node_counts, blas_tris, blas_tri_counts,
               shade_normals, shade_mat, blas_count, xform, mats, n_mats, lights, n_lights, env, bounces, ext);
    Pcg pcg{0ull + 1442695040888963407ull};
    pcg.next();
    for (uint64_t i = 0; i < pcg_draws; i++) pcg.next();
    std::vector<float> seeds(size_t(bounces) + 1);
    for (int32_t b = 0; b <= bounces; b++) seeds[b] = pcg.fast_uniform();
    const float sc = float(sample_index);
    const float jx = rand2(sc, 0.0f), jy = rand2(sc, 1.0f);
    const V3 loc{cam->location[0], cam->location[1], cam->location[2]};
    const size_t per_rec = 20 * size_t(bounces + 1);
    parallel_pixels(width * height, threads, [&](uint32_t lo, uint32_t hi) {
        std::vector<int32_t> stack;
        for (uint32_t pix = lo; pix < hi; pix++) {
            const bool tr = mix32(pix * 0x9e3779b1u + frame_index * 0x632be5abu) % train_stride == 0;
            train[pix] = tr ? 1 : 0;
            CacheOut co;
            if (tr) {
                co.rec = rec + per_rec * pix;
                std::fill(co.rec, co.rec + per_rec, 0.0f);
            } else {
                co.query = query + 16 * size_t(pix);
                std::fill(co.query, co.query + 16, 0.0f);
            }
            V3 L = trace_path(S, loc, camera_dir(cam, pix, width, height, jx, jy), float(pix), seeds.data(), stack, co);
            radiance[4 * size_t(pix)] = L.x;
            radiance[4 * size_t(pix) + 1] = L.y;
            radiance[4 * size_t(pix) + 2] = L.z;
            radiance[4 * size_t(pix) + 3] = 1.0f;
        }
    });
}

// IEEE binary16 of a float, round to nearest even (the R16G16_SFLOAT store).
uint16_t oracle_f2h(float f) {
    uint32_t x;
    std::memcpy(&x, &f, 4);
    const uint32_t sign = (x >> 16) & 0x8000u, mant = x & 0x7FFFFFu;
    const int32_t exp = int32_t((x >> 23) & 0xFFu);
    if (exp == 0xFF) return uint16_t(sign | 0x7C00u | (mant ? 0x200u | (mant >> 13) : 0u));
    const int32_t e = exp - 127 + 15;
    if (e >= 31) return uint16_t(sign | 0x7C00u);
    if (e <= 0) {
        if (e < -10) return uint16_t(sign);
        const uint32_t m = mant | 0x800000u;
        const uint32_t shift = uint32_t(14 - e);
        uint32_t h = m >> shift;
        const uint32_t rem = m & ((1u << shift) - 1u), half = 1u << (shift - 1u);
        if (rem > half || (rem == half && (h & 1u))) h++;
        return uint16_t(sign | h);
    }
    uint32_t h = (uint32_t(e) << 10) | (mant >> 13);
    const uint32_t rem = mant & 0x1FFFu;
    if (rem > 0x1000u || (rem == 0x1000u && (h & 1u))) h++;
    return uint16_t(sign | h);
}

float oracle_h2f(uint16_t h) {
    const uint32_t sign = uint32_t(h & 0x8000u) << 16;
    const uint32_t exp = (h >> 10) & 0x1Fu, mant = h & 0x3FFu;
    uint32_t x;
    if (exp == 0) {
        if (mant == 0) {
            x = sign;
        } else {  // subnormal: renormalise
            int32_t e = -1;
            uint32_t m = mant;
            do {
                e++;
                m <<= 1;
            } while (!(m & 0x400u));
            x = sign | (uint32_t(127 - 15 - e) << 23) | ((m & 0x3FFu) << 13);
        }
    } else if (exp == 31) {
        x = sign | 0x7F800000u | (mant << 13);
    } else {
        x = sign | ((exp - 15 + 127) << 23) | (mant << 13);
    }
    float f;
    std::memcpy(&f, &x, 4);
    return f;
}

// One realtime frame (PathTracingRenderer with realTime; rayGen.csh and
// rayHit.csh with REALTIME): spf paths per pixel (ray ID = pixel * spf +
// sample) through the pixel centre plus jitter / 4, seeds after pcg_draws
// draws; per pixel the radiance as temporal.csh's FetchTexel reads it
// (frame[4]), and from sample 0's first bounce the G-buffer: view-space depth
// (1e12 on a miss), velocity and EncodeNormal as binary16 pairs (low half =
// first component), material ID (65535 on a miss).  view / proj / last_vp:
// glm column-major mat4; last_xform: last frame's object-to-world rows (12
// floats per instance slot).
void oracle_pathtrace_realtime(const atlas_rt_gpu_bvh_node* tlas, uint32_t tlas_count,
                               const atlas_rt_gpu_bvh_instance* inst, uint32_t inst_count,
                               const atlas_rt_gpu_bvh_node* const* blas_nodes, const uint32_t* blas_node_counts,
                               const atlas_rt_gpu_bvh_triangle* const* blas_tris, const uint32_t* blas_tri_counts,
                               const float* const* shade_normals, const int32_t* const* shade_mat,
                               uint32_t blas_count, const float* xform, const atlas_rt_material* mats,
                               uint32_t n_mats, const atlas_rt_light* lights, uint32_t n_lights, const float* env,
                               const atlas_rt_camera* cam, uint32_t width, uint32_t height, int32_t bounces,
                               uint32_t spf, uint64_t pcg_draws, const float* view, const float* proj,
                               const float* last_vp, const float* jitter, const float* last_xform,
                               float max_radiance, float* frame, float* depth, uint32_t* velocity,
                               uint32_t* normal, uint16_t* material, int32_t threads, const OracleSurface* ext) {
    Scene S;
    std::vector<BlasView> views;
    make_scene(S, views, tlas, tlas_count, inst, inst_count, blas_nodes, blas_node_counts, blas_tris, blas_tri_counts,
               shade_normals, shade_mat, blas_count, xform, mats, n_mats, lights, n_lights, env, bounces, ext);
    Pcg pcg{0ull + 1442695040888963407ull};
    pcg.next();
    for (uint64_t i = 0; i < pcg_draws; i++) pcg.next();
    std::vector<float> seeds(size_t(bounces) + 1);
    for (int32_t b = 0; b <= bounces; b++) seeds[b] = pcg.fast_uniform();
    const V3 loc{cam->location[0], cam->location[1], cam->location[2]};
    const V3 org{cam->origin[0], cam->origin[1], cam->origin[2]};
    const V3 rgt{cam->right[0], cam->right[1], cam->right[2]};
    const V3 bot{cam->bottom[0], cam->bottom[1], cam->bottom[2]};
    auto gmax = [](float x, float y) { return (x < y) ? y : x; };
    auto gmin = [](float x, float y) { return (y < x) ? y : x; };
    auto mat4 = [](const float* m, float x, float y, float z, float w, float out[4]) {
        for (int k = 0; k < 4; k++) out[k] = (m[k] * x + m[4 + k] * y) + (m[8 + k] * z + m[12 + k] * w);
    };
    parallel_pixels(width * height, threads, [&](uint32_t lo, uint32_t hi) {
        std::vector<int32_t> stack;
        for (uint32_t pix = lo; pix < hi; pix++) {
            const uint32_t px = pix % width, py = pix / width;
            const float cx = jitter[0] * 0.25f + (float(px) + 0.5f) / float(width);
            const float cy = jitter[1] * 0.25f + (float(py) + 0.5f) / float(height);
            const V3 dir = nrm(((org + mul(rgt, cx)) + mul(bot, cy)) - loc);
            Primary prim;
            std::vector<V3> L(spf);
            for (uint32_t k = 0; k < spf; k++)
                L[k] = trace_path(S, loc, dir, float(pix * spf + k), seeds.data(), stack, CacheOut{},
                                  k == 0 ? &prim : nullptr);
            float c[3];
            if (spf == 1) {
                c[0] = L[0].x;
                c[1] = L[0].y;
                c[2] = L[0].z;
            } else {
                const uint32_t per = 0xFFFFFFFFu / spf;
                uint32_t sum[3] = {0u, 0u, 0u};
                for (uint32_t k = 0; k < spf; k++) {
                    const float v[3] = {L[k].x, L[k].y, L[k].z};
                    for (int a = 0; a < 3; a++) {
                        const float q = gmin(gmax(v[a] / max_radiance, 0.0f), 1.0f) * float(per);
                        const uint32_t u = q >= 0.0f ? uint32_t(q) : 0u;
                        sum[a] += std::min(u, per);
                    }
                }
                for (int a = 0; a < 3; a++) c[a] = float(sum[a]) / float(per) / float(spf) * max_radiance;
            }
            for (int a = 0; a < 3; a++) frame[4 * size_t(pix) + a] = gmax(c[a], 0.0f);
            frame[4 * size_t(pix) + 3] = 0.0f;
            // G-buffer (rayHit.csh:76-117)
            V3 P, Ng;
            float inv[12] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0}, last[12] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0};
            uint32_t mat_id = 65535u;
            if (prim.hit) {
                P = prim.P;
                Ng = prim.Ng;
                mat_id = uint32_t(prim.material_id) & 0xFFFFu;
                std::memcpy(inv, inst[prim.instance].inverse_matrix, 48);
                std::memcpy(last, last_xform + 12 * size_t(prim.instance), 48);
            } else {
                P = loc + mul(dir, 100.0f);
                Ng = V3{1.0f, 0.0f, 0.0f};
            }
            const V3 obj = xf_glm(inv, P, 1.0f);
            const V3 lp = xf(last, obj, 1.0f);
            float vs[4], pc[4], pl[4];
            mat4(view, P.x, P.y, P.z, 1.0f, vs);
            mat4(proj, vs[0], vs[1], vs[2], vs[3], pc);
            mat4(last_vp, lp.x, lp.y, lp.z, 1.0f, pl);
            const float vx = (pl[0] / pl[3] - pc[0] / pc[3]) * 0.5f;
            const float vy = (pl[1] / pl[3] - pc[1] / pc[3]) * 0.5f;
            velocity[pix] = uint32_t(oracle_f2h(vx)) | (uint32_t(oracle_f2h(vy)) << 16);
            depth[pix] = prim.hit ? vs[2] : 1.0e12f;
            material[pix] = uint16_t(mat_id);
            const float l1 = (std::fabs(Ng.x) + std::fabs(Ng.y)) + std::fabs(Ng.z);
            float nx = Ng.x / l1, nz = Ng.z / l1;
            if (Ng.y < 0.0f) {
                const float a = nx, b = nz;
                nx = (a >= 0.0f ? 1.0f : -1.0f) * (1.0f - std::fabs(b));
                nz = (b >= 0.0f ? 1.0f : -1.0f) * (1.0f - std::fabs(a));
            }
            normal[pix] = uint32_t(oracle_f2h(nx)) | (uint32_t(oracle_f2h(nz)) << 16);
        }
    });
}

// The restatement's sin / cos of sample angles and pow(x, 5), for tests.
void oracle_sincos(const float* x, float* sn, float* cs, float* p5, uint32_t n) {
    for (uint32_t i = 0; i < n; i++) {
        oracle::sincos_ph(x[i], sn[i], cs[i]);
        p5[i] = oracle::pow5f(x[i]);
    }
}

// Traversal work of the last oracle_pathtrace call: rays traced (closest
// hits and shadow rays), inner nodes fetched (TLAS and BLAS), triangles
// tested, instances entered.
void oracle_pathtrace_counts(uint64_t* out) {
    out[0] = g_count.rays;
    out[1] = g_count.nodes;
    out[2] = g_count.tris;
    out[3] = g_count.insts;
}

// UnpackTriangle (raytracer/common.hsh:17-44) of GPUTriangle records:
// per record the vertex normals (9 floats, unpackUnitVector of v0..v2.w),
// the material index (d0.w bits), the opacity (d2.w), the texture
// coordinates (6 floats, unpackHalf2x16 of d0.xyz), the tangent and
// bitangent (6 floats, unpackUnitVector of d1.xy) and the vertex colours
// (12 floats, unpackUnorm4x8 of d2.xyz).
void oracle_unpack_triangles(const atlas_rt_gpu_triangle* r, uint32_t n, float* normals, int32_t* material,
                             float* opacity, float* uvs, float* tangents, float* colors) {
    auto bits = [](float f) {
        uint32_t u;
        std::memcpy(&u, &f, 4);
        return u;
    };
    // unpackUnitVector (common/packing.hsh:2-12): k / 1023 * 2 - 1 per 10-bit field
    auto unit = [](uint32_t c, float* o) {
        for (int k = 0; k < 3; k++) o[k] = float((c >> (10 * k)) & 1023u) / 1023.0f * 2.0f - 1.0f;
    };
    for (uint32_t i = 0; i < n; i++) {
        const atlas_rt_gpu_triangle& t = r[i];
        unit(bits(t.v0[3]), normals + 9 * size_t(i));
        unit(bits(t.v1[3]), normals + 9 * size_t(i) + 3);
        unit(bits(t.v2[3]), normals + 9 * size_t(i) + 6);
        material[i] = int32_t(bits(t.d0[3]));
        opacity[i] = t.d2[3];
        for (int k = 0; k < 3; k++) {
            const uint32_t h = bits(t.d0[k]);
            uvs[6 * size_t(i) + 2 * k] = oracle_h2f(uint16_t(h & 0xFFFFu));
            uvs[6 * size_t(i) + 2 * k + 1] = oracle_h2f(uint16_t(h >> 16));
        }
        unit(bits(t.d1[0]), tangents + 6 * size_t(i));
        unit(bits(t.d1[1]), tangents + 6 * size_t(i) + 3);
        for (int k = 0; k < 3; k++) {
            const uint32_t c = bits(t.d2[k]);
            for (int j = 0; j < 4; j++) colors[12 * size_t(i) + 4 * k + j] = float((c >> (8 * j)) & 0xFFu) / 255.0f;
        }
    }
}

}  // extern "C"
