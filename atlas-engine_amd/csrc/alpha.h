// Device code (HIP sources only).  Triangle opacity as the opacity-aware
// traversals see it: the value baked per BVH slot (MeshData.cpp:141: the
// material's opacity, or -1 where the material has an opacity map) and, for
// -1, GetOpacity (raytracer/surface.hsh:147-160) sampling the material's
// opacity texture.
#pragma once

#include <hip/hip_runtime.h>

#include "pt_types.h"
#include "world.h"

namespace atlas_rt {
namespace pt {

// textureLod with the engine's global sampler (linear filter, repeat
// addressing, MainRenderer.cpp:922-929) at level 0 over UNORM8 texels, as
// the Sample*Bilinear functions of raytracer/texture.hsh read it; channels a
// texture lacks read g = b = 0, a = 1.  Callers check the index.
__device__ __forceinline__ float4 sample_texture(const AlphaArgs& A, int32_t id, float x, float y) {
    const TexDesc T = A.tex[id];
    const uint32_t ch = T.channels ? T.channels : 1u;
    const float xw = x - floorf(x), yw = y - floorf(y);
    const float fx = xw * float(T.width) - 0.5f, fy = yw * float(T.height) - 0.5f;
    const float x0 = floorf(fx), y0 = floorf(fy);
    const float a = fx - x0, b = fy - y0;
    const int32_t ix = int32_t(x0), iy = int32_t(y0);
    const uint32_t i0 = ix < 0 ? T.width - 1 : min(uint32_t(ix), T.width - 1);
    const uint32_t j0 = iy < 0 ? T.height - 1 : min(uint32_t(iy), T.height - 1);
    const uint32_t i1 = i0 + 1 >= T.width ? 0u : i0 + 1, j1 = j0 + 1 >= T.height ? 0u : j0 + 1;
    const uint8_t* px = A.texels + T.offset;
    float out[4] = {0.0f, 0.0f, 0.0f, 1.0f};
    for (uint32_t c = 0; c < ch; c++) {
        const float t00 = float(px[(size_t(j0) * T.width + i0) * ch + c]) / 255.0f;
        const float t10 = float(px[(size_t(j0) * T.width + i1) * ch + c]) / 255.0f;
        const float t01 = float(px[(size_t(j1) * T.width + i0) * ch + c]) / 255.0f;
        const float t11 = float(px[(size_t(j1) * T.width + i1) * ch + c]) / 255.0f;
        const float r0 = t00 * (1.0f - a) + t10 * a;
        const float r1 = t01 * (1.0f - a) + t11 * a;
        out[c] = r0 * (1.0f - b) + r1 * b;
    }
    return make_float4(out[0], out[1], out[2], out[3]);
}

__device__ __forceinline__ bool has_texture(const AlphaArgs& A, int32_t id) {
    return id >= 0 && uint32_t(id) < A.tex_count;
}

// SampleOpacityBilinear (raytracer/texture.hsh:18-25): .r, 1 without a map.
__device__ __forceinline__ float sample_opacity(const AlphaArgs& A, int32_t id, float x, float y) {
    return has_texture(A, id) ? sample_texture(A, id, x, y).x : 1.0f;
}

// r * uv0 + s * uv1 + t * uv2 with r = 1 - s - t (surface.hsh:81-87).
__device__ __forceinline__ float2 uv_at(const float4* uv, int32_t slot, float s, float t) {
    if (!uv) return make_float2(0.0f, 0.0f);
    const float4 a = uv[2 * size_t(slot)], b = uv[2 * size_t(slot) + 1];
    const float r = (1.0f - s) - t;
    return make_float2((r * a.x + s * a.z) + t * b.x, (r * a.y + s * a.w) + t * b.y);
}

// Opacity of (instance slot, BVH slot) at the object-space barycentrics.
struct TriOpacity {
    WorldDev W;
    AlphaArgs A;
    __device__ __forceinline__ float operator()(int32_t inst, int32_t slot, float u, float v) const {
        const float4 meta = W.instances[4 * size_t(inst) + 3];
        const int32_t mesh = __float_as_int(meta.x);
        const float4* sp = A.shade[mesh] + 3 * size_t(slot);
        const float op = sp[1].w;
        if (op >= 0.0f || !A.materials) return op >= 0.0f ? op : 1.0f;
        const Material& m = A.materials[__float_as_int(sp[0].w) + __float_as_int(meta.y)];
        float2 tc = uv_at(A.uv ? A.uv[mesh] : nullptr, slot, u, v);
        if (m.invert_uvs > 0) tc.y = 1.0f - tc.y;
        return sample_opacity(A, m.opacity_texture, tc.x, tc.y) * m.opacity;
    }
};

// HitClosestTransparency's leaf filter (bvh.hsh:127-133): opacity > 0.
struct AcceptOpaque {
    TriOpacity op;
    __device__ __forceinline__ bool operator()(int32_t inst, int32_t slot, float u, float v) const {
        return op(inst, slot, u, v) > 0.0f;
    }
};

}  // namespace pt
}  // namespace atlas_rt
