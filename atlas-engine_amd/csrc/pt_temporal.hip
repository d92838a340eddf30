// Temporal resolve of the realtime path tracer, gfx950
// (data/shader/pathtracer/temporal.csh).
//
// One 16 x 16 block per tile; the tile plus a 5-pixel border of this frame's
// radiance and depth is staged in LDS (26 x 26 float4, border texels clamped
// to the image), as the shader's LoadGroupSharedData does.  Per pixel:
//   - depth-weighted mean and standard deviation of the radiance in YCoCg
//     over the 11 x 11 neighbourhood (ComputeVarianceMinMax);
//   - the history at the reprojected position: bilinear over the 2 x 2
//     pixels whose G-buffer matches (normal, material, depth), else the mean
//     of the matching ones of the 3 x 3 neighbourhood (SampleHistory); its
//     alpha is the history length;
//   - a 4 x 4 Catmull-Rom history where enough of it is valid
//     (SampleCatmullRom);
//   - clip of the history against mean +- std, clamp of the current value to
//     mean +- currentClipFactor * std, blend weight from the history length.
// GLSL's min / max / clamp / mix are written out as the specification
// defines them (min(x, y) = y < x ? y : x, max(x, y) = x < y ? y : x,
// mix(x, y, a) = x * (1 - a) + y * a); pow(x, 16) is four squarings.
// The shader's FindNearest3x3 result is unused there and is not computed.
#include <hip/hip_fp16.h>

#include "pathtracer.h"

#pragma clang fp contract(off)

namespace atlas_rt {
namespace pt {

namespace {

constexpr int kTile = 16;
constexpr int kRadius = 5;
constexpr int kSide = kTile + 2 * kRadius;  // 26

__device__ __forceinline__ float gmin(float x, float y) { return (y < x) ? y : x; }
__device__ __forceinline__ float gmax(float x, float y) { return (x < y) ? y : x; }
__device__ __forceinline__ float gclamp(float x, float lo, float hi) { return gmin(gmax(x, lo), hi); }
__device__ __forceinline__ int iclamp(int x, int lo, int hi) { return x < lo ? lo : (x > hi ? hi : x); }

struct F3 {
    float x, y, z;
};

__device__ __forceinline__ F3 f3(float x, float y, float z) { return F3{x, y, z}; }

// RGBToYCoCgMatrix * rgb (common/ycocg.hsh): columns (0.25, 0.5, -0.25),
// (0.5, 0, 0.5), (0.25, -0.5, -0.25).
__device__ __forceinline__ F3 to_ycocg(F3 c) {
    return f3((0.25f * c.x + 0.5f * c.y) + 0.25f * c.z, (0.5f * c.x + 0.0f * c.y) + -0.5f * c.z,
              (-0.25f * c.x + 0.5f * c.y) + -0.25f * c.z);
}

// YCoCgToRGBMatrix * ycocg: columns (1, 1, 1), (1, 0, -1), (-1, 1, -1).
__device__ __forceinline__ F3 to_rgb(F3 c) {
    return f3((1.0f * c.x + 1.0f * c.y) + -1.0f * c.z, (1.0f * c.x + 0.0f * c.y) + 1.0f * c.z,
              (1.0f * c.x + -1.0f * c.y) + -1.0f * c.z);
}

__device__ __forceinline__ float half_lo(uint32_t v) { return __half2float(__ushort_as_half(uint16_t(v & 0xFFFFu))); }
__device__ __forceinline__ float half_hi(uint32_t v) { return __half2float(__ushort_as_half(uint16_t(v >> 16))); }

// DecodeNormal (common/normalencode.hsh:18-30).
__device__ __forceinline__ F3 decode_normal(uint32_t packed) {
    float cx = half_lo(packed), cy = half_hi(packed);
    const float y = 1.0f - (fabsf(cx) * 1.0f + fabsf(cy) * 1.0f);
    if (y < 0.0f) {
        const float a = cx, b = cy;
        cx = (a >= 0.0f ? 1.0f : -1.0f) * (1.0f - fabsf(b));
        cy = (b >= 0.0f ? 1.0f : -1.0f) * (1.0f - fabsf(a));
    }
    const float l = sqrtf((cx * cx + y * y) + cy * cy);
    return f3(cx / l, y / l, cy / l);
}

struct Pixel {
    float depth;
    uint32_t material;
    F3 normal;
};

// IsHistoryPixelValid: normal agreement ^ 16, same material, depth agreement.
__device__ __forceinline__ bool history_valid(const TemporalArgs& T, uint32_t idx, const Pixel& c) {
    const F3 hn = decode_normal(T.hnormal[idx]);
    const float d = gmax((hn.x * c.normal.x + hn.y * c.normal.y) + hn.z * c.normal.z, 0.0f);
    const float d2 = d * d, d4 = d2 * d2, d8 = d4 * d4;
    float confidence = d8 * d8;
    confidence *= uint32_t(T.hmaterial[idx]) != c.material ? 0.0f : 1.0f;
    const float phi = 16.0f / fabsf(c.depth);
    confidence *= gmin(1.0f, expf(-fabsf(c.depth - T.hdepth[idx]) * phi));
    return confidence > 0.1f;
}

__device__ __forceinline__ uint32_t at(const TemporalArgs& T, int x, int y) {
    return uint32_t(iclamp(y, 0, int(T.height) - 1)) * T.width + uint32_t(iclamp(x, 0, int(T.width) - 1));
}

__device__ __forceinline__ float4 add4(float4 a, float4 b) { return make_float4(a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w); }
__device__ __forceinline__ float4 mul4(float4 a, float s) { return make_float4(a.x * s, a.y * s, a.z * s, a.w * s); }
__device__ __forceinline__ float4 div4(float4 a, float s) { return make_float4(a.x / s, a.y / s, a.z / s, a.w / s); }

__global__ __launch_bounds__(kTile * kTile) void k_temporal(TemporalArgs T) {
    __shared__ float4 tile[kSide * kSide];
    const int x0 = int(blockIdx.x) * kTile - kRadius, y0 = int(blockIdx.y) * kTile - kRadius;
    for (int i = int(threadIdx.x); i < kSide * kSide; i += kTile * kTile) {
        const uint32_t idx = at(T, x0 + i % kSide, y0 + i / kSide);
        // each tap's YCoCg once per tile entry instead of once per pixel that reads it
        const float4 c = T.frame[idx];
        const F3 r = to_ycocg(f3(c.x, c.y, c.z));
        tile[i] = make_float4(r.x, r.y, r.z, T.depth[idx]);
    }
    __syncthreads();
    const int lx = int(threadIdx.x) % kTile, ly = int(threadIdx.x) / kTile;
    const int px = int(blockIdx.x) * kTile + lx, py = int(blockIdx.y) * kTile + ly;
    if (px >= int(T.width) || py >= int(T.height)) return;
    const uint32_t idx = uint32_t(py) * T.width + uint32_t(px);
    Pixel c;
    c.depth = T.depth[idx];
    c.material = T.material[idx];
    c.normal = decode_normal(T.normal[idx]);

    // ComputeVarianceMinMax: i over x, j over y
    F3 m1 = f3(0.0f, 0.0f, 0.0f), m2 = f3(0.0f, 0.0f, 0.0f);
    float total = 0.0f;
    // the per-tap division by phi as a product with its reciprocal, and exp
    // on the hardware exp2 (GLSL allows 2.5 ulp for a division and
    // 3 + 2 |x| ulp for exp)
    const float inv_phi = 1.0f / gmax(1.0f, fabsf(0.025f * c.depth));
    for (int i = -kRadius; i <= kRadius; i++)
        for (int j = -kRadius; j <= kRadius; j++) {
            const float4 t = tile[(ly + kRadius + j) * kSide + (lx + kRadius + i)];
            const F3 r = f3(t.x, t.y, t.z);
            const float w = gmin(1.0f, __expf(-fabsf(c.depth - t.w) * inv_phi));
            m1 = f3(m1.x + r.x * w, m1.y + r.y * w, m1.z + r.z * w);
            m2 = f3(m2.x + (r.x * r.x) * w, m2.y + (r.y * r.y) * w, m2.z + (r.z * r.z) * w);
            total += w;
        }
    const F3 mean = f3(m1.x / total, m1.y / total, m1.z / total);
    const F3 sd = f3(sqrtf(gmax(m2.x / total - mean.x * mean.x, 0.0f)), sqrtf(gmax(m2.y / total - mean.y * mean.y, 0.0f)),
                     sqrtf(gmax(m2.z / total - mean.z * mean.z, 0.0f)));

    const float4 cur4 = T.frame[idx];
    F3 cur = f3(cur4.x, cur4.y, cur4.z);
    const uint32_t vel = T.velocity[idx];
    const float vx = half_lo(vel), vy = half_hi(vel);
    const float inv_w = 1.0f / float(T.width), inv_h = 1.0f / float(T.height);
    const float hu = (float(px) + 0.5f) * inv_w + vx, hv = (float(py) + 0.5f) * inv_h + vy;
    const float hpx = float(px) + vx * float(T.width), hpy = float(py) + vy * float(T.height);

    // SampleHistory
    float4 history = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    bool valid = false;
    {
        const float fx = hpx - floorf(hpx), fy = hpy - floorf(hpy);
        const float w[4] = {(1.0f - fx) * (1.0f - fy), fx * (1.0f - fy), (1.0f - fx) * fy, fx * fy};
        const int bx = int(hpx), by = int(hpy);
        float tw = 0.0f;
        for (int k = 0; k < 4; k++) {
            const uint32_t h = at(T, bx + (k & 1), by + (k >> 1));
            if (history_valid(T, h, c)) {
                tw += w[k];
                history = add4(history, mul4(T.hradiance[h], w[k]));
            }
        }
        if (tw > 0.0f) {
            history = div4(history, tw);
            valid = true;
        } else {
            for (int k = 0; k < 9; k++) {
                const uint32_t h = at(T, bx + k % 3 - 1, by + k / 3 - 1);
                if (history_valid(T, h, c)) {
                    tw += 1.0f;
                    history = add4(history, T.hradiance[h]);
                }
            }
            if (tw > 0.0f) {
                history = div4(history, tw);
                valid = true;
            } else {
                history = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            }
        }
    }
    float history_length = history.w;
    F3 hist = f3(history.x, history.y, history.z);

    // SampleCatmullRom (Jimenez, "Filmic SMAA", SIGGRAPH 2016)
    {
        const float posx = hu * float(T.width), posy = hv * float(T.height);
        const float cenx = floorf(posx - 0.5f) + 0.5f, ceny = floorf(posy - 0.5f) + 0.5f;
        const float f[2] = {posx - cenx, posy - ceny};
        float wts[4][2];
        int uvs[4][2];
        for (int a = 0; a < 2; a++) {
            const float f1 = f[a], f2 = f1 * f1, f3v = f2 * f1;
            const float w0 = f2 - 0.5f * (f3v + f1);
            const float w1 = (1.5f * f3v - 2.5f * f2) + 1.0f;
            const float w3 = 0.5f * (f3v - f2);
            const float w2 = ((1.0f - w0) - w1) - w3;
            wts[0][a] = w0;
            wts[1][a] = w1;
            wts[2][a] = w2;
            wts[3][a] = w3;
            const float cen = a == 0 ? cenx : ceny;
            uvs[0][a] = int(cen - 1.0f);
            uvs[1][a] = int(cen);
            uvs[2][a] = int(cen + 1.0f);
            uvs[3][a] = int(cen + 2.0f);
        }
        float4 cr = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        float tw = 0.0f;
        for (int xi = 0; xi <= 3; xi++)
            for (int yi = 0; yi <= 3; yi++) {
                float w = wts[xi][0] * wts[yi][1];
                const uint32_t h = at(T, uvs[xi][0], uvs[yi][1]);
                w *= history_valid(T, h, c) ? 1.0f : 0.0f;
                cr = add4(cr, mul4(T.hradiance[h], w));
                tw += w;
            }
        if (tw > 0.5f && valid) {
            cr = div4(cr, tw);
            hist = f3(gmax(cr.x, 0.0f), gmax(cr.y, 0.0f), gmax(cr.z, 0.0f));
        }
    }

    hist = to_ycocg(hist);
    cur = to_ycocg(cur);
    const F3 hmin = f3(mean.x - sd.x, mean.y - sd.y, mean.z - sd.z);
    const F3 hmax = f3(mean.x + sd.x, mean.y + sd.y, mean.z + sd.z);
    const float ccf = T.current_clip_factor;
    const F3 cmin = f3(mean.x - ccf * sd.x, mean.y - ccf * sd.y, mean.z - ccf * sd.z);
    const F3 cmax = f3(mean.x + ccf * sd.x, mean.y + ccf * sd.y, mean.z + ccf * sd.z);
    // ClipBoundingBox
    float clip;
    {
        const float kMinDir = 1.0f / 32767.0f;
        float d[3] = {cur.x - hist.x, cur.y - hist.y, cur.z - hist.z};
        for (int a = 0; a < 3; a++) d[a] = fabsf(d[a]) < kMinDir ? kMinDir : d[a];
        const float o[3] = {hist.x, hist.y, hist.z};
        const float lo[3] = {hmin.x, hmin.y, hmin.z}, hi[3] = {hmax.x, hmax.y, hmax.z};
        float in[3];
        for (int a = 0; a < 3; a++) {
            const float inv = 1.0f / d[a];
            in[a] = gmin((lo[a] - o[a]) * inv, (hi[a] - o[a]) * inv);
        }
        clip = gmax(in[0], gmax(in[1], in[2]));
    }
    const float adj = gclamp(clip, 0.0f, T.history_clip_max);
    cur = f3(gclamp(cur.x, cmin.x, cmax.x), gclamp(cur.y, cmin.y, cmax.y), gclamp(cur.z, cmin.z, cmax.z));
    if (!valid) cur = mean;
    const F3 hrgb = to_rgb(hist), crgb = to_rgb(cur);
    const float tw = (T.max_history_length - 1.0f) / T.max_history_length;
    const float a = 1.0f - adj;
    float factor = 0.0f * (1.0f - a) + tw * a;
    factor = (hu < 0.0f || hv < 0.0f || hu > 1.0f || hv > 1.0f) ? 0.0f : factor;
    if (factor < 0.1f || !valid) history_length = 0.0f;
    factor = gmin(factor, history_length / (history_length + 1.0f));
    const float k = 1.0f - factor;
    T.out[idx] = make_float4(crgb.x * k + hrgb.x * factor, crgb.y * k + hrgb.y * factor, crgb.z * k + hrgb.z * factor,
                             history_length + 1.0f);
}

}  // namespace

atlas_rt_status launch_temporal(const TemporalArgs& T, hipStream_t s) {
    const dim3 grid((T.width + kTile - 1) / kTile, (T.height + kTile - 1) / kTile);
    hipLaunchKernelGGL(k_temporal, grid, dim3(kTile * kTile), 0, s, T);
    ATLAS_HIP_TRY(hipGetLastError());
    return ATLAS_RT_OK;
}

}  // namespace pt
}  // namespace atlas_rt
