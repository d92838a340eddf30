// TEST INFRASTRUCTURE - the oracle.  Only tests/, __graft_entry__.smoke() and
// bench.py's cpu_baseline leg may load it, and only as the checker.
//
// Serial CPU restatement of the realtime path tracer's temporal resolve,
// data/shader/pathtracer/temporal.csh, one pixel at a time: neighbourhood
// reads clamp their coordinates to the image (what the shader's shared-memory
// tile holds), GLSL's min / max / clamp / mix as the specification defines
// them, pow(x, 16) as four squarings.  No FMA contraction.
#include <cmath>
#include <cstdint>
#include <cstring>

extern "C" float oracle_h2f(uint16_t h);

namespace {

float gmin(float x, float y) { return (y < x) ? y : x; }
float gmax(float x, float y) { return (x < y) ? y : x; }
float gclamp(float x, float lo, float hi) { return gmin(gmax(x, lo), hi); }

struct F3 {
    float x, y, z;
};

F3 to_ycocg(F3 c) {  // RGBToYCoCgMatrix * rgb (common/ycocg.hsh)
    return F3{(0.25f * c.x + 0.5f * c.y) + 0.25f * c.z, (0.5f * c.x + 0.0f * c.y) + -0.5f * c.z,
              (-0.25f * c.x + 0.5f * c.y) + -0.25f * c.z};
}

F3 to_rgb(F3 c) {  // YCoCgToRGBMatrix * ycocg
    return F3{(1.0f * c.x + 1.0f * c.y) + -1.0f * c.z, (1.0f * c.x + 0.0f * c.y) + 1.0f * c.z,
              (1.0f * c.x + -1.0f * c.y) + -1.0f * c.z};
}

F3 decode_normal(uint32_t packed) {  // DecodeNormal (common/normalencode.hsh:18-30)
    float cx = oracle_h2f(uint16_t(packed & 0xFFFFu)), cy = oracle_h2f(uint16_t(packed >> 16));
    const float y = 1.0f - (std::fabs(cx) * 1.0f + std::fabs(cy) * 1.0f);
    if (y < 0.0f) {
        const float a = cx, b = cy;
        cx = (a >= 0.0f ? 1.0f : -1.0f) * (1.0f - std::fabs(b));
        cy = (b >= 0.0f ? 1.0f : -1.0f) * (1.0f - std::fabs(a));
    }
    const float l = std::sqrt((cx * cx + y * y) + cy * cy);
    return F3{cx / l, y / l, cy / l};
}

struct In {
    uint32_t w, h;
    const float* frame;
    const float* depth;
    const uint32_t* velocity;
    const uint32_t* normal;
    const uint16_t* material;
    const float* hdepth;
    const uint32_t* hnormal;
    const uint16_t* hmaterial;
    const float* hradiance;
};

uint32_t at(const In& I, int x, int y) {
    x = x < 0 ? 0 : (x > int(I.w) - 1 ? int(I.w) - 1 : x);
    y = y < 0 ? 0 : (y > int(I.h) - 1 ? int(I.h) - 1 : y);
    return uint32_t(y) * I.w + uint32_t(x);
}

// IsHistoryPixelValid
bool valid_at(const In& I, uint32_t idx, float depth, uint32_t material, F3 n) {
    const F3 hn = decode_normal(I.hnormal[idx]);
    const float d = gmax((hn.x * n.x + hn.y * n.y) + hn.z * n.z, 0.0f);
    const float d2 = d * d, d4 = d2 * d2, d8 = d4 * d4;
    float confidence = d8 * d8;
    confidence *= uint32_t(I.hmaterial[idx]) != material ? 0.0f : 1.0f;
    const float phi = 16.0f / std::fabs(depth);
    confidence *= gmin(1.0f, std::exp(-std::fabs(depth - I.hdepth[idx]) * phi));
    return confidence > 0.1f;
}

}  // namespace

extern "C" {

// frame / hradiance / out: float4 per pixel; velocity, normal: binary16
// pairs (low half first); params: historyClipMax, currentClipFactor,
// maxHistoryLength.
void oracle_temporal(uint32_t width, uint32_t height, const float* frame, const float* depth,
                     const uint32_t* velocity, const uint32_t* normal, const uint16_t* material, const float* hdepth,
                     const uint32_t* hnormal, const uint16_t* hmaterial, const float* hradiance,
                     const float* params, float* out) {
    const In I{width, height, frame, depth, velocity, normal, material, hdepth, hnormal, hmaterial, hradiance};
    const float clip_max = params[0], ccf = params[1], max_len = params[2];
    for (uint32_t py = 0; py < height; py++)
        for (uint32_t px = 0; px < width; px++) {
            const uint32_t idx = py * width + px;
            const float depth_c = depth[idx];
            const uint32_t mat_c = material[idx];
            const F3 n_c = decode_normal(normal[idx]);
            // ComputeVarianceMinMax
            F3 m1{0, 0, 0}, m2{0, 0, 0};
            float total = 0.0f;
            for (int i = -5; i <= 5; i++)
                for (int j = -5; j <= 5; j++) {
                    const uint32_t k = at(I, int(px) + i, int(py) + j);
                    const F3 r = to_ycocg(F3{frame[4 * k], frame[4 * k + 1], frame[4 * k + 2]});
                    const float phi = gmax(1.0f, std::fabs(0.025f * depth_c));
                    const float w = gmin(1.0f, std::exp(-std::fabs(depth_c - depth[k]) / phi));
                    m1 = F3{m1.x + r.x * w, m1.y + r.y * w, m1.z + r.z * w};
                    m2 = F3{m2.x + (r.x * r.x) * w, m2.y + (r.y * r.y) * w, m2.z + (r.z * r.z) * w};
                    total += w;
                }
            const F3 mean{m1.x / total, m1.y / total, m1.z / total};
            const F3 sd{std::sqrt(gmax(m2.x / total - mean.x * mean.x, 0.0f)),
                        std::sqrt(gmax(m2.y / total - mean.y * mean.y, 0.0f)),
                        std::sqrt(gmax(m2.z / total - mean.z * mean.z, 0.0f))};
            F3 cur{frame[4 * idx], frame[4 * idx + 1], frame[4 * idx + 2]};
            const float vx = oracle_h2f(uint16_t(velocity[idx] & 0xFFFFu));
            const float vy = oracle_h2f(uint16_t(velocity[idx] >> 16));
            const float hu = (float(px) + 0.5f) * (1.0f / float(width)) + vx;
            const float hv = (float(py) + 0.5f) * (1.0f / float(height)) + vy;
            const float hpx = float(px) + vx * float(width), hpy = float(py) + vy * float(height);
            // SampleHistory
            float hist4[4] = {0, 0, 0, 0};
            bool valid = false;
            {
                const float fx = hpx - std::floor(hpx), fy = hpy - std::floor(hpy);
                const float w[4] = {(1.0f - fx) * (1.0f - fy), fx * (1.0f - fy), (1.0f - fx) * fy, fx * fy};
                const int bx = int(hpx), by = int(hpy);
                const int ox[4] = {0, 1, 0, 1}, oy[4] = {0, 0, 1, 1};
                float tw = 0.0f;
                for (int k = 0; k < 4; k++) {
                    const uint32_t h = at(I, bx + ox[k], by + oy[k]);
                    if (valid_at(I, h, depth_c, mat_c, n_c)) {
                        tw += w[k];
                        for (int a = 0; a < 4; a++) hist4[a] = hist4[a] + hradiance[4 * h + a] * w[k];
                    }
                }
                if (tw > 0.0f) {
                    for (int a = 0; a < 4; a++) hist4[a] = hist4[a] / tw;
                    valid = true;
                } else {
                    const int dx[9] = {-1, 0, 1, -1, 0, 1, -1, 0, 1}, dy[9] = {-1, -1, -1, 0, 0, 0, 1, 1, 1};
                    for (int k = 0; k < 9; k++) {
                        const uint32_t h = at(I, bx + dx[k], by + dy[k]);
                        if (valid_at(I, h, depth_c, mat_c, n_c)) {
                            tw += 1.0f;
                            for (int a = 0; a < 4; a++) hist4[a] = hist4[a] + hradiance[4 * h + a];
                        }
                    }
                    if (tw > 0.0f) {
                        for (int a = 0; a < 4; a++) hist4[a] = hist4[a] / tw;
                        valid = true;
                    } else {
                        for (int a = 0; a < 4; a++) hist4[a] = 0.0f;
                    }
                }
            }
            float hlen = hist4[3];
            F3 hist{hist4[0], hist4[1], hist4[2]};
            // SampleCatmullRom
            {
                const float pos[2] = {hu * float(width), hv * float(height)};
                float wts[4][2];
                int uvs[4][2];
                for (int a = 0; a < 2; a++) {
                    const float cen = std::floor(pos[a] - 0.5f) + 0.5f;
                    const float f1 = pos[a] - cen, f2 = f1 * f1, f3 = f2 * f1;
                    const float w0 = f2 - 0.5f * (f3 + f1);
                    const float w1 = (1.5f * f3 - 2.5f * f2) + 1.0f;
                    const float w3 = 0.5f * (f3 - f2);
                    wts[0][a] = w0;
                    wts[1][a] = w1;
                    wts[2][a] = ((1.0f - w0) - w1) - w3;
                    wts[3][a] = w3;
                    uvs[0][a] = int(cen - 1.0f);
                    uvs[1][a] = int(cen);
                    uvs[2][a] = int(cen + 1.0f);
                    uvs[3][a] = int(cen + 2.0f);
                }
                float cr[4] = {0, 0, 0, 0}, tw = 0.0f;
                for (int xi = 0; xi <= 3; xi++)
                    for (int yi = 0; yi <= 3; yi++) {
                        float w = wts[xi][0] * wts[yi][1];
                        const uint32_t h = at(I, uvs[xi][0], uvs[yi][1]);
                        w *= valid_at(I, h, depth_c, mat_c, n_c) ? 1.0f : 0.0f;
                        for (int a = 0; a < 4; a++) cr[a] = cr[a] + hradiance[4 * h + a] * w;
                        tw += w;
                    }
                if (tw > 0.5f && valid)
                    hist = F3{gmax(cr[0] / tw, 0.0f), gmax(cr[1] / tw, 0.0f), gmax(cr[2] / tw, 0.0f)};
            }
            hist = to_ycocg(hist);
            cur = to_ycocg(cur);
            const float lo[3] = {mean.x - sd.x, mean.y - sd.y, mean.z - sd.z};
            const float hi[3] = {mean.x + sd.x, mean.y + sd.y, mean.z + sd.z};
            // ClipBoundingBox
            float in[3];
            {
                const float kMinDir = 1.0f / 32767.0f;
                float d[3] = {cur.x - hist.x, cur.y - hist.y, cur.z - hist.z};
                const float o[3] = {hist.x, hist.y, hist.z};
                for (int a = 0; a < 3; a++) {
                    d[a] = std::fabs(d[a]) < kMinDir ? kMinDir : d[a];
                    const float inv = 1.0f / d[a];
                    in[a] = gmin((lo[a] - o[a]) * inv, (hi[a] - o[a]) * inv);
                }
            }
            const float clip = gmax(in[0], gmax(in[1], in[2]));
            const float adj = gclamp(clip, 0.0f, clip_max);
            cur = F3{gclamp(cur.x, mean.x - ccf * sd.x, mean.x + ccf * sd.x),
                     gclamp(cur.y, mean.y - ccf * sd.y, mean.y + ccf * sd.y),
                     gclamp(cur.z, mean.z - ccf * sd.z, mean.z + ccf * sd.z)};
            if (!valid) cur = mean;
            const F3 hrgb = to_rgb(hist), crgb = to_rgb(cur);
            const float tw = (max_len - 1.0f) / max_len;
            const float a = 1.0f - adj;
            float factor = 0.0f * (1.0f - a) + tw * a;
            factor = (hu < 0.0f || hv < 0.0f || hu > 1.0f || hv > 1.0f) ? 0.0f : factor;
            if (factor < 0.1f || !valid) hlen = 0.0f;
            factor = gmin(factor, hlen / (hlen + 1.0f));
            const float k = 1.0f - factor;
            out[4 * idx] = crgb.x * k + hrgb.x * factor;
            out[4 * idx + 1] = crgb.y * k + hrgb.y * factor;
            out[4 * idx + 2] = crgb.z * k + hrgb.z * factor;
            out[4 * idx + 3] = hlen + 1.0f;
        }
}

}  // extern "C"
