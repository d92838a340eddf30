// Radiance-cache kernels for gfx950: fused hash-grid encoding + MLP
// inference, fused forward/backward training with per-block weight-gradient
// sums in LDS, and the Adam updates.  Layouts and the lane/unit mapping are
// described in nrc.h.
#include <hip/hip_runtime.h>

#include "nrc.h"

namespace atlas_rt {
namespace nrc {

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

constexpr float kBeta1 = 0.9f;
constexpr float kBeta2 = 0.99f;
constexpr float kEps = 1e-15f;
constexpr float kLossFloor = 0.01f;
constexpr int kStride = 40;              // bf16 per staged row: 32 queries + 8 pad (80 B)
constexpr int kStage = 64 * kStride;     // one staged 64-unit x 32-query matrix

__device__ __forceinline__ f32x4 mfma(u32x4 a, u32x4 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a), __builtin_bit_cast(bf16x8, b), c, 0,
                                                   0, 0);
}

// Round to nearest even (finite inputs).
__device__ __forceinline__ uint32_t bf16_bits(float f) {
    uint32_t u = __float_as_uint(f);
    return (u + 0x7fffu + ((u >> 16) & 1u)) >> 16;
}
__device__ __forceinline__ uint32_t pack_bf16(float lo, float hi) { return bf16_bits(lo) | (bf16_bits(hi) << 16); }
__device__ __forceinline__ float bf16_lo(uint32_t w) { return __uint_as_float(w << 16); }
__device__ __forceinline__ float bf16_hi(uint32_t w) { return __uint_as_float(w & 0xffff0000u); }
__device__ __forceinline__ u32x4 pack_frag(f32x4 a, f32x4 b) {
    return u32x4{pack_bf16(a[0], a[1]), pack_bf16(a[2], a[3]), pack_bf16(b[0], b[1]), pack_bf16(b[2], b[3])};
}
__device__ __forceinline__ f32x4 relu4(f32x4 a) {
    return f32x4{fmaxf(a[0], 0.0f), fmaxf(a[1], 0.0f), fmaxf(a[2], 0.0f), fmaxf(a[3], 0.0f)};
}
__device__ __forceinline__ float half_to_float(uint32_t bits16) {
    return (float)__builtin_bit_cast(_Float16, (uint16_t)bits16);
}
__device__ __forceinline__ uint16_t float_to_half(float f) { return __builtin_bit_cast(uint16_t, (_Float16)f); }

__device__ __forceinline__ u32x4 load_frag(const uint16_t* base, int frag, int lane) {
    return *reinterpret_cast<const u32x4*>(base + ((size_t)frag * 64 + lane) * 8);
}

// ---------------------------------------------------------------------------
// Hash-grid encoding
// ---------------------------------------------------------------------------

// The 8 corners of the cell around normalised position (x, y, z) at `lev`:
// table entry index (level-major) and trilinear weight.
__device__ __forceinline__ void level_corners(const NrcDev& P, int lev, float x, float y, float z, uint32_t (&idx)[8],
                                              float (&w)[8]) {
    const float R = P.res[lev];
    const float sx = x * R, sy = y * R, sz = z * R;
    const float cx = fminf(floorf(sx), R - 1.0f), cy = fminf(floorf(sy), R - 1.0f), cz = fminf(floorf(sz), R - 1.0f);
    const float fx = sx - cx, fy = sy - cy, fz = sz - cz;
    const uint32_t ix = (uint32_t)cx, iy = (uint32_t)cy, iz = (uint32_t)cz;
    const uint32_t T = 1u << P.log2_t;
    const bool dense = (P.dense_mask >> lev) & 1u;
    const uint32_t R1 = (uint32_t)R + 1u;
#pragma unroll
    for (int c = 0; c < 8; c++) {
        const uint32_t px = ix + (c & 1), py = iy + ((c >> 1) & 1), pz = iz + (c >> 2);
        const uint32_t h = dense ? px + R1 * (py + R1 * pz) : ((px ^ (py * 2654435761u) ^ (pz * 805459861u)) & (T - 1u));
        idx[c] = (uint32_t)lev * T + h;
        const float wx = (c & 1) ? fx : 1.0f - fx;
        const float wy = (c & 2) ? fy : 1.0f - fy;
        const float wz = (c & 4) ? fz : 1.0f - fz;
        w[c] = (wx * wy) * wz;
    }
}

__device__ __forceinline__ float clamp01(float v) { return fminf(fmaxf(v, 0.0f), 1.0f); }

// Level held in feature slot j (features 2j, 2j+1 of the lane's 8) by lane
// group g, or -1 when the slot holds direction/normal/roughness inputs.
__device__ __forceinline__ int slot_level(int g, int j) {
    if (j < 2) return 2 * g + j;
    return g < 2 ? 8 + 2 * g + (j - 2) : -1;
}

// The 8 input features lane group g feeds to K slots i = 0..7 for query qi:
// features 4g + 0..3 and 16 + 4g + 0..3 of the 32-input vector.
__device__ __forceinline__ void lane_features(const NrcDev& P, const float4* __restrict__ q, uint32_t qi, bool valid,
                                              int g, float (&f)[8]) {
#pragma unroll
    for (int k = 0; k < 8; k++) f[k] = 0.0f;
    if (!valid) return;
    const float4 a = q[3 * (size_t)qi], b = q[3 * (size_t)qi + 1], c = q[3 * (size_t)qi + 2];
    const float x = clamp01((a.x - P.bmin[0]) * P.binv[0]);
    const float y = clamp01((a.y - P.bmin[1]) * P.binv[1]);
    const float z = clamp01((a.z - P.bmin[2]) * P.binv[2]);
    const uint32_t* tab = reinterpret_cast<const uint32_t*>(P.table);
#pragma unroll
    for (int j = 0; j < 4; j++) {
        const int lev = slot_level(g, j);
        if (lev < 0 || lev >= (int)P.levels) continue;
        uint32_t idx[8];
        float w[8];
        level_corners(P, lev, x, y, z, idx, w);
        float f0 = 0.0f, f1 = 0.0f;
#pragma unroll
        for (int cc = 0; cc < 8; cc++) {
            const uint32_t e = tab[idx[cc]];
            f0 += w[cc] * half_to_float(e & 0xffffu);
            f1 += w[cc] * half_to_float(e >> 16);
        }
        f[2 * j] = f0;
        f[2 * j + 1] = f1;
    }
    if (g == 2) {
        f[4] = b.x;
        f[5] = b.y;
        f[6] = b.z;
        f[7] = c.x;
    } else if (g == 3) {
        f[4] = c.y;
        f[5] = c.z;
        f[6] = a.w;
        f[7] = 1.0f;
    }
}

// Scatter d(loss)/d(feature) of the lane's hash features into the FP32
// table gradient.
__device__ __forceinline__ void lane_table_grad(const NrcDev& P, const float4* __restrict__ q, uint32_t qi, int g,
                                                const float (&df)[8], float* __restrict__ grad) {
    const float4 a = q[3 * (size_t)qi];
    const float x = clamp01((a.x - P.bmin[0]) * P.binv[0]);
    const float y = clamp01((a.y - P.bmin[1]) * P.binv[1]);
    const float z = clamp01((a.z - P.bmin[2]) * P.binv[2]);
#pragma unroll
    for (int j = 0; j < 4; j++) {
        const int lev = slot_level(g, j);
        if (lev < 0 || lev >= (int)P.levels) continue;
        const float d0 = df[2 * j], d1 = df[2 * j + 1];
        if (d0 == 0.0f && d1 == 0.0f) continue;
        uint32_t idx[8];
        float w[8];
        level_corners(P, lev, x, y, z, idx, w);
#pragma unroll
        for (int cc = 0; cc < 8; cc++) {
            atomicAdd(grad + 2 * (size_t)idx[cc], w[cc] * d0);
            atomicAdd(grad + 2 * (size_t)idx[cc] + 1, w[cc] * d1);
        }
    }
}

// ---------------------------------------------------------------------------
// MLP building blocks (all lanes of the wave active)
// ---------------------------------------------------------------------------

// z[t][q] = W_L^T x for output tiles t, query tiles q.
template <int L, int QT>
__device__ __forceinline__ void fwd_layer(const uint16_t* __restrict__ wn, const u32x4 (&x)[kK[L] / 32][QT],
                                          f32x4 (&z)[kNp[L] / 16][QT], int lane) {
    constexpr int S = kK[L] / 32, T = kNp[L] / 16;
    const uint16_t* w = wn + wn_offset(L);
#pragma unroll
    for (int t = 0; t < T; t++)
#pragma unroll
        for (int q = 0; q < QT; q++) z[t][q] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int s = 0; s < S; s++)
#pragma unroll
        for (int t = 0; t < T; t++) {
            const u32x4 a = load_frag(w, t * S + s, lane);
#pragma unroll
            for (int q = 0; q < QT; q++) z[t][q] = mfma(a, x[s][q], z[t][q]);
        }
}

// dx[t][q] = W_L dz for input tiles t.
template <int L, int QT>
__device__ __forceinline__ void bwd_layer(const uint16_t* __restrict__ wk, const u32x4 (&dz)[nk_cols(L) / 32][QT],
                                          f32x4 (&dx)[kK[L] / 16][QT], int lane) {
    constexpr int S = nk_cols(L) / 32, T = kK[L] / 16;
    const uint16_t* w = wk + wk_offset(L);
#pragma unroll
    for (int t = 0; t < T; t++)
#pragma unroll
        for (int q = 0; q < QT; q++) dx[t][q] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int s = 0; s < S; s++)
#pragma unroll
        for (int t = 0; t < T; t++) {
            const u32x4 a = load_frag(w, t * S + s, lane);
#pragma unroll
            for (int q = 0; q < QT; q++) dx[t][q] = mfma(a, dz[s][q], dx[t][q]);
        }
}

template <int QT>
__device__ __forceinline__ void relu_pack(const f32x4 (&z)[4][QT], u32x4 (&h)[2][QT]) {
#pragma unroll
    for (int s = 0; s < 2; s++)
#pragma unroll
        for (int q = 0; q < QT; q++) h[s][q] = pack_frag(relu4(z[2 * s][q]), relu4(z[2 * s + 1][q]));
}

__device__ __forceinline__ float gate(uint32_t bits_word, bool hi, float v) {
    const float h = hi ? bf16_hi(bits_word) : bf16_lo(bits_word);
    return h > 0.0f ? v : 0.0f;
}

// dz = dx * relu'(z) with the gate read from the packed forward activations.
template <int QT>
__device__ __forceinline__ void gate_pack(const f32x4 (&dx)[4][QT], const u32x4 (&h)[2][QT], u32x4 (&dz)[2][QT]) {
#pragma unroll
    for (int s = 0; s < 2; s++)
#pragma unroll
        for (int q = 0; q < QT; q++) {
            const u32x4 hv = h[s][q];
            const f32x4 a = dx[2 * s][q], b = dx[2 * s + 1][q];
            const f32x4 ga{gate(hv[0], false, a[0]), gate(hv[0], true, a[1]), gate(hv[1], false, a[2]),
                           gate(hv[1], true, a[3])};
            const f32x4 gb{gate(hv[2], false, b[0]), gate(hv[2], true, b[1]), gate(hv[3], false, b[2]),
                           gate(hv[3], true, b[3])};
            dz[s][q] = pack_frag(ga, gb);
        }
}

__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Transpose B-operand fragments (units P(s,g,i) x query 16 q + r) into a
// [unit][query] bf16 matrix in LDS.
template <int S, int QT>
__device__ __forceinline__ void stage(uint16_t* st, const u32x4 (&x)[S][QT], int g, int r) {
#pragma unroll
    for (int s = 0; s < S; s++)
#pragma unroll
        for (int q = 0; q < QT; q++)
#pragma unroll
            for (int i = 0; i < 8; i++) {
                const uint32_t w = x[s][q][i >> 1];
                const uint16_t v = (i & 1) ? (uint16_t)(w >> 16) : (uint16_t)(w & 0xffffu);
                const int n = 32 * s + 16 * (i >> 2) + 4 * g + (i & 3);
                st[n * kStride + 16 * q + r] = v;
            }
}

// dW_L += X_L^T dZ_L over the wave's 32 queries (K of the product), added
// into the block's LDS gradient.
template <int L>
__device__ __forceinline__ void dw_layer(const uint16_t* xs, const uint16_t* ds, float* dW, int g, int r) {
    constexpr int MT = kK[L] / 16, NT = kNp[L] / 16, N = kN[L];
    float* dst = dW + master_offset(L);
    u32x4 b[NT];
#pragma unroll
    for (int to = 0; to < NT; to++) b[to] = *reinterpret_cast<const u32x4*>(ds + (16 * to + r) * kStride + 8 * g);
#pragma unroll
    for (int ti = 0; ti < MT; ti++) {
        const u32x4 a = *reinterpret_cast<const u32x4*>(xs + (16 * ti + r) * kStride + 8 * g);
#pragma unroll
        for (int to = 0; to < NT; to++) {
            const f32x4 d = mfma(a, b[to], f32x4{0.0f, 0.0f, 0.0f, 0.0f});
            const int out = 16 * to + r;
            if (out < N) {
#pragma unroll
                for (int i = 0; i < 4; i++) atomicAdd(&dst[(16 * ti + 4 * g + i) * N + out], d[i]);
            }
        }
    }
}

template <int L, int S, int QT>
__device__ __forceinline__ void dw_step(uint16_t* xs, uint16_t* ds, const u32x4 (&x)[kK[L] / 32][QT],
                                        const u32x4 (&dz)[S][QT], float* dW, int g, int r) {
    stage<kK[L] / 32, QT>(xs, x, g, r);
    stage<S, QT>(ds, dz, g, r);
    wave_sync();
    dw_layer<L>(xs, ds, dW, g, r);
    wave_sync();
}

// ---------------------------------------------------------------------------
// Kernels
// ---------------------------------------------------------------------------

constexpr int kInferQT = kInferQueriesPerWave / 16;
constexpr int kTrainQT = kTrainQueriesPerWave / 16;
static_assert(kTrainQT == 2, "the weight-gradient product takes K = 32 queries per chunk");

// count = *d_count when d_count is set (device-resident counts), else n.
__global__ __launch_bounds__(256) void k_nrc_infer(NrcDev P, const float4* __restrict__ q, float4* __restrict__ out,
                                                   const uint32_t* __restrict__ d_count, uint32_t n) {
    constexpr int QT = kInferQT;
    const int lane = threadIdx.x & 63, g = lane >> 4, r = lane & 15;
    const uint32_t count = d_count ? min(*d_count, n) : n;
    const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const uint32_t base = wave * (16 * QT);
    if (base >= count) return;  // wave-uniform
    u32x4 x0[1][QT];
#pragma unroll
    for (int t = 0; t < QT; t++) {
        const uint32_t qi = base + 16 * t + r;
        float f[8];
        lane_features(P, q, qi, qi < count, g, f);
        x0[0][t] = u32x4{pack_bf16(f[0], f[1]), pack_bf16(f[2], f[3]), pack_bf16(f[4], f[5]), pack_bf16(f[6], f[7])};
    }
    f32x4 z[4][QT];
    u32x4 h[2][QT];
    fwd_layer<0, QT>(P.wn, x0, z, lane);
    relu_pack<QT>(z, h);
    fwd_layer<1, QT>(P.wn, h, z, lane);
    relu_pack<QT>(z, h);
    fwd_layer<2, QT>(P.wn, h, z, lane);
    relu_pack<QT>(z, h);
    fwd_layer<3, QT>(P.wn, h, z, lane);
    relu_pack<QT>(z, h);
    fwd_layer<4, QT>(P.wn, h, z, lane);
    relu_pack<QT>(z, h);
    f32x4 y[1][QT];
    fwd_layer<5, QT>(P.wn, h, y, lane);
    if (g == 0) {
#pragma unroll
        for (int t = 0; t < QT; t++) {
            const uint32_t qi = base + 16 * t + r;
            if (qi < count) out[qi] = make_float4(y[0][t][0], y[0][t][1], y[0][t][2], 0.0f);
        }
    }
}

// Trains on slice `slice` of `slices` equal slices of the samples; the
// sample count is *d_count when d_count is set (at most n), else n.
__global__ __launch_bounds__(kTrainBlock, 1) void k_nrc_train(NrcDev P, const float4* __restrict__ q,
                                                               const float4* __restrict__ tgt,
                                                               const uint32_t* __restrict__ d_count, uint32_t n,
                                                               uint32_t slice, uint32_t slices,
                                                               float* __restrict__ partial,
                                                               float* __restrict__ grad_table, float* loss) {
    constexpr int QT = kTrainQT;
    constexpr int kWaves = kTrainBlock / 64;
    const uint32_t total = d_count ? min(*d_count, n) : n;
    const uint32_t first = uint32_t(uint64_t(total) * slice / slices);
    const uint32_t count = uint32_t(uint64_t(total) * (slice + 1) / slices) - first;
    const float inv_norm = count ? 1.0f / (3.0f * float(count)) : 0.0f;
    q += 3 * size_t(first);
    tgt += first;
    __shared__ float dW[kMasterSize];
    __shared__ __align__(16) uint16_t staging[kWaves][2][kStage];
    for (int j = threadIdx.x; j < kMasterSize; j += kTrainBlock) dW[j] = 0.0f;
    __syncthreads();

    const int lane = threadIdx.x & 63, g = lane >> 4, r = lane & 15, wv = threadIdx.x >> 6;
    uint16_t* xs = staging[wv][0];
    uint16_t* ds = staging[wv][1];
    const uint32_t chunks = (count + kTrainQueriesPerWave - 1) / kTrainQueriesPerWave;
    float lane_loss = 0.0f;

    for (uint32_t c = blockIdx.x * kWaves + wv; c < chunks; c += gridDim.x * kWaves) {  // wave-uniform
        const uint32_t base = c * kTrainQueriesPerWave;
        // forward, keeping every layer's input as packed bf16
        u32x4 x0[1][QT], h0[2][QT], h1[2][QT], h2[2][QT], h3[2][QT], h4[2][QT];
#pragma unroll
        for (int t = 0; t < QT; t++) {
            const uint32_t qi = base + 16 * t + r;
            float f[8];
            lane_features(P, q, qi, qi < count, g, f);
            x0[0][t] =
                u32x4{pack_bf16(f[0], f[1]), pack_bf16(f[2], f[3]), pack_bf16(f[4], f[5]), pack_bf16(f[6], f[7])};
        }
        f32x4 z[4][QT];
        fwd_layer<0, QT>(P.wn, x0, z, lane);
        relu_pack<QT>(z, h0);
        fwd_layer<1, QT>(P.wn, h0, z, lane);
        relu_pack<QT>(z, h1);
        fwd_layer<2, QT>(P.wn, h1, z, lane);
        relu_pack<QT>(z, h2);
        fwd_layer<3, QT>(P.wn, h2, z, lane);
        relu_pack<QT>(z, h3);
        fwd_layer<4, QT>(P.wn, h3, z, lane);
        relu_pack<QT>(z, h4);
        f32x4 y[1][QT];
        fwd_layer<5, QT>(P.wn, h4, y, lane);

        // relative L2 loss and its gradient (outputs live in lanes g == 0)
        u32x4 dz5[1][QT];
#pragma unroll
        for (int t = 0; t < QT; t++) {
            const uint32_t qi = base + 16 * t + r;
            f32x4 d{0.0f, 0.0f, 0.0f, 0.0f};
            if (g == 0 && qi < count) {
                const float4 tv = tgt[qi];
                const float tt[3] = {tv.x, tv.y, tv.z};
#pragma unroll
                for (int i = 0; i < 3; i++) {
                    const float yv = y[0][t][i];
                    const float den = yv * yv + kLossFloor;
                    const float diff = yv - tt[i];
                    lane_loss += diff * diff / den * inv_norm;
                    d[i] = 2.0f * diff / den * inv_norm;
                }
            }
            dz5[0][t] = pack_frag(d, f32x4{0.0f, 0.0f, 0.0f, 0.0f});
        }

        // backward
        f32x4 dx[4][QT];
        u32x4 dz[2][QT];
        dw_step<5, 1, QT>(xs, ds, h4, dz5, dW, g, r);
        bwd_layer<5, QT>(P.wk, dz5, dx, lane);
        gate_pack<QT>(dx, h4, dz);
        dw_step<4, 2, QT>(xs, ds, h3, dz, dW, g, r);
        bwd_layer<4, QT>(P.wk, dz, dx, lane);
        gate_pack<QT>(dx, h3, dz);
        dw_step<3, 2, QT>(xs, ds, h2, dz, dW, g, r);
        bwd_layer<3, QT>(P.wk, dz, dx, lane);
        gate_pack<QT>(dx, h2, dz);
        dw_step<2, 2, QT>(xs, ds, h1, dz, dW, g, r);
        bwd_layer<2, QT>(P.wk, dz, dx, lane);
        gate_pack<QT>(dx, h1, dz);
        dw_step<1, 2, QT>(xs, ds, h0, dz, dW, g, r);
        bwd_layer<1, QT>(P.wk, dz, dx, lane);
        gate_pack<QT>(dx, h0, dz);
        dw_step<0, 2, QT>(xs, ds, x0, dz, dW, g, r);
        f32x4 dx0[2][QT];
        bwd_layer<0, QT>(P.wk, dz, dx0, lane);

        // input-feature gradients -> hash table
#pragma unroll
        for (int t = 0; t < QT; t++) {
            const uint32_t qi = base + 16 * t + r;
            if (qi >= count) continue;
            const float df[8] = {dx0[0][t][0], dx0[0][t][1], dx0[0][t][2], dx0[0][t][3],
                                 dx0[1][t][0], dx0[1][t][1], dx0[1][t][2], dx0[1][t][3]};
            lane_table_grad(P, q, qi, g, df, grad_table);
        }
    }

    // loss: one atomic per wave
    for (int m = 32; m >= 1; m >>= 1) lane_loss += __shfl_xor(lane_loss, m, 64);
    if (lane == 0 && lane_loss != 0.0f) atomicAdd(loss, lane_loss);
    __syncthreads();
    float* dst = partial + (size_t)blockIdx.x * kMasterSize;
    for (int j = threadIdx.x; j < kMasterSize; j += kTrainBlock) dst[j] = dW[j];
}

__global__ void k_mlp_reduce(const float* __restrict__ partial, uint32_t blocks, float* __restrict__ grad_out,
                             float* __restrict__ master, float* __restrict__ m, float* __restrict__ v, float lr_t) {
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= (uint32_t)kMasterSize) return;
    float s = 0.0f;
    for (uint32_t b = 0; b < blocks; b++) s += partial[(size_t)b * kMasterSize + j];
    if (m == nullptr) {
        grad_out[j] = s;
        return;
    }
    const float mj = kBeta1 * m[j] + (1.0f - kBeta1) * s;
    const float vj = kBeta2 * v[j] + (1.0f - kBeta2) * s * s;
    m[j] = mj;
    v[j] = vj;
    master[j] -= lr_t * mj / (sqrtf(vj) + kEps);
}

__global__ void k_table_adam(float* __restrict__ t32, uint16_t* __restrict__ t16, float* __restrict__ m,
                             float* __restrict__ v, float* __restrict__ grad, uint32_t n, float lr_t) {
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    const float gj = grad[j];
    if (gj == 0.0f) return;
    grad[j] = 0.0f;
    const float mj = kBeta1 * m[j] + (1.0f - kBeta1) * gj;
    const float vj = kBeta2 * v[j] + (1.0f - kBeta2) * gj * gj;
    m[j] = mj;
    v[j] = vj;
    const float p = t32[j] - lr_t * mj / (sqrtf(vj) + kEps);
    t32[j] = p;
    t16[j] = float_to_half(p);
}

__device__ __forceinline__ uint16_t bf16_of(float f) { return (uint16_t)bf16_bits(f); }

__global__ void k_pack(const float* __restrict__ master, uint16_t* __restrict__ wn, uint16_t* __restrict__ wk) {
    const uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
    const bool fwd = e < (uint32_t)kWnSize;
    if (e >= (uint32_t)(kWnSize + kWkSize)) return;
    uint32_t local = fwd ? e : e - kWnSize;
    int L = 0;
#pragma unroll
    for (int l = 0; l < kLayers; l++) {
        const uint32_t off = fwd ? (uint32_t)wn_offset(l) : (uint32_t)wk_offset(l);
        if (local >= off) L = l;
    }
    local -= fwd ? (uint32_t)wn_offset(L) : (uint32_t)wk_offset(L);
    const int K = kK[L], N = kN[L];
    const int i = local & 7, lane = (local >> 3) & 63, frag = local >> 9;
    const int g = lane >> 4, r = lane & 15;
    int in, out;
    if (fwd) {
        const int S = K / 32;
        const int t = frag / S, s = frag % S;
        in = 32 * s + 16 * (i >> 2) + 4 * g + (i & 3);
        out = 16 * t + r;
    } else {
        const int S = nk_cols(L) / 32;
        const int t = frag / S, s = frag % S;
        in = 16 * t + r;
        out = 32 * s + 16 * (i >> 2) + 4 * g + (i & 3);
    }
    const uint16_t v = out < N ? bf16_of(master[master_offset(L) + in * N + out]) : (uint16_t)0;
    if (fwd)
        wn[e] = v;
    else
        wk[e - kWnSize] = v;
}

__global__ void k_fill(float* p, uint32_t n, float value) {
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j < n) p[j] = value;
}

__global__ void k_f32_to_f16(const float* __restrict__ a, uint16_t* __restrict__ b, uint32_t n) {
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j < n) b[j] = float_to_half(a[j]);
}

__global__ void k_f16_to_f32(const uint16_t* __restrict__ a, float* __restrict__ b, uint32_t n) {
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j < n) b[j] = half_to_float(a[j]);
}

// ---------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------

static inline uint32_t blocks_for(uint32_t n, uint32_t per) { return (n + per - 1) / per; }

hipError_t launch_infer(const NrcDev& d, const float4* queries, float4* out, const uint32_t* d_count, uint32_t count,
                        hipStream_t stream) {
    if (count == 0) return hipSuccess;
    const uint32_t per_block = 256 / 64 * kInferQueriesPerWave;
    hipLaunchKernelGGL(k_nrc_infer, dim3(blocks_for(count, per_block)), dim3(256), 0, stream, d, queries, out, d_count,
                       count);
    return hipGetLastError();
}

hipError_t launch_train(const NrcDev& d, const float4* queries, const float4* targets, const uint32_t* d_count,
                        uint32_t count, uint32_t slice, uint32_t slices, float* partial, uint32_t max_blocks,
                        float* grad_table, float* loss, uint32_t* blocks, hipStream_t stream) {
    hipLaunchKernelGGL(k_fill, dim3(1), dim3(64), 0, stream, loss, 1u, 0.0f);
    // with a device-resident count the grid covers the largest slice
    const uint32_t most = d_count ? blocks_for(count, slices) : count;
    const uint32_t chunks = blocks_for(most, kTrainQueriesPerWave);
    uint32_t nb = blocks_for(chunks, kTrainBlock / 64);
    if (nb > max_blocks) nb = max_blocks;
    if (nb == 0) nb = 1;  // an empty slice still writes zero partials
    *blocks = nb;
    hipLaunchKernelGGL(k_nrc_train, dim3(nb), dim3(kTrainBlock), 0, stream, d, queries, targets, d_count, count, slice,
                       slices, partial, grad_table, loss);
    return hipGetLastError();
}

hipError_t launch_mlp_reduce(const float* partial, uint32_t blocks, float* grad_out, float* master, float* m, float* v,
                             float lr_t, hipStream_t stream) {
    hipLaunchKernelGGL(k_mlp_reduce, dim3(blocks_for(kMasterSize, 256)), dim3(256), 0, stream, partial, blocks,
                       grad_out, master, m, v, lr_t);
    return hipGetLastError();
}

hipError_t launch_table_adam(float* table32, uint16_t* table16, float* m, float* v, float* grad, uint32_t n,
                             float lr_t, hipStream_t stream) {
    hipLaunchKernelGGL(k_table_adam, dim3(blocks_for(n, 256)), dim3(256), 0, stream, table32, table16, m, v, grad, n,
                       lr_t);
    return hipGetLastError();
}

hipError_t launch_pack(const float* master, uint16_t* wn, uint16_t* wk, hipStream_t stream) {
    hipLaunchKernelGGL(k_pack, dim3(blocks_for(kWnSize + kWkSize, 256)), dim3(256), 0, stream, master, wn, wk);
    return hipGetLastError();
}

hipError_t launch_fill(float* p, uint32_t n, float value, hipStream_t stream) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_fill, dim3(blocks_for(n, 256)), dim3(256), 0, stream, p, n, value);
    return hipGetLastError();
}

hipError_t launch_table_from_f32(const float* t32, uint16_t* t16, uint32_t n, hipStream_t stream) {
    hipLaunchKernelGGL(k_f32_to_f16, dim3(blocks_for(n, 256)), dim3(256), 0, stream, t32, t16, n);
    return hipGetLastError();
}

hipError_t launch_table_from_f16(const uint16_t* t16, float* t32, uint32_t n, hipStream_t stream) {
    hipLaunchKernelGGL(k_f16_to_f32, dim3(blocks_for(n, 256)), dim3(256), 0, stream, t16, t32, n);
    return hipGetLastError();
}

}  // namespace nrc
}  // namespace atlas_rt
