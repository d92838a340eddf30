// Neural radiance cache: network shapes, packed weight layouts, the device
// view, the host object and the kernel launchers (nrc_kernels.hip).
//
// Network: 32 inputs -> 5 hidden layers of 64 ReLU units -> 3 linear
// outputs, no biases (the constant-1 input stands in for a first-layer bias).
// Every layer runs on v_mfma_f32_16x16x32_bf16 in the transposed form
// Z^T = W^T X^T: the A operand is a 16x32 weight fragment, the B operand a
// 32-input x 16-query activation fragment, the accumulator a 16-unit x
// 16-query tile.  Lane l = 16 g + r of a wave holds, in an accumulator tile,
// query r and units 4 g + 0..3.  The B operand of the next layer is built in
// the same lane from two such tiles: K slot (g, i) of K step s stands for
// unit  P(s, g, i) = 32 s + 16 (i >> 2) + 4 g + (i & 3).
// The weight fragments are stored pre-permuted with the same P, so no data
// leaves its lane between layers.
#pragma once

#include "rt_internal.h"

namespace atlas_rt {
namespace nrc {

constexpr int kLayers = 6;
constexpr int kIn = 32;
constexpr int kHid = 64;
constexpr int kOut = 3;
constexpr int kMaxLevels = 12;
constexpr int kK[kLayers] = {32, 64, 64, 64, 64, 64};   // inputs per layer
constexpr int kN[kLayers] = {64, 64, 64, 64, 64, 3};    // outputs per layer
constexpr int kNp[kLayers] = {64, 64, 64, 64, 64, 16};  // outputs padded to 16-unit tiles

// FP32 master weights: layer by layer, [K][N] row-major (input-major).
constexpr int master_offset(int l) {
    int o = 0;
    for (int i = 0; i < l; i++) o += kK[i] * kN[i];
    return o;
}
constexpr int kMasterSize = master_offset(kLayers);  // 18,624

// BF16 forward fragments ("wn"): per layer, fragment (t, s) for output tile t
// and K step s holds 64 lanes x 8 values, value i of lane 16 g + r =
// W[P(s, g, i)][16 t + r] (zero past the real outputs).
constexpr int wn_offset(int l) {
    int o = 0;
    for (int i = 0; i < l; i++) o += kNp[i] * kK[i];
    return o;
}
// BF16 backward fragments ("wk"): per layer, fragment (t, s) for input tile
// t and output K step s, value i of lane 16 g + r = W[16 t + r][P(s, g, i)]
// (zero past the real outputs).  Output K is padded to at least 32.
constexpr int nk_cols(int l) { return kNp[l] < 32 ? 32 : kNp[l]; }
constexpr int wk_offset(int l) {
    int o = 0;
    for (int i = 0; i < l; i++) o += kK[i] * nk_cols(i);
    return o;
}
constexpr int kWnSize = wn_offset(kLayers);
constexpr int kWkSize = wk_offset(kLayers);

constexpr int kTrainBlock = 256;         // 4 waves
constexpr int kTrainQueriesPerWave = 32;  // two 16-query tiles per chunk
constexpr int kInferQueriesPerWave = 64;  // four 16-query tiles per wave

struct NrcDev {
    const uint16_t* table;   // FP16 [levels][T][2]
    const uint16_t* wn;
    const uint16_t* wk;
    float res[kMaxLevels];   // cells per axis of each level (integer valued)
    uint32_t dense_mask;     // bit l: level l is indexed densely ((res + 1)^3 <= T)
    uint32_t levels;
    uint32_t log2_t;
    float bmin[3];
    float binv[3];           // 1 / (bmax - bmin)
};

// Launchers; every launch goes to `stream`.
// Counts: with d_count set the kernels read the count from device memory
// (at most `count`), else `count` itself.
hipError_t launch_infer(const NrcDev& d, const float4* queries, float4* out, const uint32_t* d_count, uint32_t count,
                        hipStream_t stream);
// Forward + backward over slice `slice` of `slices` equal slices of the
// samples: per-block weight-gradient partial sums into
// partial[blocks][kMasterSize], table gradients added into grad_table, loss
// written to *loss.  Returns the block count used through *blocks.
hipError_t launch_train(const NrcDev& d, const float4* queries, const float4* targets, const uint32_t* d_count,
                        uint32_t count, uint32_t slice, uint32_t slices, float* partial, uint32_t max_blocks,
                        float* grad_table, float* loss, uint32_t* blocks, hipStream_t stream);
// Sum the partials; with m/v set, apply Adam to `master` (lr_t already
// bias-corrected), else write the sums to grad_out.
hipError_t launch_mlp_reduce(const float* partial, uint32_t blocks, float* grad_out, float* master, float* m, float* v,
                             float lr_t, hipStream_t stream);
// Adam on the table entries with a non-zero gradient; clears the gradient.
hipError_t launch_table_adam(float* table32, uint16_t* table16, float* m, float* v, float* grad, uint32_t n,
                             float lr_t, hipStream_t stream);
// FP32 master -> packed BF16 forward/backward fragments.
hipError_t launch_pack(const float* master, uint16_t* wn, uint16_t* wk, hipStream_t stream);
hipError_t launch_fill(float* p, uint32_t n, float value, hipStream_t stream);
hipError_t launch_table_from_f32(const float* t32, uint16_t* t16, uint32_t n, hipStream_t stream);
hipError_t launch_table_from_f16(const uint16_t* t16, float* t32, uint32_t n, hipStream_t stream);

}  // namespace nrc
}  // namespace atlas_rt

struct atlas_rt_nrc;

namespace atlas_rt {
namespace nrc {
// One Adam step (forward/backward, weight-gradient reduction + update,
// weight packing, table update) on a slice of the samples; see launch_train.
atlas_rt_status train_step(atlas_rt_nrc& N, const float4* queries, const float4* targets, const uint32_t* d_count,
                           uint32_t count, uint32_t slice, uint32_t slices, float* d_loss, hipStream_t s);
}  // namespace nrc
}  // namespace atlas_rt

struct atlas_rt_nrc {
    atlas_rt_nrc_params params{};
    atlas_rt::nrc::NrcDev dev{};
    int device = -1;
    uint32_t table_params = 0;   // levels * T * 2
    uint32_t max_blocks = 0;     // training grid cap (one block per CU)
    float* master = nullptr;     // FP32 MLP weights
    float* m_mlp = nullptr;      // Adam moments
    float* v_mlp = nullptr;
    uint16_t* wn = nullptr;
    uint16_t* wk = nullptr;
    float* table32 = nullptr;    // FP32 master table
    float* m_tab = nullptr;
    float* v_tab = nullptr;
    float* grad_tab = nullptr;   // kept zero between steps
    uint16_t* table16 = nullptr;
    float* partial = nullptr;    // [max_blocks][kMasterSize]
    float* loss = nullptr;       // internal loss slot when the caller passes none
    int32_t step = 0;
};
