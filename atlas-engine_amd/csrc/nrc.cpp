// Radiance cache host object and C ABI: parameter validation, level
// resolutions, deterministic initialisation, and the training step
// (forward/backward kernel, gradient reduction + Adam, weight packing,
// sparse table Adam).
#include "nrc.h"

#include <algorithm>
#include <cmath>
#include <vector>

using namespace atlas_rt;
using namespace atlas_rt::nrc;

namespace {

template <class T>
atlas_rt_status dev_alloc(T*& p, size_t count) {
    void* v = nullptr;
    ATLAS_HIP_TRY(hipMalloc(&v, std::max<size_t>(count, 1) * sizeof(T)));
    p = reinterpret_cast<T*>(v);
    return ATLAS_RT_OK;
}

void free_all(atlas_rt_nrc& N) {
    void* ptrs[] = {N.master, N.m_mlp, N.v_mlp, N.wn, N.wk, N.table32, N.m_tab, N.v_tab, N.grad_tab, N.table16,
                    N.partial, N.loss};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
}

// SplitMix64 stream -> uniform floats in [0, 1) from 24-bit mantissas.
struct SplitMix {
    uint64_t s;
    uint64_t next() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    float uniform() { return float(next() >> 40) * (1.0f / 16777216.0f); }
};

atlas_rt_status hip_check(hipError_t e, const char* what) { return e == hipSuccess ? ATLAS_RT_OK : hip_fail(e, what); }

atlas_rt_status reset_moments(atlas_rt_nrc& N, hipStream_t s) {
    atlas_rt_status st = hip_check(launch_fill(N.m_mlp, kMasterSize, 0.0f, s), "nrc fill");
    if (st == ATLAS_RT_OK) st = hip_check(launch_fill(N.v_mlp, kMasterSize, 0.0f, s), "nrc fill");
    if (st == ATLAS_RT_OK) st = hip_check(launch_fill(N.m_tab, N.table_params, 0.0f, s), "nrc fill");
    if (st == ATLAS_RT_OK) st = hip_check(launch_fill(N.v_tab, N.table_params, 0.0f, s), "nrc fill");
    N.step = 0;
    return st;
}

atlas_rt_status check_device(const atlas_rt_nrc* N, const char* fn) {
    int dev = -1;
    ATLAS_HIP_TRY(hipGetDevice(&dev));
    if (dev != N->device) return fail(ATLAS_RT_E_INVALID_ARG, std::string(fn) + ": current device differs from the cache's");
    return ATLAS_RT_OK;
}

}  // namespace

atlas_rt_status atlas_rt::nrc::train_step(atlas_rt_nrc& N, const float4* queries, const float4* targets,
                                          const uint32_t* d_count, uint32_t count, uint32_t slice, uint32_t slices,
                                          float* d_loss, hipStream_t s) {
    uint32_t blocks = 0;
    atlas_rt_status st = hip_check(launch_train(N.dev, queries, targets, d_count, count, slice, slices, N.partial,
                                                N.max_blocks, N.grad_tab, d_loss ? d_loss : N.loss, &blocks, s),
                                   "nrc train launch");
    if (st != ATLAS_RT_OK) return st;
    N.step++;
    // Adam bias correction folded into the step size
    const double b1t = std::pow(0.9, double(N.step)), b2t = std::pow(0.99, double(N.step));
    const float lr_t = float(double(N.params.learning_rate) * std::sqrt(1.0 - b2t) / (1.0 - b1t));
    st = hip_check(launch_mlp_reduce(N.partial, blocks, nullptr, N.master, N.m_mlp, N.v_mlp, lr_t, s), "nrc adam");
    if (st == ATLAS_RT_OK) st = hip_check(launch_pack(N.master, N.wn, N.wk, s), "nrc pack");
    if (st == ATLAS_RT_OK)
        st = hip_check(launch_table_adam(N.table32, N.table16, N.m_tab, N.v_tab, N.grad_tab, N.table_params, lr_t, s),
                       "nrc table adam");
    return st;
}

extern "C" {

atlas_rt_status atlas_rt_nrc_create(const atlas_rt_nrc_params* params, atlas_rt_nrc** out) {
    if (!params || !out) return fail(ATLAS_RT_E_INVALID_ARG, "atlas_rt_nrc_create: NULL argument");
    *out = nullptr;
    const atlas_rt_nrc_params& p = *params;
    if (p.levels < 1 || p.levels > (uint32_t)kMaxLevels)
        return fail(ATLAS_RT_E_INVALID_ARG, "atlas_rt_nrc_create: levels must be 1..12");
    if (p.log2_table < 10 || p.log2_table > 24)
        return fail(ATLAS_RT_E_INVALID_ARG, "atlas_rt_nrc_create: log2_table must be 10..24");
    if (p.base_resolution < 2 || p.max_resolution < p.base_resolution || p.max_resolution > (1u << 20))
        return fail(ATLAS_RT_E_INVALID_ARG, "atlas_rt_nrc_create: need 2 <= base_resolution <= max_resolution <= 2^20");
    if (!(p.learning_rate > 0.0f) || !std::isfinite(p.learning_rate))
        return fail(ATLAS_RT_E_INVALID_ARG, "atlas_rt_nrc_create: learning_rate must be positive");
    for (int a = 0; a < 3; a++)
        if (!(p.bounds_max[a] > p.bounds_min[a]) || !std::isfinite(p.bounds_min[a]) || !std::isfinite(p.bounds_max[a]))
            return fail(ATLAS_RT_E_INVALID_ARG, "atlas_rt_nrc_create: bounds_max must exceed bounds_min on every axis");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
        return fail(ATLAS_RT_E_NO_DEVICE, "atlas_rt_nrc_create: no HIP device");

    auto* nrc = new atlas_rt_nrc();
    atlas_rt_nrc& N = *nrc;
    N.params = p;
    (void)hipGetDevice(&N.device);
    const uint64_t T = 1ull << p.log2_table;
    N.table_params = uint32_t(p.levels * T * 2);
    NrcDev& d = N.dev;
    d.levels = p.levels;
    d.log2_t = p.log2_table;
    // level l has floor(base * b^l) cells per axis, b = (max / base)^(1 / (levels - 1))
    const double b = p.levels > 1 ? std::exp((std::log(double(p.max_resolution)) - std::log(double(p.base_resolution))) /
                                             double(p.levels - 1))
                                  : 1.0;
    d.dense_mask = 0;
    for (uint32_t l = 0; l < (uint32_t)kMaxLevels; l++) {
        const double r = l < p.levels ? std::floor(double(p.base_resolution) * std::pow(b, double(l)) + 1e-9) : 1.0;
        d.res[l] = float(r);
        const double cells = (r + 1.0) * (r + 1.0) * (r + 1.0);
        if (l < p.levels && cells <= double(T)) d.dense_mask |= 1u << l;
    }
    for (int a = 0; a < 3; a++) {
        d.bmin[a] = p.bounds_min[a];
        d.binv[a] = float(1.0 / (double(p.bounds_max[a]) - double(p.bounds_min[a])));
    }
    N.max_blocks = uint32_t(std::max(1, device_cu_count()));

    atlas_rt_status st = ATLAS_RT_OK;
    if (st == ATLAS_RT_OK) st = dev_alloc(N.master, kMasterSize);
    if (st == ATLAS_RT_OK) st = dev_alloc(N.m_mlp, kMasterSize);
    if (st == ATLAS_RT_OK) st = dev_alloc(N.v_mlp, kMasterSize);
    if (st == ATLAS_RT_OK) st = dev_alloc(N.wn, kWnSize);
    if (st == ATLAS_RT_OK) st = dev_alloc(N.wk, kWkSize);
    if (st == ATLAS_RT_OK) st = dev_alloc(N.table32, N.table_params);
    if (st == ATLAS_RT_OK) st = dev_alloc(N.m_tab, N.table_params);
    if (st == ATLAS_RT_OK) st = dev_alloc(N.v_tab, N.table_params);
    if (st == ATLAS_RT_OK) st = dev_alloc(N.grad_tab, N.table_params);
    if (st == ATLAS_RT_OK) st = dev_alloc(N.table16, N.table_params);
    if (st == ATLAS_RT_OK) st = dev_alloc(N.partial, size_t(N.max_blocks) * kMasterSize);
    if (st == ATLAS_RT_OK) st = dev_alloc(N.loss, 1);
    if (st != ATLAS_RT_OK) {
        free_all(N);
        delete nrc;
        return st;
    }
    d.table = N.table16;
    d.wn = N.wn;
    d.wk = N.wk;

    // Initialisation: weights U(-sqrt(6 / (fan_in + fan_out)), +...), table U(-1e-4, 1e-4).
    SplitMix rng{0xA71A5ull * 0x100000001B3ull + p.seed};
    std::vector<float> w(kMasterSize);
    for (int l = 0; l < kLayers; l++) {
        const float lim = std::sqrt(6.0f / float(kK[l] + kN[l]));
        for (int j = 0; j < kK[l] * kN[l]; j++) w[master_offset(l) + j] = (2.0f * rng.uniform() - 1.0f) * lim;
    }
    std::vector<float> tab(N.table_params);
    for (auto& x : tab) x = (2.0f * rng.uniform() - 1.0f) * 1e-4f;
    hipStream_t s = nullptr;
    if (st == ATLAS_RT_OK) st = hip_check(hipMemcpy(N.master, w.data(), w.size() * 4, hipMemcpyHostToDevice), "hipMemcpy");
    if (st == ATLAS_RT_OK)
        st = hip_check(hipMemcpy(N.table32, tab.data(), tab.size() * 4, hipMemcpyHostToDevice), "hipMemcpy");
    if (st == ATLAS_RT_OK) st = hip_check(launch_table_from_f32(N.table32, N.table16, N.table_params, s), "nrc init");
    if (st == ATLAS_RT_OK) st = hip_check(launch_pack(N.master, N.wn, N.wk, s), "nrc pack");
    if (st == ATLAS_RT_OK) st = hip_check(launch_fill(N.grad_tab, N.table_params, 0.0f, s), "nrc fill");
    if (st == ATLAS_RT_OK) st = reset_moments(N, s);
    if (st == ATLAS_RT_OK) st = hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
    if (st != ATLAS_RT_OK) {
        free_all(N);
        delete nrc;
        return st;
    }
    *out = nrc;
    return ATLAS_RT_OK;
}

atlas_rt_status atlas_rt_nrc_destroy(atlas_rt_nrc* nrc) {
    if (!nrc) return fail(ATLAS_RT_E_INVALID_ARG, "atlas_rt_nrc_destroy: NULL handle");
    (void)hipDeviceSynchronize();
    free_all(*nrc);
    delete nrc;
    return ATLAS_RT_OK;
}

atlas_rt_status atlas_rt_nrc_infer(atlas_rt_nrc* nrc, const float* d_queries, float* d_out, uint32_t count,
                                   void* stream) {
    if (!nrc) return fail(ATLAS_RT_E_INVALID_ARG, "atlas_rt_nrc_infer: NULL handle");
    if (count == 0) return ATLAS_RT_OK;
    if (!d_queries || !d_out) return fail(ATLAS_RT_E_INVALID_ARG, "atlas_rt_nrc_infer: NULL buffer");
    if ((reinterpret_cast<uintptr_t>(d_queries) | reinterpret_cast<uintptr_t>(d_out)) & 15u)
        return fail(ATLAS_RT_E_INVALID_ARG, "atlas_rt_nrc_infer: buffers must be 16-byte aligned");
    atlas_rt_status st = check_device(nrc, "atlas_rt_nrc_infer");
    if (st != ATLAS_RT_OK) return st;
    return hip_check(launch_infer(nrc->dev, reinterpret_cast<const float4*>(d_queries), reinterpret_cast<float4*>(d_out),
                                  nullptr, count, static_cast<hipStream_t>(stream)),
                     "nrc infer launch");
}

static atlas_rt_status check_train_args(atlas_rt_nrc* nrc, const float* q, const float* t, uint32_t count,
                                        const char* fn) {
    if (!nrc) return fail(ATLAS_RT_E_INVALID_ARG, std::string(fn) + ": NULL handle");
    if (count && (!q || !t)) return fail(ATLAS_RT_E_INVALID_ARG, std::string(fn) + ": NULL buffer");
    if ((reinterpret_cast<uintptr_t>(q) | reinterpret_cast<uintptr_t>(t)) & 15u)
        return fail(ATLAS_RT_E_INVALID_ARG, std::string(fn) + ": buffers must be 16-byte aligned");
    return check_device(nrc, fn);
}

atlas_rt_status atlas_rt_nrc_train(atlas_rt_nrc* nrc, const float* d_queries, const float* d_targets, uint32_t count,
                                   float* d_loss, void* stream) {
    atlas_rt_status st = check_train_args(nrc, d_queries, d_targets, count, "atlas_rt_nrc_train");
    if (st != ATLAS_RT_OK) return st;
    if (count == 0) return ATLAS_RT_OK;
    return train_step(*nrc, reinterpret_cast<const float4*>(d_queries), reinterpret_cast<const float4*>(d_targets),
                      nullptr, count, 0, 1, d_loss, static_cast<hipStream_t>(stream));
}

atlas_rt_status atlas_rt_nrc_gradients(atlas_rt_nrc* nrc, const float* d_queries, const float* d_targets,
                                       uint32_t count, float* d_grad_mlp, float* d_grad_table, void* stream) {
    atlas_rt_status st = check_train_args(nrc, d_queries, d_targets, count, "atlas_rt_nrc_gradients");
    if (st != ATLAS_RT_OK) return st;
    if (!d_grad_mlp || !d_grad_table) return fail(ATLAS_RT_E_INVALID_ARG, "atlas_rt_nrc_gradients: NULL output");
    atlas_rt_nrc& N = *nrc;
    hipStream_t s = static_cast<hipStream_t>(stream);
    st = hip_check(launch_fill(d_grad_table, N.table_params, 0.0f, s), "nrc fill");
    uint32_t blocks = 0;
    if (st == ATLAS_RT_OK)
        st = hip_check(launch_train(N.dev, reinterpret_cast<const float4*>(d_queries),
                                    reinterpret_cast<const float4*>(d_targets), nullptr, count, 0, 1, N.partial,
                                    N.max_blocks, d_grad_table, N.loss, &blocks, s),
                       "nrc train launch");
    if (st == ATLAS_RT_OK)
        st = hip_check(launch_mlp_reduce(N.partial, blocks, d_grad_mlp, nullptr, nullptr, nullptr, 0.0f, s),
                       "nrc reduce");
    return st;
}

atlas_rt_status atlas_rt_nrc_get_params(atlas_rt_nrc* nrc, float* d_mlp, uint16_t* d_table) {
    if (!nrc) return fail(ATLAS_RT_E_INVALID_ARG, "atlas_rt_nrc_get_params: NULL handle");
    atlas_rt_status st = check_device(nrc, "atlas_rt_nrc_get_params");
    if (st != ATLAS_RT_OK) return st;
    if (d_mlp) ATLAS_HIP_TRY(hipMemcpy(d_mlp, nrc->master, kMasterSize * sizeof(float), hipMemcpyDeviceToDevice));
    if (d_table)
        ATLAS_HIP_TRY(hipMemcpy(d_table, nrc->table16, size_t(nrc->table_params) * 2, hipMemcpyDeviceToDevice));
    return ATLAS_RT_OK;
}

atlas_rt_status atlas_rt_nrc_set_params(atlas_rt_nrc* nrc, const float* d_mlp, const uint16_t* d_table) {
    if (!nrc) return fail(ATLAS_RT_E_INVALID_ARG, "atlas_rt_nrc_set_params: NULL handle");
    atlas_rt_status st = check_device(nrc, "atlas_rt_nrc_set_params");
    if (st != ATLAS_RT_OK) return st;
    atlas_rt_nrc& N = *nrc;
    if (d_mlp) {
        ATLAS_HIP_TRY(hipMemcpy(N.master, d_mlp, kMasterSize * sizeof(float), hipMemcpyDeviceToDevice));
        st = hip_check(launch_pack(N.master, N.wn, N.wk, nullptr), "nrc pack");
        if (st != ATLAS_RT_OK) return st;
    }
    if (d_table) {
        ATLAS_HIP_TRY(hipMemcpy(N.table16, d_table, size_t(N.table_params) * 2, hipMemcpyDeviceToDevice));
        st = hip_check(launch_table_from_f16(N.table16, N.table32, N.table_params, nullptr), "nrc table");
        if (st != ATLAS_RT_OK) return st;
    }
    st = reset_moments(N, nullptr);
    if (st != ATLAS_RT_OK) return st;
    ATLAS_HIP_TRY(hipStreamSynchronize(nullptr));
    return ATLAS_RT_OK;
}

atlas_rt_status atlas_rt_nrc_sizes(const atlas_rt_nrc* nrc, uint32_t* mlp_params, uint32_t* table_params) {
    if (!nrc) return fail(ATLAS_RT_E_INVALID_ARG, "atlas_rt_nrc_sizes: NULL handle");
    if (mlp_params) *mlp_params = kMasterSize;
    if (table_params) *table_params = nrc->table_params;
    return ATLAS_RT_OK;
}

}  // extern "C"
