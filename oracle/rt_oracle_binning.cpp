// TEST INFRASTRUCTURE - the oracle.  Only tests/, __graft_entry__.smoke() and
// bench.py's cpu_baseline leg may load it, and only as the checker.
//
// Serial CPU restatement of ray coherence sorting (the product's
// atlas_rt_sort_rays_by_direction and the path tracer's ray binning):
// the bin of a direction is DetermineRayBin (data/shader/raytracer/
// tracing.hsh:18-21) over UnitVectorToOctahedron (data/shader/common/
// octahedron.hsh:23-37), with the cell clamped to the 8 x 8 grid; rays are
// ordered by bin (binningOffset.csh: exclusive prefix sum of the bin counts)
// and keep their input order inside a bin (binning.csh appends with atomics,
// so the engine's order inside a bin is arbitrary).  No FMA contraction.
#include <cmath>
#include <cstdint>
#include <vector>

#include "../include/atlas_rt.h"

extern "C" {

uint32_t oracle_ray_bin(float x, float y, float z) {
    const float l1 = (std::fabs(x) + std::fabs(y)) + std::fabs(z);
    float ox = x / l1, oz = z / l1;
    if (y < 0.0f) {  // lower hemisphere: fold over the diagonals
        const float a = ox, b = oz;
        ox = (a >= 0.0f ? 1.0f : -1.0f) * (1.0f - std::fabs(b));
        oz = (b >= 0.0f ? 1.0f : -1.0f) * (1.0f - std::fabs(a));
    }
    // saturate(0.5 * . + 0.5); a NaN coordinate saturates to 0
    float u = 0.5f * ox + 0.5f, v = 0.5f * oz + 0.5f;
    u = std::fmin(std::fmax(u, 0.0f), 1.0f);
    v = std::fmin(std::fmax(v, 0.0f), 1.0f);
    uint32_t cx = uint32_t(u * 8.0f), cy = uint32_t(v * 8.0f);
    if (cx > 7u) cx = 7u;
    if (cy > 7u) cy = 7u;
    return cy * 8u + cx;  // Flatten2D(coord, ivec2(8))
}

void oracle_ray_bins(const float* dirs, uint32_t count, uint32_t* bins) {
    for (uint32_t i = 0; i < count; i++) bins[i] = oracle_ray_bin(dirs[3 * i], dirs[3 * i + 1], dirs[3 * i + 2]);
}

// out[i] = in[perm[i]]; offsets[65]: first index of each bin, then count.
void oracle_sort_rays_by_direction(const atlas_rt_ray* in, uint32_t count, atlas_rt_ray* out, uint32_t* perm,
                                   uint32_t* offsets) {
    std::vector<uint32_t> bins(count), start(65, 0);
    for (uint32_t i = 0; i < count; i++) {
        bins[i] = oracle_ray_bin(in[i].direction[0], in[i].direction[1], in[i].direction[2]);
        start[bins[i] + 1]++;
    }
    for (int b = 0; b < 64; b++) start[size_t(b) + 1] += start[size_t(b)];
    for (int b = 0; b <= 64; b++) offsets[b] = start[size_t(b)];
    for (uint32_t i = 0; i < count; i++) {
        const uint32_t dst = start[bins[i]]++;
        out[dst] = in[i];
        perm[dst] = i;
    }
}

}  // extern "C"
