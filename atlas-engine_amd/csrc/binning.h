// Ray coherence sorting: rays grouped by the octahedral bin of their
// direction (raytracer/tracing.hsh:18-29, binning.csh, binningOffset.csh;
// dispatched by RayTracingHelper.cpp:304-344, where the engine has it
// switched off).
//
// Bin: the direction's octahedral map (common/octahedron.hsh:23-37: divide x
// and z by |x| + |y| + |z|, fold the lower hemisphere, then 0.5 * . + 0.5
// clamped to [0, 1]) scaled by 8 and truncated, cell (cx, cy) -> cy * 8 + cx.
// A coordinate of exactly 1 truncates to 8 in the engine, one past its 8x8
// grid (and up to bin 72 of its 64 counters); here it is clamped to 7.  A NaN
// direction falls in bin 0.
//
// Order: the engine appends with atomics (an arbitrary order inside a bin);
// this sort is stable (rays keep their order inside a bin), so the output is
// deterministic.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace atlas_rt {

constexpr uint32_t kRayBins = 64;

// Device workspace bytes for sorting up to `cap` records.
size_t bin_sort_workspace(uint32_t cap);

// Stable counting sort by direction bin of `rec`-float4 records (rec = 2: an
// atlas_rt_ray; rec = 3: a path record), direction in float4 [1].xyz.
// The count is *d_count when d_count is non-null (a device counter), else
// count; at most cap.  perm (optional): out[i] = in[perm[i]].  bin_offsets
// (optional, 65 entries): the first output index of each bin, then the count
// (all zero when cap is 0).  The input and output must not overlap.
hipError_t launch_bin_sort(const float4* in, float4* out, int rec, const uint32_t* d_count, uint32_t count,
                           uint32_t cap, void* workspace, uint32_t* perm, uint32_t* bin_offsets, hipStream_t s);

}  // namespace atlas_rt
