// Gather gridding onto the (oversampled) Cartesian grid, driven by a Plan2.
#pragma once

#include "plan.hpp"

namespace ksj::gpu {

constexpr uint32_t kCoilGroup = 8;  // coils accumulated per thread in the gather

inline uint32_t padded_coils(uint32_t coils) { return (coils + kCoilGroup - 1) / kCoilGroup * kCoilGroup; }

// Applies the plan's per-sample weight and moves channel-major samples [C][S]
// into sorted, coil-fastest rows out[k][Cp] (Cp = padded_coils(C), pad zeroed).
void grid_prep(const Plan2& plan, const float2* d_kspace, uint32_t coils, float2* d_sorted, hipStream_t stream);

// Gathers every grid cell of all coils: grid[c][G_r][G_c] (plane-major).  Cells
// with more than heavy_threshold contributing samples are reduced by one
// workgroup each in a fixed order; all others by one thread per cell and coil
// group.  Every cell is written (no prior clear needed) and the result is
// bit-reproducible for a given plan.
void grid_gather(const Plan2& plan, const float2* d_sorted, uint32_t coils, float2* d_grid, hipStream_t stream);

}  // namespace ksj::gpu
