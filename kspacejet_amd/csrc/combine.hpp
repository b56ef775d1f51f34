// Coil combination and layout kernels.
#pragma once

#include "common.hpp"

namespace ksj::gpu {

// Root sum of squares over the contiguous coil axis of coil-last images [P][C]:
// out[p] = sqrt(sum_c |x[p][c]|^2), accumulated in coil order in float
// (reductions.hpp:973-983 then eigen_stats_common.hpp:376-381).
void rss_coil_last(const float2* d_in, float* d_out, uint64_t pixels, uint32_t coils, hipStream_t stream);

// Coil-major planes [C][P] -> optional coil-last [P][C] and optional RSS [P],
// staged through LDS so both the reads and the writes are contiguous.
void coil_major_to_last(const float2* d_in, float2* d_coil_last, float* d_rss, uint64_t pixels, uint32_t coils,
                        hipStream_t stream);

constexpr uint32_t kMaxCoilGroups = 64;

// Outputs of coil_major_split: images stacked along the coil axis in groups of
// `coils` planes; group g writes coil-last images and/or its RSS image (null skips).
struct CoilGroupOutputs {
  float2* coil_last[kMaxCoilGroups];
  float* rss[kMaxCoilGroups];
  uint32_t groups = 0;
  uint32_t coils = 0;
};

// Coil-major planes [groups * coils][P] -> per group coil-last [P][coils] and/or RSS [P].
void coil_major_split(const float2* d_in, const CoilGroupOutputs& out, uint64_t pixels, hipStream_t stream);

// Items of an uneven batch, one launch: item i's coil-major [coils[i]][pixels[i]]
// planes -> coil-last and/or RSS (either may be null), as coil_major_split with one
// group.  The argument table goes to d_table (coil_split_fused_table_bytes() long).
size_t coil_split_fused_table_bytes();
void coil_split_fused(const float2* const* d_in, float2* const* d_coil_last, float* const* d_rss,
                      const uint64_t* pixels, const uint32_t* coils, uint32_t count, void* d_table,
                      hipStream_t stream);

// x[p] = sum_c conj(S[p][c]) I[p][c] / sum_c |S[p][c]|^2, 0 where the denominator is 0.
void sense_combine(const float2* d_images, const float2* d_sens, float2* d_out, uint64_t pixels, uint32_t coils,
                   hipStream_t stream);

}  // namespace ksj::gpu
