// K-space conditioning ahead of reconstruction: channel mixing (noise prewhitening and
// coil compression), per-channel readout phase correction and readout crop of
// channel-major k-space (providers/kspacejet-kspace-conditioning/src/operators/).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace ksj::gpu {

constexpr uint32_t kMixMaxChannels = 64;

// y[o][s] = sum_i w[o][i] x[i][s]: w row-major [out][in], x and y channel-major
// [channel][samples], complex float.  FP32 fused multiply-adds in input-channel order.
// x and y must not overlap.
void channel_mix(const float2* d_w, const float2* d_x, float2* d_y, uint32_t out_channels, uint32_t in_channels,
                 uint64_t samples, hipStream_t stream);

// y[c][r][k] = x[c][r][k] * p[c][k], each product and sum rounded to float (no
// contraction), the rounding of std::complex<float> multiplication of finite values.
void phase_correct(const float2* d_x, const float2* d_p, float2* d_y, uint32_t channels, uint32_t rows,
                   uint32_t cols, hipStream_t stream);

// y[l][k] = x[l][offset + k] for `lines` lines of in_cols values.
void readout_crop(const float2* d_x, float2* d_y, uint64_t lines, uint32_t in_cols, uint32_t out_cols,
                  uint32_t offset, hipStream_t stream);

}  // namespace ksj::gpu
