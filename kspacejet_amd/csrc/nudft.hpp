// Exact non-uniform DFT kernels.
#pragma once

#include "common.hpp"

namespace ksj::gpu {

// Coil-major output image[c][rows][cols] (float arithmetic, reference semantics).
void nudft2_adjoint(const float* d_traj, const float2* d_samples, float2* d_image, uint32_t samples,
                    uint32_t channels, uint32_t rows, uint32_t cols, double row_origin, double col_origin,
                    hipStream_t stream);

// FP64 forward model y[c][s] from image[c][rows][cols] (complex128), trajectory float32 [S][2].
void nudft2_forward_f64(const float* d_traj, const double2* d_image, double2* d_out, uint32_t samples,
                        uint32_t channels, uint32_t rows, uint32_t cols, double row_origin, double col_origin,
                        hipStream_t stream);

}  // namespace ksj::gpu
