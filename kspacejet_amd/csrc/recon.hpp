// Reconstruction pipelines assembled from the kernels: Cartesian multi-coil
// inverse FFT and the 2-D adjoint NUFFT (gridding -> pruned 2-D inverse FFT ->
// deapodize/crop -> coil layout / RSS).
#pragma once

#include "plan.hpp"

namespace ksj::gpu {

enum : int { kLayoutCoilLast = 0, kLayoutCoilMajor = 1 };

size_t ifft2_workspace_bytes(uint32_t channels, uint32_t rows, uint32_t cols);
// Channel-major k-space [C][R][Cols] -> images, 1/cols then 1/rows scaling
// (Normalization::inverse), uncentred or centred (fftshift(ifft2(ifftshift))).
void ifft2_coil_images(const float2* d_kspace, float2* d_images, uint32_t channels, uint32_t rows, uint32_t cols,
                       bool centred, int layout, void* d_ws, size_t ws_bytes, hipStream_t stream);

size_t adjoint2_workspace_bytes(const Plan2& plan, uint32_t channels);
// Channel-major samples [C][S] -> coil images (layout) and/or RSS image [N_r][N_c].
void adjoint2_execute(const Plan2& plan, const float2* d_kspace, uint32_t channels, float2* d_images, int layout,
                      float* d_rss, void* d_ws, size_t ws_bytes, hipStream_t stream);

size_t nudft2_adjoint_workspace_bytes(uint32_t channels, uint32_t rows, uint32_t cols);

}  // namespace ksj::gpu
