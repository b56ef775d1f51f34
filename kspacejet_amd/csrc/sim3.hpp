// FP64 forward NUFFT of a 3-D volume (data simulation).
#pragma once

#include "common.hpp"

namespace ksj::gpu {

// Workspace of nufft3_forward_f64: the 2x-oversampled complex128 grid.
size_t nufft3_forward_f64_workspace_bytes(const uint32_t image[3]);

// y_s ~ sum_v x_v exp(-i w_s . (v - origin)) for a volume x [Nz][Ny][Nx] (complex128) and
// a trajectory [S][3] (z, y, x) in rad/px, by a Kaiser-Bessel NUFFT of width 4, 6 or 8 on
// the 2x grid (each side twice the volume's, a power of two), all in FP64; y [S] complex128.
void nufft3_forward_f64(const double2* d_image, const uint32_t image[3], const double origin[3], const float* d_traj,
                        uint64_t samples, int width, double2* d_out, void* d_ws, size_t ws_bytes,
                        hipStream_t stream);

}  // namespace ksj::gpu
