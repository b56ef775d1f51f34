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

// One scan of an adjoint call: channel-major samples [C][S] in, coil images
// (layout) and/or RSS image [N_r][N_c] out (null skips an output).
struct AdjointScan {
  const float2* kspace = nullptr;
  float2* images = nullptr;
  float* rss = nullptr;
};

// Scans on one plan are stacked along the coil axis, up to 64 coils per pass.
uint32_t adjoint2_scans_per_pass(uint32_t channels);
// Workspace of a whole call: one pass workspace per overlapping pass lane.
size_t adjoint2_workspace_bytes(const Plan2& plan, uint32_t channels, uint32_t scans = 1);
size_t adjoint2_pass_workspace_bytes(const Plan2& plan, uint32_t channels, uint32_t scans = 1);
void adjoint2_execute(const Plan2& plan, const AdjointScan* scans, uint32_t count, uint32_t channels, int layout,
                      void* d_ws, size_t ws_bytes, hipStream_t stream);

// One item of a multi-plan call: its own plan (geometry and trajectory), channels and outputs.
struct MultiItem {
  const Plan2* plan = nullptr;
  uint32_t channels = 0;
  AdjointScan scan;
};

// Items of different geometry, coil count and trajectory: each item's pipeline runs on
// one of `multi_lanes` streams (largest items first, round-robin), the lanes joined to
// `stream`.  Enqueue-only (no host synchronisation, no allocation once the FFT twiddles
// of every grid length exist), so the call can be captured into a HIP graph.
size_t adjoint2_multi_workspace_bytes(const MultiItem* items, uint32_t count);
void adjoint2_execute_multi(const MultiItem* items, uint32_t count, int layout, void* d_ws, size_t ws_bytes,
                            hipStream_t stream);

size_t nudft2_adjoint_workspace_bytes(uint32_t channels, uint32_t rows, uint32_t cols);

}  // namespace ksj::gpu
