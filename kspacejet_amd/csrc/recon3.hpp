// 3-D KB adjoint NUFFT pipeline over a resident Plan3 (recon3.hip).
#pragma once

#include "plan3.hpp"

namespace ksj::gpu {

// Workspace bytes of adjoint3_execute for `channels` coils taken `coils_per_pass` at
// a time (1..64; the grid of one pass holds G^3 x padded(coils_per_pass) values).
size_t adjoint3_workspace_bytes(const Plan3& plan, uint32_t channels, uint32_t coils_per_pass);

// k-space [channels][S] complex64 (channel-major) -> coil images coil-last
// [N_z][N_y][N_x][channels] and/or the RSS volume [N_z][N_y][N_x] (float, summed in
// channel order).  Enqueued on `stream`; allocates nothing.
void adjoint3_execute(const Plan3& plan, const float2* d_kspace, uint32_t channels, uint32_t coils_per_pass,
                      float2* d_images, float* d_rss, void* d_workspace, size_t workspace_bytes, hipStream_t stream);

}  // namespace ksj::gpu
