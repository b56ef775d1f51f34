// Exact non-uniform DFTs: the FP32 direct adjoint of the reference's
// noncartesian_adjoint_reconstruct operator and an FP64 direct forward used to
// simulate k-space.  Both stage samples (or image rows) in LDS so every thread
// of a block reads the same element at the same time (LDS broadcast).
#include "nudft.hpp"

namespace ksj::gpu {
namespace {

constexpr int kChunk = 256;
constexpr int kAdjGroup = 8;

// image[c][r][col] = sum_s y[c][s] exp(+i(w_r (r - r0) + w_c (col - c0)))
// (eigen_nufft_direct_nudft.cpp:77-97, float arithmetic, no normalisation).
__global__ __launch_bounds__(256) void adjoint_kernel(const float* __restrict__ traj, const float2* __restrict__ y,
                                                      float2* __restrict__ image, uint32_t samples, uint32_t channels,
                                                      uint32_t rows, uint32_t cols, float r0, float c0) {
  __shared__ float2 tr[kChunk];
  __shared__ float2 val[kChunk][kAdjGroup];
  const uint64_t pixels = static_cast<uint64_t>(rows) * cols;
  const uint64_t p = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  const uint32_t g = blockIdx.y * kAdjGroup;
  const float rp = p < pixels ? static_cast<float>(p / cols) - r0 : 0.f;
  const float cp = p < pixels ? static_cast<float>(p % cols) - c0 : 0.f;
  float2 acc[kAdjGroup];
#pragma unroll
  for (int c = 0; c < kAdjGroup; ++c) acc[c] = make_float2(0.f, 0.f);
  for (uint32_t s0 = 0; s0 < samples; s0 += kChunk) {
    const uint32_t n = samples - s0 < kChunk ? samples - s0 : kChunk;
    for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) tr[i] = make_float2(traj[2 * (s0 + i)], traj[2 * (s0 + i) + 1]);
    for (uint32_t i = threadIdx.x; i < n * kAdjGroup; i += blockDim.x) {
      const uint32_t c = i / n, s = i % n;
      val[s][c] = g + c < channels ? y[static_cast<uint64_t>(g + c) * samples + s0 + s] : make_float2(0.f, 0.f);
    }
    __syncthreads();
    for (uint32_t s = 0; s < n; ++s) {
      const float phase = tr[s].x * rp + tr[s].y * cp;
      float sn, cs;
      sincosf(phase, &sn, &cs);
#pragma unroll
      for (int c = 0; c < kAdjGroup; ++c) {
        const float2 v = val[s][c];
        acc[c].x += v.x * cs - v.y * sn;
        acc[c].y += v.x * sn + v.y * cs;
      }
    }
    __syncthreads();
  }
  if (p >= pixels) return;
#pragma unroll
  for (int c = 0; c < kAdjGroup; ++c)
    if (g + c < channels) image[static_cast<uint64_t>(g + c) * pixels + p] = acc[c];
}

constexpr int kFwdGroup = 4;
constexpr int kFwdSeed = 16;

// y[c][s] = sum_{r,col} x[c][r][col] exp(-i(w_r (r - r0) + w_c (col - c0))) in FP64
// (eigen_nufft_direct_nudft.cpp:52-74).  The column phasor advances by a
// rotation re-seeded from sincos every kFwdSeed columns.
__global__ __launch_bounds__(256) void forward_f64_kernel(const float* __restrict__ traj,
                                                          const double2* __restrict__ image, double2* __restrict__ y,
                                                          uint32_t samples, uint32_t channels, uint32_t rows,
                                                          uint32_t cols, double r0, double c0) {
  extern __shared__ double2 row_vals[];  // [kFwdGroup][cols]
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t g = blockIdx.y * kFwdGroup;
  const double wr = s < samples ? static_cast<double>(traj[2 * s]) : 0.0;
  const double wc = s < samples ? static_cast<double>(traj[2 * s + 1]) : 0.0;
  double sstep, cstep;
  sincos(-wc, &sstep, &cstep);
  double2 acc[kFwdGroup];
#pragma unroll
  for (int c = 0; c < kFwdGroup; ++c) acc[c] = make_double2(0.0, 0.0);
  const uint64_t plane = static_cast<uint64_t>(rows) * cols;
  for (uint32_t r = 0; r < rows; ++r) {
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < kFwdGroup * cols; i += blockDim.x) {
      const uint32_t c = i / cols, col = i % cols;
      row_vals[i] = g + c < channels ? image[(g + c) * plane + static_cast<uint64_t>(r) * cols + col]
                                     : make_double2(0.0, 0.0);
    }
    __syncthreads();
    double sa, ca;
    sincos(-wr * (static_cast<double>(r) - r0), &sa, &ca);
    for (uint32_t cb = 0; cb < cols; cb += kFwdSeed) {
      double sb, cbv;
      sincos(-wc * (static_cast<double>(cb) - c0), &sb, &cbv);
      const uint32_t ce = cb + kFwdSeed < cols ? cb + kFwdSeed : cols;
      for (uint32_t col = cb; col < ce; ++col) {
        const double er = ca * cbv - sa * sb, ei = ca * sb + sa * cbv;
#pragma unroll
        for (int c = 0; c < kFwdGroup; ++c) {
          const double2 v = row_vals[c * cols + col];
          acc[c].x += v.x * er - v.y * ei;
          acc[c].y += v.x * ei + v.y * er;
        }
        const double nb = cbv * cstep - sb * sstep;
        sb = cbv * sstep + sb * cstep;
        cbv = nb;
      }
    }
  }
  if (s >= samples) return;
#pragma unroll
  for (int c = 0; c < kFwdGroup; ++c)
    if (g + c < channels) y[static_cast<uint64_t>(g + c) * samples + s] = acc[c];
}

}  // namespace

void nudft2_adjoint(const float* d_traj, const float2* d_samples, float2* d_image, uint32_t samples,
                    uint32_t channels, uint32_t rows, uint32_t cols, double row_origin, double col_origin,
                    hipStream_t stream) {
  require(rows >= 1 && cols >= 1 && channels >= 1, "direct adjoint needs a non-empty image");
  const uint64_t pixels = static_cast<uint64_t>(rows) * cols;
  dim3 grid(ceil_div(pixels, 256), ceil_div(channels, kAdjGroup));
  hipLaunchKernelGGL(adjoint_kernel, grid, dim3(256), 0, stream, d_traj, d_samples, d_image, samples, channels, rows,
                     cols, static_cast<float>(row_origin), static_cast<float>(col_origin));
  KSJ_LAUNCH_CHECK();
}

void nudft2_forward_f64(const float* d_traj, const double2* d_image, double2* d_out, uint32_t samples,
                        uint32_t channels, uint32_t rows, uint32_t cols, double row_origin, double col_origin,
                        hipStream_t stream) {
  require(rows >= 1 && cols >= 1 && channels >= 1 && samples >= 1, "direct forward needs non-empty inputs");
  require(static_cast<size_t>(kFwdGroup) * cols * sizeof(double2) <= 64 * 1024, "direct forward supports cols <= 1024");
  dim3 grid(ceil_div(samples, 256), ceil_div(channels, kFwdGroup));
  const size_t shmem = static_cast<size_t>(kFwdGroup) * cols * sizeof(double2);
  hipLaunchKernelGGL(forward_f64_kernel, grid, dim3(256), shmem, stream, d_traj, d_image, d_out, samples, channels,
                     rows, cols, row_origin, col_origin);
  KSJ_LAUNCH_CHECK();
}

}  // namespace ksj::gpu
