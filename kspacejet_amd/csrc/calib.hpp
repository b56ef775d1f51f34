// Calibration estimators on the GPU: channel covariance of channel-major samples, a
// Hermitian eigensolver, and the three estimates of
// providers/kspacejet-calibration/src/operators/ built from them (whitening noise model,
// coil-compression basis, readout phase correction).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace ksj::gpu {

constexpr uint32_t kCalibMaxChannels = 64;

enum CovarianceKind : int {
  kCovarianceNoise = 0,       // sum_s conj(x_r - m_r) (x_c - m_c) / S, m = per-channel mean (population)
  kCovarianceSecondMoment = 1 // sum_s x_r conj(x_c) / S (uncentred)
};

// Workspace of channel_covariance and of the estimates for `channels` x `samples`.
size_t calibration_workspace_bytes(uint32_t channels, uint64_t samples);

// Hermitian covariance [C][C] (double2, row-major) of x [C][S] (complex float).  Per-chunk
// FP32 products summed in FP64, in a fixed order.
void channel_covariance(const float2* d_x, uint32_t channels, uint64_t samples, int kind, double2* d_cov,
                        void* d_ws, size_t ws_bytes, hipStream_t stream);

// Eigen-decomposition of a Hermitian n x n matrix (double2, row-major, n <= 64) in FP64:
// eigenvalues ascending in d_values[n], eigenvector k in column k of d_vectors [n][n].
// *d_status = 0 on success, 1 when the matrix holds a non-finite value, 2 when the
// rotations did not converge.  One workgroup; asynchronous.
void hermitian_eigen(const double2* d_matrix, uint32_t n, double* d_values, double2* d_vectors, uint32_t* d_status,
                     hipStream_t stream);

// The estimates.  Each synchronises `stream` and throws Error(kErrArgument) when the
// eigen-decomposition fails (non-finite samples, no convergence).
// W = V diag(1 / sqrt(max(lambda, floor))) V^H of the noise covariance, complex float [C][C].
void noise_model_estimate(const float2* d_x, uint32_t channels, uint64_t samples, float eigenvalue_floor,
                          float2* d_whitening, void* d_ws, size_t ws_bytes, hipStream_t stream);
// B[v][p] = conj(eigenvector p of the v-th largest eigenvalue) of the second moment, [V][P].
void coil_compression_basis_estimate(const float2* d_x, uint32_t physical, uint32_t virtual_channels,
                                     uint64_t samples, float2* d_basis, void* d_ws, size_t ws_bytes,
                                     hipStream_t stream);
// p[c][k] = conj(s) / |s| with s = sum over lines of ref[c][line][k] (1 + 0i where s = 0).
void phase_correction_estimate(const float2* d_ref, uint32_t channels, uint32_t lines, uint32_t readout,
                               float2* d_phase, hipStream_t stream);

// Sensitivity maps by Walsh's adaptive method (Walsh, Gmitro & Marcellin, MRM 2000): at
// each pixel the dominant eigenvector of the coil covariance summed over a window x window
// neighbourhood (clipped at the edges), by power iteration from R e_0 (`iterations`
// normalised steps) in FP32, rotated so that coil 0 is real and >= 0; zero where the
// neighbourhood is zero.  Coil-last images and maps [rows][cols][coils].  Asynchronous.
void sensitivity_maps_walsh(const float2* d_images, uint32_t rows, uint32_t cols, uint32_t coils, uint32_t window,
                            uint32_t iterations, float2* d_maps, hipStream_t stream);

}  // namespace ksj::gpu
