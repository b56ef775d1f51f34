/*
 * ksj_gpu.h -- C ABI of the MI355X (gfx950) reconstruction library libksj_gpu.so.
 *
 * Plain pointers and sizes only.  Every d_* argument is device memory on the
 * current HIP device; `stream` is a hipStream_t (NULL = the default stream).
 * All work is enqueued on `stream`; a call returns when the work is enqueued,
 * except where a function states that it synchronises.
 *
 * Complex data is interleaved float32 (re, im), i.e. the reference's cf32 /
 * complex_float32 payload encoding.  Layouts follow the reference type registry
 * (types/registry.json):
 *   k-space (Cartesian)   [channel][ky][kx]        channel-major   (ksj.kspace-frame)
 *   k-space (non-Cart.)   [channel][sample]        channel-major   (ksj.noncartesian-kspace-frame)
 *   trajectory            [sample][ky, kx] float32 rad/pixel       (ksj.trajectory-frame)
 *   coil images           [ky][kx][channel]        coil-last       (ksj.coil-image-frame)
 *   image                 [ky][kx] float32                         (ksj.image-frame)
 * Reference citations are relative to the KSpaceJet source tree.
 *
 * Every int-returning function returns KSJ_GPU_OK or an error code, and on error
 * ksj_gpu_last_error() describes it (per calling thread).  No function throws.
 */
#ifndef KSJ_GPU_H_
#define KSJ_GPU_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define KSJ_GPU_OK 0
#define KSJ_GPU_INVALID_ARGUMENT 1
#define KSJ_GPU_UNSUPPORTED 2
#define KSJ_GPU_DEVICE_ERROR 3
#define KSJ_GPU_RESOURCE_EXHAUSTED 4

#define KSJ_GPU_LAYOUT_COIL_LAST 0  /* [ky][kx][channel], the reference's coil-image-frame */
#define KSJ_GPU_LAYOUT_COIL_MAJOR 1 /* [channel][ky][kx] */

#define KSJ_GPU_KERNEL_LINEAR 0        /* 2x2 bilinear (the reference radial route) */
#define KSJ_GPU_KERNEL_KAISER_BESSEL 1

#define KSJ_GPU_DENSITY_NONE 0
#define KSJ_GPU_DENSITY_RAMP 1  /* hypot(k_row, k_col), radial_analytic_ramp */
#define KSJ_GPU_DENSITY_GIVEN 2

/* ------------------------------------------------------------------ library */

/* Selects the HIP device for the calling thread and warms the runtime.
 * No reference counterpart (the reference is CPU-only); the Provider calls it
 * from execution_context_create with ksj_execution_context_descriptor.device_ordinal
 * (libs/recon/kspacejet-provider-sdk/include/kspacejet/provider/provider.h:363-371). */
int ksj_gpu_init(int device_ordinal);
const char* ksj_gpu_last_error(void);
/* Waits for `stream`. */
int ksj_gpu_stream_synchronize(void* stream);

/* ------------------------------------------------------------------ Cartesian */

/* Multi-coil 2-D inverse FFT, 1/cols then 1/rows (Normalization::inverse).
 * centred = 0: uncentred, as the reference
 *   (providers/kspacejet-cartesian-recon/src/operators/cartesian_ifft2_coil_images.cpp:375-412
 *    over ksj::fft::fft2_inplace_with_workspace,
 *    libs/numerics/kspacejet-fft/include/kspacejet/fft/transforms.hpp:1003-1009);
 * centred = 1: fftshift(ifft2(ifftshift(k))) per coil.
 * Replaces the per-channel copy + fft2_inplace_with_workspace + strided coil-last store. */
size_t ksj_gpu_ifft2_workspace_bytes(uint32_t channels, uint32_t rows, uint32_t cols);
int ksj_gpu_ifft2_coil_images(const void* d_kspace, void* d_images, uint32_t channels, uint32_t rows, uint32_t cols,
                              int centred, int layout, void* d_workspace, size_t workspace_bytes, void* stream);

/* ------------------------------------------------------------------ coil combine */

/* out[p] = sqrt(sum_c |x[p][c]|^2) over coil-last images.  Replaces
 * ksj::stats::root_sum_of_squares_across(CubeView, MatrixView, Dim::dim2)
 * (libs/numerics/kspacejet-stats/include/kspacejet/stats/norms.hpp:207-219) as called by
 * providers/kspacejet-coil-combine/src/operators/coil_combine_rss.cpp:308. */
int ksj_gpu_rss(const void* d_coil_images, float* d_image, uint64_t pixels, uint32_t channels, void* stream);

/* Sensitivity-weighted combine x = sum_c conj(S_c) I_c / sum_c |S_c|^2 (complex output).
 * The reference only plans this operator
 * (providers/interfaces/org.kspacejet.coil-combine/coil_combine_adaptive.json: coil_images + sensitivity_maps
 *  -> ksj.complex-image-frame). */
int ksj_gpu_sense_combine(const void* d_coil_images, const void* d_sensitivities, void* d_image, uint64_t pixels,
                          uint32_t channels, void* stream);

/* ------------------------------------------------------------------ non-Cartesian */

/* Analytic ramp DCF hypot(k_row, k_col) in float.  Replaces ksj::nufft::radial_analytic_ramp_dcf2
 * (libs/numerics/kspacejet-nufft/include/kspacejet/nufft/radial_gridding.hpp:33). */
int ksj_gpu_ramp_dcf2(const float* d_traj, float* d_dcf, uint64_t samples, void* stream);

/* Exact adjoint NUDFT x(r,c) = sum_s y_s exp(+i(w_r (r - r0) + w_c (c - c0))) per channel, float.
 * Replaces the per-channel ksj::nufft::direct_nudft2_adjoint loop of
 * providers/kspacejet-noncartesian-recon/src/operators/noncartesian_adjoint_reconstruct.cpp:346-368
 * (libs/numerics/kspacejet-nufft/src/eigen/eigen_nufft_direct_nudft.cpp:77-97). */
size_t ksj_gpu_nudft2_adjoint_workspace_bytes(uint32_t channels, uint32_t rows, uint32_t cols);
int ksj_gpu_nudft2_adjoint(const float* d_traj, const void* d_samples, void* d_images, uint32_t channels,
                           uint32_t samples, uint32_t rows, uint32_t cols, double row_origin, double col_origin,
                           int layout, void* d_workspace, size_t workspace_bytes, void* stream);

/* Exact forward NUDFT in FP64 (complex128 in and out), for simulating k-space.
 * Counterpart of ksj::nufft::direct_nudft2_forward
 * (libs/numerics/kspacejet-nufft/src/eigen/eigen_nufft_direct_nudft.cpp:52-74). */
int ksj_gpu_nudft2_forward_f64(const float* d_traj, const void* d_image, void* d_samples, uint32_t channels,
                               uint32_t samples, uint32_t rows, uint32_t cols, double row_origin, double col_origin,
                               void* stream);

/* Golden-angle radial trajectory, FP64 on the device, rounded to float [S][ky, kx] rad/pixel:
 * spoke p at angle p * pi / phi (111.246 deg), k_j = (j - M/2) / M cycles per pixel.  The
 * reference receives trajectories from ISMRMRD files and has no generator. */
int ksj_gpu_golden_angle2(float* d_traj, uint32_t spokes, uint32_t readout, void* stream);

/* ------------------------------------------------------------------ gridding plans */

typedef struct ksj_gpu_plan2_desc {
  uint32_t struct_size;     /* sizeof(ksj_gpu_plan2_desc) */
  uint32_t image_rows, image_cols;
  uint32_t grid_rows, grid_cols;   /* oversampled grid (= image for the linear route) */
  uint32_t kernel;                 /* KSJ_GPU_KERNEL_* */
  uint32_t kernel_width;           /* 2 (linear) or 4/6/8 (Kaiser-Bessel) */
  uint32_t density;                /* KSJ_GPU_DENSITY_* */
  uint32_t centred;                /* 1: crop the centre of the oversampled grid; 0: unshifted (linear route) */
  uint32_t samples;
  uint32_t item_visits;            /* gather work-item size in visits, 0 = the schedule's tuning key
                                      ("matrix_item_visits" or "item_visits") */
  uint32_t reserved0;
  double beta;                     /* <= 0: Beatty et al. 2005 value for the oversampling */
  double row_origin, col_origin;   /* image origin; the reference uses ((R-1)/2, (C-1)/2) */
} ksj_gpu_plan2_desc;

typedef struct ksj_gpu_plan2_info {
  uint32_t struct_size;
  uint32_t items;          /* gather work items */
  uint64_t visits;         /* (sample, 2x2 cell block) visits */
  uint64_t device_bytes;   /* resident table bytes */
  double beta;
} ksj_gpu_plan2_info;

typedef struct ksj_gpu_plan2 ksj_gpu_plan2;

/* A plan holds everything derived from one trajectory: FP64 grid coordinates,
 * the stable sample-to-cell sort, density x origin-phase weights, kernel taps and
 * deapodization.  The grid-coordinate rule is that of
 * libs/numerics/kspacejet-nufft/src/eigen/eigen_radial_gridding.cpp:262-277; with
 * kernel = LINEAR, grid = image and centred = 0 the plan reproduces
 * ksj::nufft::radial_linear_gridding2_adjoint
 * (libs/numerics/kspacejet-nufft/include/kspacejet/nufft/radial_gridding.hpp:59). */
int ksj_gpu_plan2_create(const ksj_gpu_plan2_desc* desc, ksj_gpu_plan2** out_plan);
/* Builds the tables for a trajectory (d_dcf read only for DENSITY_GIVEN).  Synchronises
 * `stream`.  Fails with INVALID_ARGUMENT on a non-finite sample or one outside [-pi, pi]. */
int ksj_gpu_plan2_set_trajectory(ksj_gpu_plan2* plan, const float* d_traj, const float* d_dcf, void* stream);
int ksj_gpu_plan2_get_info(const ksj_gpu_plan2* plan, ksj_gpu_plan2_info* out_info);
/* Copies the sort result: d_cell_start[grid_rows*grid_cols + 1] and d_perm[samples]
 * (sorted position -> sample index), for bit-exact checks of the binning. */
int ksj_gpu_plan2_export_bins(const ksj_gpu_plan2* plan, uint32_t* d_cell_start, uint32_t* d_perm, void* stream);
void ksj_gpu_plan2_destroy(ksj_gpu_plan2* plan);

/* Adjoint NUFFT of channel-major samples [C][S]: weight, gather onto the grid,
 * pruned inverse 2-D FFT, deapodize and crop; writes coil images in `layout`
 * (if d_images) and/or the RSS image (if d_rss).  For the KB plan this is the
 * reconstruction that replaces the bilinear route of
 * providers/kspacejet-noncartesian-recon/src/operators/radial_gridding_reconstruct.cpp:391-413. */
size_t ksj_gpu_adjoint2_workspace_bytes(const ksj_gpu_plan2* plan, uint32_t channels);
int ksj_gpu_adjoint2_execute(const ksj_gpu_plan2* plan, const void* d_kspace, uint32_t channels, void* d_images,
                             int layout, float* d_rss, void* d_workspace, size_t workspace_bytes, void* stream);

/* The same adjoint for callers that cannot allocate device memory: code under HIP graph
 * capture, tight serving loops, hosts that forbid hipMalloc.  No counterpart in the
 * reference (its operators run on host memory); it exists for those callers only.
 * The workspace is a file-scope device array of KSJ_GPU_STATIC_SCRATCH_BYTES (128 MiB) per
 * GPU, held until the process ends; the call makes no hipMalloc/hipFree and no other
 * allocating call (the plan's FFT tables are made when its trajectory is set).  It fails
 * with KSJ_GPU_RESOURCE_EXHAUSTED, before enqueuing anything, when
 * ksj_gpu_adjoint2_workspace_bytes(plan, channels) exceeds the array (the 256x256,
 * 16-channel, 205,824-sample scan needs about 90 MB; 64 channels at that shape do not fit).
 * Every call on one GPU uses that GPU's one array: calls that overlap on one GPU (another
 * stream or thread before the previous call's work has finished) race and must be
 * serialised by the caller.  The array keeps the last call's contents; no kernel relies on
 * it starting at zero.  Results are bit-identical to ksj_gpu_adjoint2_execute. */
#define KSJ_GPU_STATIC_SCRATCH_BYTES (128u * 1024u * 1024u)
int ksj_gpu_adjoint2_execute_static(const ksj_gpu_plan2* plan, const void* d_kspace, uint32_t channels,
                                    void* d_images, int layout, float* d_rss, void* stream);

/* Batched adjoint of `count` scans that share the plan's trajectory: the scans are
 * stacked along the coil axis (up to 64 coils per gridding pass), so one pass of
 * gather/FFT/deapodize kernels serves several scans.  d_kspace[i] is scan i's
 * [C][S] samples; d_images[i] / d_rss[i] its outputs (either array may be NULL,
 * or hold NULL entries; every scan needs at least one output).  Each scan's result
 * is the same as its single-scan ksj_gpu_adjoint2_execute result up to the float
 * summation order of the shared passes.  No reference counterpart: the reference
 * processes one frame per firing (radial_gridding_reconstruct.cpp:400-413). */
size_t ksj_gpu_adjoint2_batch_workspace_bytes(const ksj_gpu_plan2* plan, uint32_t channels, uint32_t scans);
int ksj_gpu_adjoint2_execute_batch(const ksj_gpu_plan2* plan, uint32_t count, const void* const* d_kspace,
                                   uint32_t channels, void* const* d_images, int layout, float* const* d_rss,
                                   void* d_workspace, size_t workspace_bytes, void* stream);

/* Uneven batch: `count` scans, each with its own plan (geometry, kernel, trajectory)
 * and channel count.  Two forms with bit-identical results (tuning key "multi_fused"):
 *   0: every item's pipeline runs on one of "multi_lanes" streams (largest items first,
 *      round-robin), joined back to `stream`;
 *   1: fused, for up to 64 items of the 2x2-block or matrix schedule with <= 64 channels:
 *      each stage is one launch per shape class (padded coil count for gridding, length
 *      for each FFT pass) over all items, with per-item argument tables written into the
 *      workspace by small argument-carrying kernels;
 *   2: as 1, with prep, gather and reduce of every item (all padded coil counts) in one
 *      persistent launch fed by a device work queue per XCD (matrix-schedule plans with
 *      packed visits and matrix_batch 16; other batches take form 1).  Same bits as 0 and 1.
 * The call only enqueues work (no host synchronisation, no allocation once every grid
 * length has been run once), so it can be captured with ksj_gpu_graph_begin/end and
 * replayed.  The workspace size covers both forms.
 * Item i writes d_images[i] (layout) and/or d_rss[i]; either array may be NULL or
 * hold NULL entries.  No reference counterpart: the reference runs one frame per firing. */
size_t ksj_gpu_adjoint2_multi_workspace_bytes(uint32_t count, const ksj_gpu_plan2* const* plans,
                                              const uint32_t* channels);
int ksj_gpu_adjoint2_execute_multi(uint32_t count, const ksj_gpu_plan2* const* plans, const void* const* d_kspace,
                                   const uint32_t* channels, void* const* d_images, int layout, float* const* d_rss,
                                   void* d_workspace, size_t workspace_bytes, void* stream);
/* Launches of the multi_fused = 2 work queue on the current device whose bounded waits
 * ran out (each such launch stops early and its results are undefined); 0 in correct
 * operation.  The waits are bounded so that a broken queue cannot hang the GPU. */
int ksj_gpu_chain_faults(uint64_t* out_count);

/* Route trajectory normalisation: raw ISMRMRD-order [S][kx, ky] -> the Provider payload
 * order [S][ky, kx] in rad/px, each axis normalised with its own encoded dimension, float
 * arithmetic as the reference.  Replaces normalize_radial_coordinate(_pair) of
 * libs/recon/kspacejet-recon-runtime/src/noncartesian_rss_hdf5.cpp:776-834 (same units,
 * same inclusive Nyquist bounds).  Returns KSJ_GPU_INVALID_ARGUMENT with the first failing
 * coordinate's reason (sample order, kx before ky).  Synchronises `stream`. */
#define KSJ_GPU_TRAJ_UNITS_UNSPECIFIED 0
#define KSJ_GPU_TRAJ_UNITS_CYCLES_PER_FOV 1
#define KSJ_GPU_TRAJ_UNITS_RADIANS_PER_PIXEL 2
#define KSJ_GPU_TRAJ_UNITS_ENCODED_MATRIX_INDEX 3
int ksj_gpu_normalize_trajectory2(const float* d_raw, uint64_t samples, int units, uint32_t encoded_rows,
                                  uint32_t encoded_cols, float* d_out, void* stream);

/* ISMRMRD acquisitions -> the k-space frame of a non-Cartesian route.  d_acquisitions
 * holds the acquisitions one after another, each in the ISMRMRD sample layout
 * samples[sample + channel * ns_a] (complex float; libs/io/kspacejet-ismrmrd/include/
 * kspacejet/ismrmrd/dataset_reader.hpp:57) with ns_a = samples_per_acquisition[a] (host
 * array); d_kspace receives the channel-major frame [channels][S], S = sum of ns_a, acquisition
 * a at columns [o_a, o_a + ns_a) in acquisition order.  Replaces the per-channel copy loop of
 * the route's replay (libs/recon/kspacejet-recon-runtime/src/noncartesian_rss_hdf5.cpp:
 * 1077-1083) with its checks (:880-884): every acquisition has samples, channels in [1, 64],
 * S below 2^32.  The trajectory needs no reordering: ISMRMRD's [sample][dims] rows of the
 * acquisitions, concatenated, are the raw [S][kx, ky] input of
 * ksj_gpu_normalize_trajectory2.  When every acquisition has the same count the call is
 * asynchronous on `stream` and needs no workspace; otherwise the acquisition offsets go to
 * d_workspace (at least ksj_gpu_assemble_workspace_bytes, 4-byte aligned) and `stream` is
 * synchronised.  No reference entry point takes device memory: the reference assembles the
 * frame in host memory while the file streams. */
size_t ksj_gpu_assemble_workspace_bytes(uint32_t acquisitions);
int ksj_gpu_assemble_acquisitions(const void* d_acquisitions, const uint32_t* samples_per_acquisition,
                                  uint32_t acquisitions, uint32_t channels, void* d_kspace, void* d_workspace,
                                  size_t workspace_bytes, void* stream);

/* Trajectory check on the device, for hosts that hand trajectories over in HBM.  Writes to
 * the device word *d_flags: KSJ_GPU_TRAJ_DIFFERS when some of the `values` floats at d_traj
 * differ bitwise from those at d_resident (always when d_resident is NULL), and
 * KSJ_GPU_TRAJ_INVALID when some value is not finite or lies outside [-bound, bound]
 * (bound = FLT_MAX: finite only).  The value rule is the host check of the reference's
 * non-Cartesian operators (providers/kspacejet-noncartesian-recon/src/operators/
 * radial_gridding_reconstruct.cpp:219-221, finite and within [-pi, pi];
 * noncartesian_adjoint_reconstruct.cpp:300, finite), moved to where the payload lives.
 * No reference entry point: the reference takes host payloads only.  Asynchronous on
 * `stream`; no memset, so it may be captured. */
#define KSJ_GPU_TRAJ_DIFFERS 1u
#define KSJ_GPU_TRAJ_INVALID 2u
int ksj_gpu_trajectory_check(const float* d_traj, const float* d_resident, uint64_t values, float bound,
                             uint32_t* d_flags, void* stream);

/* ------------------------------------------------------------------ k-space conditioning */

/* Channel mixing of channel-major k-space: y[o][s] = sum_i m[o][i] x[i][s], complex float,
 * m row-major [out_channels][in_channels], x [in_channels][samples], y [out_channels][samples]
 * (samples = rows * cols for a Cartesian frame), 1 <= channels <= 64, FP32 accumulation in
 * input-channel order; y must not overlap x.  Replaces the ksj::linalg::matmul(W, X, Y) of
 * providers/kspacejet-kspace-conditioning/src/operators/noise_prewhiten.cpp:141 and
 * noncartesian_noise_prewhiten.cpp:133 (m = the ksj.noise-model whitening matrix) and of
 * coil_compress.cpp:159 and noncartesian_coil_compress.cpp:148 (m = the
 * ksj.coil-compression-basis). */
int ksj_gpu_channel_mix(const void* d_matrix, const void* d_kspace, void* d_output, uint32_t out_channels,
                        uint32_t in_channels, uint64_t samples, void* stream);

/* Per-channel readout phase: y[c][r][k] = x[c][r][k] * p[c][k], every product and sum
 * rounded to float (bit-identical to std::complex<float> multiplication of finite values).
 * Replaces the transform_indexed of phase_correct.cpp:140-142 (p = the ksj.phase-model). */
int ksj_gpu_phase_correct(const void* d_kspace, const void* d_phase, void* d_output, uint32_t channels,
                          uint32_t rows, uint32_t cols, void* stream);

/* Readout crop: y[l][k] = x[l][offset + k] for `lines` = channels * rows lines of in_cols
 * values, offset + out_cols <= in_cols.  Replaces the per-line std::copy_n of
 * readout_oversampling_remove.cpp:174-180. */
int ksj_gpu_readout_crop(const void* d_kspace, void* d_output, uint64_t lines, uint32_t in_cols, uint32_t out_cols,
                         uint32_t offset, void* stream);

/* ------------------------------------------------------------------ calibration */

/* Estimates of providers/kspacejet-calibration/src/operators/ from channel-major complex
 * float samples x[C][S] (C <= 64).  The covariance is summed on the GPU (FP32 products per
 * chunk of 64 samples, FP64 sums in a fixed order) and decomposed in FP64 by a cyclic
 * Jacobi solver on one workgroup; outputs are rounded to float once.  The estimate calls
 * synchronise `stream` and fail with KSJ_GPU_INVALID_ARGUMENT when a sample is not finite
 * or the decomposition does not converge (the reference reports the same cases as "could
 * not estimate").  Workspace: ksj_gpu_calibration_workspace_bytes(C, S). */
size_t ksj_gpu_calibration_workspace_bytes(uint32_t channels, uint64_t samples);

/* Hermitian covariance [C][C] (complex128, row-major).  centred = 1: the mean-centred
 * population covariance sum_s conj(x_r - m_r)(x_c - m_c) / S of
 * ksj::stats::covariance_channel_major_with_workspace
 * (libs/numerics/kspacejet-stats/include/kspacejet/stats/moments.hpp:99-140, as called with
 * VarianceNormalization::population by noise_model_estimate.cpp:182); centred = 0: the
 * uncentred sum_s x_r conj(x_c) / S of coil_compression_basis_estimate.cpp:207-227.
 * Asynchronous. */
int ksj_gpu_channel_covariance(const void* d_samples, uint32_t channels, uint64_t samples, int centred,
                               void* d_covariance, void* d_workspace, size_t workspace_bytes, void* stream);

/* Eigen-decomposition of a Hermitian n x n complex128 matrix (n <= 64): eigenvalues
 * ascending (float64) and eigenvector k in column k of d_vectors (complex128 [n][n]).
 * *d_status: 0 = converged, 1 = non-finite input, 2 = no convergence.  Replaces
 * ksj::linalg::self_adjoint_eigen_decomposition_with_workspace
 * (libs/numerics/kspacejet-linalg/include/kspacejet/linalg/decompositions.hpp:503-626): the
 * same rotation per pivot, applied in the cyclic round-robin order rather than largest
 * pivot first, so each eigenvector agrees with the reference's up to a unit phase.
 * Asynchronous. */
int ksj_gpu_hermitian_eigen(const void* d_matrix, uint32_t n, double* d_values, void* d_vectors, uint32_t* d_status,
                            void* stream);

/* Noise model: W = V diag(1 / sqrt(max(lambda, floor))) V^H of the centred covariance,
 * complex float [C][C] row-major, applied as y = W x.  Replaces noise_model_estimate.cpp:158-203
 * (floor 1e-6 there) over ksj::linalg::whitening_matrix_from_self_adjoint_eigen_with_workspace
 * (libs/numerics/kspacejet-linalg/include/kspacejet/linalg/whitening.hpp:24-50).  S >= 2. */
int ksj_gpu_noise_model_estimate(const void* d_samples, uint32_t channels, uint64_t samples, float eigenvalue_floor,
                                 void* d_whitening, void* d_workspace, size_t workspace_bytes, void* stream);

/* Coil-compression basis B[v][p] = conj(V[p][P-1-v]) of the uncentred covariance (the
 * conjugated eigenvectors of the largest eigenvalues), complex float [V][P] row-major,
 * applied as y = B x.  Replaces coil_compression_basis_estimate.cpp:177-244. */
int ksj_gpu_coil_compression_basis_estimate(const void* d_samples, uint32_t physical_channels,
                                            uint32_t virtual_channels, uint64_t samples, void* d_basis,
                                            void* d_workspace, size_t workspace_bytes, void* stream);

/* Sensitivity maps [rows][cols][channels] (ksj.sensitivity-map) from coil-last coil images
 * (ksj.coil-image-frame) by Walsh's adaptive method: at each pixel the dominant eigenvector
 * of the coil covariance summed over a window x window neighbourhood (odd, <= 31, clipped
 * at the image edges), by `iterations` (<= 64) normalised power-iteration steps from
 * R e_0 in FP32, rotated so that coil 0 is real and >= 0; zero where the neighbourhood is
 * zero.  For the planned interface providers/interfaces/org.kspacejet.calibration/
 * sensitivity_map_estimate.json (no reference implementation); its maps feed
 * ksj_gpu_sense_combine (coil_combine_adaptive.json).  Asynchronous. */
int ksj_gpu_sensitivity_maps_walsh(const void* d_coil_images, uint32_t rows, uint32_t cols, uint32_t channels,
                                   uint32_t window, uint32_t iterations, void* d_maps, void* stream);

/* Readout phase model p[c][k] = conj(s) / |s|, s = sum over reference lines of
 * ref[c][line][k] in float in line order (1 + 0i where s = 0), |s| = the FP64 square root
 * of re^2 + im^2 rounded to float.  Replaces phase_correction_estimate.cpp:141-163.
 * Asynchronous. */
int ksj_gpu_phase_correction_estimate(const void* d_reference, uint32_t channels, uint32_t lines, uint32_t readout,
                                      void* d_phase, void* stream);

/* ------------------------------------------------------------------ 3-D (kooshball) */

/* 3-D KB gridding adjoint.  No reference counterpart: the reference reconstructs 2-D
 * frames only (ksj::nufft::Grid2D, include/kspacejet/nufft/types.hpp:9-14); these
 * extend its 2-D definitions axis by axis.  Volumes are [z][y][x]; trajectories
 * [S][3] float32 in (z, y, x) order, radians per pixel; origin per axis (the 2-D
 * providers' rule is ((N-1)/2) per axis). */
#define KSJ_GPU_DENSITY_RAMP_SQUARED 3 /* |k|^2, the 3-D radial density compensation */

typedef struct ksj_gpu_plan3_desc {
  uint32_t struct_size;          /* sizeof(ksj_gpu_plan3_desc) */
  uint32_t image_z, image_y, image_x;
  uint32_t grid_z, grid_y, grid_x; /* oversampled grid, each <= 2048, fewer than 2^32 points */
  uint32_t kernel_width;         /* 2, 4, 6 or 8 (Kaiser-Bessel) */
  uint32_t density;              /* KSJ_GPU_DENSITY_NONE, _RAMP_SQUARED or _GIVEN */
  uint32_t samples;
  uint32_t item_visits;          /* gather work-item size in visits, 0 = the schedule's tuning key
                                    ("matrix3_item_visits" or "item_visits") */
  uint32_t reserved0;
  double beta;                   /* <= 0: Beatty et al. 2005 value for the oversampling */
  double origin_z, origin_y, origin_x;
} ksj_gpu_plan3_desc;

typedef struct ksj_gpu_plan3_info {
  uint32_t struct_size;
  uint32_t items;          /* gather work items */
  uint64_t visits;         /* (sample, 2x2x2 cell block) visits */
  uint64_t device_bytes;   /* resident table bytes */
  double beta;
} ksj_gpu_plan3_info;

typedef struct ksj_gpu_plan3 ksj_gpu_plan3;

int ksj_gpu_plan3_create(const ksj_gpu_plan3_desc* desc, ksj_gpu_plan3** out_plan);
/* Builds every table from d_traj ([S][3] float32); d_dcf is read only for DENSITY_GIVEN.
 * Synchronises `stream`.  Rejects non-finite coordinates and |w| > pi. */
int ksj_gpu_plan3_set_trajectory(ksj_gpu_plan3* plan, const float* d_traj, const float* d_dcf, void* stream);
int ksj_gpu_plan3_get_info(const ksj_gpu_plan3* plan, ksj_gpu_plan3_info* out_info);
/* d_cell_start[grid_z*grid_y*grid_x + 1] and d_perm[samples], for bit-exact checks of the binning. */
int ksj_gpu_plan3_export_bins(const ksj_gpu_plan3* plan, uint32_t* d_cell_start, uint32_t* d_perm, void* stream);
void ksj_gpu_plan3_destroy(ksj_gpu_plan3* plan);

/* Channel-major samples [C][S] -> coil images coil-last [z][y][x][C] (if d_images)
 * and/or the RSS volume (if d_rss), coils taken coils_per_pass (1..64) at a time:
 * one pass's grid holds G_z*G_y*G_x*padded(coils_per_pass) complex values. */
size_t ksj_gpu_adjoint3_workspace_bytes(const ksj_gpu_plan3* plan, uint32_t channels, uint32_t coils_per_pass);
int ksj_gpu_adjoint3_execute(const ksj_gpu_plan3* plan, const void* d_kspace, uint32_t channels,
                             uint32_t coils_per_pass, void* d_images, float* d_rss, void* d_workspace,
                             size_t workspace_bytes, void* stream);

/* Golden-means kooshball [spokes*readout][3]: spoke p along (c, sqrt(1-c^2) sin a,
 * sqrt(1-c^2) cos a), c = frac(0.4656 p), a = 2 pi frac(0.6823 p); readout
 * k_j = (j - M/2)/M cycles per pixel; FP64 on the device, rounded to float. */
int ksj_gpu_golden_means3(float* d_traj, uint32_t spokes, uint32_t readout, void* stream);

/* FP64 forward NUFFT of a volume, for simulating k-space: y_s ~ sum_v x_v exp(-i w_s . (v - o))
 * for x [Nz][Ny][Nx] complex128 (image = {Nz, Ny, Nx}, each side 1..512 with twice the side a
 * power of two), origin o = {oz, oy, ox} and a trajectory [S][3] (z, y, x) float rad/px;
 * y [S] complex128.  Kaiser-Bessel width 4, 6 or 8 on the 2x grid (Beatty beta), the same
 * kernel and deapodization as the adjoint, FP64 throughout (scatter of x / phi_hat, a radix-2
 * FFT per axis, W^3-tap interpolation).  No reference counterpart: the reference has no 3-D
 * path; the 3-D workload's data are defined as an FP64 width-6 NUFFT of a phantom.
 * Workspace: ksj_gpu_nufft3_forward_f64_workspace_bytes (the complex128 grid).  Asynchronous. */
size_t ksj_gpu_nufft3_forward_f64_workspace_bytes(const uint32_t* image);
int ksj_gpu_nufft3_forward_f64(const void* d_image, const uint32_t* image, const double* origin, const float* d_traj,
                               uint64_t samples, uint32_t width, void* d_out, void* d_workspace,
                               size_t workspace_bytes, void* stream);

/* ------------------------------------------------------------------ HIP graphs */

/* Captures the work enqueued on `stream` (and on streams it forks to through events)
 * between begin and end into an executable graph.  `stream` must not be the null
 * stream.  Replaying the graph repeats the captured launches with the same pointers. */
typedef struct ksj_gpu_graph ksj_gpu_graph;
int ksj_gpu_graph_begin(void* stream);
int ksj_gpu_graph_end(void* stream, ksj_gpu_graph** out_graph);
int ksj_gpu_graph_launch(ksj_gpu_graph* graph, void* stream);
void ksj_gpu_graph_destroy(ksj_gpu_graph* graph);

/* ------------------------------------------------------------------ serving */

/* A continuous-batching server on one GPU.  No reference counterpart: the reference
 * runs one frame per synchronous firing (providers/kspacejet-noncartesian-recon/src/
 * operators/radial_gridding_reconstruct.cpp:228-437); this is the served form of the
 * same operators for a stream of scans whose inputs already sit in HBM.
 *
 * A scheduler thread admits requests in arrival order: a batch is launched when one of
 * `lanes` streams is free and its workspace fits the HBM budget next to the resident
 * plans.  Radial requests that share a plan (same geometry and trajectory key) are
 * stacked along the coil axis into one launch of up to max_batch_scans scans;
 * Cartesian requests of one shape share a launch.  Plans stay resident per
 * (geometry, trajectory key), least recently used evicted first.  Requests that arrive
 * while others run join the next free lane; batches on different lanes overlap.  Each
 * result equals the result of the same request run alone. */
#define KSJ_GPU_SERVE_RADIAL2D 0     /* KB adjoint NUFFT, grid 2N, ramp or no density weights */
#define KSJ_GPU_SERVE_CARTESIAN2D 1  /* inverse FFT (Normalization::inverse), centred or not */
#define KSJ_GPU_SERVE_RADIAL3D 2     /* 3-D KB adjoint of a volume, grid 2N per axis, |k|^2 or no
                                        density weights; runs alone, coils in passes of 8 */

typedef struct ksj_gpu_serve_config {
  uint32_t struct_size;
  int32_t device;
  uint64_t hbm_budget_bytes;   /* 0 = 80% of the device's free memory */
  uint32_t lanes;              /* batches in flight, 1..8 (0 = 2) */
  uint32_t max_batch_scans;    /* scans per radial launch, 1..64 (0 = 8) */
  uint32_t plan_cache_entries; /* 0 = 16 */
  uint32_t host_staging_mib;   /* pinned staging of host-memory requests in MiB, allocated at the first
                                  such request (0 = 1024) */
} ksj_gpu_serve_config;

#define KSJ_GPU_SERVE_DEVICE_MEMORY 0 /* request buffers are device memory of the server's GPU */
#define KSJ_GPU_SERVE_HOST_MEMORY 1   /* request buffers are host memory (pageable or pinned) */

typedef struct ksj_gpu_serve_request {
  uint32_t struct_size;
  uint32_t kind;                 /* KSJ_GPU_SERVE_* */
  uint32_t image_rows, image_cols, channels;
  uint32_t samples;              /* radial */
  uint32_t kernel_width;         /* radial: 4, 6 or 8 */
  uint32_t density;              /* radial: KSJ_GPU_DENSITY_NONE or _RAMP */
  uint32_t centred;              /* Cartesian */
  uint32_t image_depth;          /* 3-D: N_z (image_rows = N_y, image_cols = N_x); 0 otherwise */
  uint64_t trajectory_key;       /* radial: names the trajectory bytes; 0 = the server hashes them */
  const float* d_traj;           /* radial: [S][ky,kx] rad/px; 3-D: [S][kz,ky,kx] */
  const void* d_kspace;          /* [C][S] (radial) or [C][R][Cols] (Cartesian) cf32 */
  void* d_images;                /* coil-last [R][Cols][C] cf32, or NULL */
  float* d_rss;                  /* [R][Cols] float, or NULL */
  uint32_t memory;               /* KSJ_GPU_SERVE_*_MEMORY: where the four buffers above lie (2-D kinds
                                    only may be host memory).  A host-memory request is copied into the
                                    server's pinned staging when it is submitted (the call blocks while
                                    the staging is full), sent to the GPU on its lane's stream ahead of
                                    its kernels (so the transfer overlaps other lanes' work), and its
                                    outputs are copied back to the caller's buffers when it completes.
                                    A request struct without this field (an older, shorter struct_size)
                                    is device memory. */
  uint32_t reserved1;
} ksj_gpu_serve_request;

typedef struct ksj_gpu_serve_stats {
  uint32_t struct_size, reserved0;
  uint64_t submitted, completed, failed, batches;
  uint64_t plan_builds, plan_hits, plan_evictions;
  uint64_t peak_device_bytes, plan_bytes, max_batch_seen;
} ksj_gpu_serve_stats;

typedef struct ksj_gpu_server ksj_gpu_server;

int ksj_gpu_serve_create(const ksj_gpu_serve_config* config, ksj_gpu_server** out_server);
/* Queues a request (its buffers must stay valid until it completes); *ticket > 0. */
int ksj_gpu_serve_submit(ksj_gpu_server* server, const ksj_gpu_serve_request* request, uint64_t* ticket);
/* Queues `count` requests at once (all validated first; none is queued on an error). */
int ksj_gpu_serve_submit_batch(ksj_gpu_server* server, uint32_t count, const ksj_gpu_serve_request* requests,
                               uint64_t* tickets);
/* Blocks until the ticket completes; returns its status, and the submit-to-completion
 * latency in milliseconds.  Each ticket is waited on once. */
int ksj_gpu_serve_wait(ksj_gpu_server* server, uint64_t ticket, double* latency_ms);
/* Blocks until every submitted request has completed. */
int ksj_gpu_serve_drain(ksj_gpu_server* server);
/* on != 0: launch nothing new until called again with 0 (requests queue up meanwhile
 * and are batched when released).  Requests still queued at destroy are dropped. */
int ksj_gpu_serve_hold(ksj_gpu_server* server, int on);
int ksj_gpu_serve_get_stats(ksj_gpu_server* server, ksj_gpu_serve_stats* out_stats);
void ksj_gpu_serve_destroy(ksj_gpu_server* server);

/* ------------------------------------------------------------------ tuning */

/* Process-wide launch parameters chosen by the autotuner (kspacejet_amd/autotune.py).
 * No reference counterpart.  Keys and legal values:
 *   "gather_map"           0 = work items in launch order, 1 = one contiguous item range per XCD,
 *                          2 = contiguous ranges of equal cost (visits) per XCD
 *   "item_visits"          1..4096, default gather work-item size of plans created afterwards
 *   "fft_lines_per_block"  0 = per-length default, else 1..64 (power of two) lines per FFT block
 *   "gather_batch"         4, 8 or 16 gather visits per batch (sample-row loads in flight per lane)
 *   "pass_lanes"           1..3 streams over which the passes of a batched adjoint overlap (the
 *                          batch workspace holds one pass workspace per lane)
 *   "multi_lanes"          1..8 streams over which the items of ksj_gpu_adjoint2_execute_multi spread
 *   "gather_coils_per_lane" 1 or 2 coils per gather lane (one 8- or 16-byte load per visit) from 32
 *                          stacked coils up
 *   "block_group"          1, 2 or 4: side of the groups of 2x2 blocks numbered together (new plans)
 *   "multi_fused"          0, 1 or 2: form of ksj_gpu_adjoint2_execute_multi (see there)
 *   "fft_schedule"         0 (default), 1 or 2: radix schedule of every LDS FFT length; 3: 1 for
 *                          length 512, 0 for the others (for overlapped serving; rounding only)
 *   "gather_scheme"        0 = 2x2 blocks, 1 = 64-cell LDS tiles, 2 = 4x4 blocks on the matrix
 *                          cores (new plans; summation order)
 *   "matrix_run_cost"      1..64: work per gather wave of the matrix schedule, in groups of four
 *                          visits plus one per item (new plans; default 12)
 *   "matrix_item_visits"   1..4096: work-item size of new matrix-schedule plans (summation order)
 *   "matrix_batch"         4, 8 or 16 visits per batch of the matrix gathers
 *   "matrix_pipeline"      1 or 2 batches of sample rows in flight per wave (2-D matrix gather)
 *   "matrix_nt_store"      0 or 1: grid cells of the 2-D matrix gather written non-temporally
 *   "matrix_slab"          1..4096 consecutive workgroups per XCD slab (2-D matrix gather)
 *   "matrix_block_width"   4 or 8: 4x4 or 4x8 cell blocks of new 2-D matrix plans (summation order)
 *   "gather3_scheme"       0 = 2x2x2 blocks, 2 = 4x4x4 blocks on the matrix cores (new 3-D plans;
 *                          summation order)
 *   "matrix3_item_visits"  1..4096: work-item size of new 3-D matrix plans (summation order)
 *   "matrix3_run_cost"     1..64: work per gather wave of new 3-D matrix plans
 *   "tile_width"           4 or 8 columns per tile (new plans; summation order)
 *   "tile_item_samples"    window samples per tile work item (new plans; summation order)
 *   "mix_scheme"           1 = ksj_gpu_channel_mix on the matrix cores (default), 0 = vector FMAs
 *                          (summation order)
 *   "matrix_persistent"    0 = one gather wave per run (default); 1..8 = that many resident workgroups per CU,
 *                          each XCD's waves looping over a contiguous eighth of the runs (same bits)
 *   "matrix_tap_share"     1 = each gather lane loads one word of a visit's taps and takes its two taps from
 *                          the visit's lanes by bpermute, the packed visits of a batch likewise (default);
 *                          0 = each lane loads its own taps and visits (same bits)
 *   "matrix3_tap_share"    3-D matrix gather: 1 = one or two words of a visit's 3W taps per lane, handed round
 *                          by bpermute; 0 = each lane loads its own six taps (default; same bits)
 *   "matrix3_map"          3-D matrix gather: 0 = runs in launch order (default), 1 = each XCD the contiguous
 *                          eighth of the runs, s > 1 = slabs of s workgroups dealt round-robin to the XCDs
 *                          (same bits)
 *   "matrix_weights"       1 = new 2-D matrix plans store a 4-byte (sample, window bin) visit and the
 *                          gather forms the cell weights from the sample's taps (default), 0 = every
 *                          visit stores its 16 or 32 cell weights (same bits; plan size and speed only)
 *   "fft_lds_pad"          register-form FFT: LDS element stride = lines per block + pad (1..4, default 1;
 *                          same bits)
 *   "fft_register"         1 = FFT passes whose loads and stores both run across contiguous lines keep
 *                          the first and last radix pass in registers, and 2-D radial coil-last
 *                          outputs take one column pass that writes the coil images and RSS itself
 *                          (default); 0 = every pass staged through LDS (same bits)
 *   "row_order"            coil-fastest row-pass output of the 2-D column-epilogue pipeline: 0 = grid-row
 *                          major (default), 1 = column major (same bits)
 *   "assemble_rows"        ksj_gpu_assemble_acquisitions: channel rows of one acquisition per workgroup, 1,
 *                          2 or 4 (default; the largest of them dividing the channels is taken; same bits)
 *   "grid_prune"           1 = cells of the oversampled grid that no sample reaches are neither written by
 *                          the 2-D / 3-D matrix gathers nor read by the first FFT pass, which takes them as
 *                          zero (default); 0 = every cell written and read (same bits)
 *   "serve_copy_threads"   worker threads (0..15, default 7) a server created afterwards adds to the
 *                          submitting thread for memcpy of host-memory requests into and out of its pinned
 *                          staging (copies of 4 MiB or more; same bits)
 * Every setting gives bit-identical results, except item_visits, matrix_item_visits,
 * matrix3_item_visits, matrix_block_width, the tile keys, gather_scheme and gather3_scheme (float summation
 * order of the gather), fft_schedule (FFT rounding) and mix_scheme (summation order of the
 * channel mix); each of those is tested against the FP64 oracle.  Not thread-safe
 * against concurrent launches. */
int ksj_gpu_tune_set(const char* key, int64_t value);
int ksj_gpu_tune_get(const char* key, int64_t* value);

/* ------------------------------------------------------------------ stage timing */

/* Per-stage timing with HIP events recorded on the stream each stage runs on
 * (stages: grid_prep, grid_gather, grid_reduce, fft_rows, fft_cols_deapod, coil_combine).
 * Off by default; no reference counterpart (the reference's kspacejet-performance
 * spans are compiled out unless KSJ_PERFORMANCE_LEVEL is set).  profile_read
 * synchronises the pending events of that stage. */
int ksj_gpu_profile_enable(int on);
int ksj_gpu_profile_reset(void);
int ksj_gpu_profile_read(const char* stage, double* total_ms, uint64_t* count);

#ifdef __cplusplus
}
#endif

#endif /* KSJ_GPU_H_ */
