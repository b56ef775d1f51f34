/*
 * FP64 CPU oracle for the KSpaceJet reconstruction hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (kspacejet_amd/) may link,
 * load or call this library; it exists so tests/, the smoke check and the
 * bench's CPU-baseline leg can check the HIP path against a plain restatement
 * of the reference algorithms.
 *
 * Complex arrays are interleaved (re, im) doubles.  Trajectories are [S][2]
 * rows of (row_frequency = ky, col_frequency = kx) in radians per pixel, the
 * layout of the reference ksj.trajectory-frame payload.
 *
 * Each function cites the reference file:line it restates.  Paths are
 * relative to the KSpaceJet source tree.
 */
#ifndef KSJ_ORACLE_H_
#define KSJ_ORACLE_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Exact non-uniform DFTs.
 * libs/numerics/kspacejet-nufft/src/eigen/eigen_nufft_direct_nudft.cpp:52-74 (forward, sign -i)
 * and :77-97 (adjoint, sign +i, no normalisation, no density weights). */
void ora_nudft2_forward(size_t rows, size_t cols, double row_origin, double col_origin, size_t samples,
                        const double* traj, const double* image, double* out);
void ora_nudft2_adjoint(size_t rows, size_t cols, double row_origin, double col_origin, size_t samples,
                        const double* traj, const double* values, double* image);

/* Analytic ramp density compensation hypot(k_row, k_col).
 * libs/numerics/kspacejet-nufft/src/eigen/eigen_radial_gridding.cpp:279-295 */
void ora_ramp_dcf2(size_t samples, const double* traj, double* dcf);

/* Reference radial route: periodic 2x2 bilinear gridding on the unoversampled
 * rows x cols grid with the origin phase, then an unscaled +i inverse DFT per axis.
 * libs/numerics/kspacejet-nufft/src/eigen/eigen_radial_gridding.cpp:262-277 (grid coordinate),
 * :297-363 (gridding and separable inverse transform).  Returns 0, or -1 on a
 * trajectory outside [-pi, pi]. */
int ora_radial_linear_gridding2(size_t rows, size_t cols, double row_origin, double col_origin, size_t samples,
                                const double* values, const double* traj, const double* dcf, double* image);

/* In-place unscaled DFT of length n.  sign = +1 computes sum x e^{+2 pi i jk/n}
 * (the inverse direction), sign = -1 the forward direction.  Mixed-radix
 * Stockham with direct prime-radix butterflies; any n >= 1. */
void ora_fft1(size_t n, double* data, int sign);

/* Inverse 2-D transform with Normalization::inverse (1/cols per row pass then
 * 1/rows per column pass), uncentred, as
 * libs/numerics/kspacejet-fft/src/eigen/eigen_fft_transforms.cpp:14-27,148-205.
 * centred != 0 computes fftshift(ifft2(ifftshift(x))) instead. */
void ora_ifft2(size_t rows, size_t cols, double* data, int centred);

/* Root sum of squares over the contiguous (last) coil axis, accumulated in coil
 * order, then sqrt.  libs/numerics/kspacejet-array/include/kspacejet/array/reductions.hpp:973-983,
 * libs/numerics/kspacejet-stats/src/eigen/eigen_stats_common.hpp:376-381. */
void ora_rss(size_t pixels, size_t coils, const double* coil_last, double* out);

/* Sensitivity-weighted combine x = sum_c conj(S_c) I_c / sum_c |S_c|^2 (0 where the
 * denominator is 0).  The reference only plans this interface
 * (providers/interfaces/org.kspacejet.coil-combine/coil_combine_adaptive.json). */
void ora_sense_combine(size_t pixels, size_t coils, const double* images, const double* sens, double* out);

/* Kaiser-Bessel pieces (the reference has no KB gridding; Beatty et al. 2005). */
double ora_i0(double x);
double ora_kb_beta(int width, double sigma);
double ora_kb_kernel(double distance, int width, double beta);
double ora_kb_ft(double index, size_t grid, int width, double beta);

/* Grid coordinate of a float32 radian-per-pixel frequency on a grid of `extent`
 * cells: u = (double)w * extent / (2 pi), wrapped once into [0, extent); then the
 * cell floor(u) and the fraction u - floor(u) rounded to float.  Same rule as
 * libs/numerics/kspacejet-nufft/src/eigen/eigen_radial_gridding.cpp:262-277, in FP64.
 * Returns -1 if the coordinate cannot be mapped. */
int ora_grid_coordinate(float w, uint32_t extent, uint32_t* cell, float* frac);

/* Bins every sample to cell index cell_row * grid_cols + cell_col and stably sorts
 * by bin (ties keep sample order).  cell_start has grid_rows*grid_cols+1 entries.
 * frac is [S][2] (row, col) in sample order.  Returns -1 on a bad coordinate. */
int ora_bin2(size_t samples, const float* traj, uint32_t grid_rows, uint32_t grid_cols, uint32_t* bins,
             float* frac, uint32_t* cell_start, uint32_t* perm);

/* FP64 KB adjoint NUFFT of one coil:
 *   c_s = value_s * dcf_s * exp(+i(w_r (N_r/2 - r0) + w_c (N_c/2 - c0)))
 *   spread c_s with phi(d_r) phi(d_c) onto G_r x G_c (taps at cell-W/2+1 .. cell+W/2)
 *   unscaled +i inverse DFT of the grid, crop indices j in [-N/2, N/2),
 *   divide by phi_hat(j_r) phi_hat(j_c); output pixel r = j_r + N_r/2.
 * This equals sum_s value_s dcf_s exp(+i(w_r (r - r0) + w_c (c - c0))) up to the KB
 * approximation error.  dcf may be NULL (unit weights).  Returns -1 on a bad coordinate. */
int ora_kb_adjoint2(size_t n_rows, size_t n_cols, size_t g_rows, size_t g_cols, int width, double beta,
                    double row_origin, double col_origin, size_t samples, const double* traj, const double* values,
                    const double* dcf, double* image);

/* Golden-angle radial trajectory: spoke p at angle p * pi / phi (111.246 deg),
 * readout k_j = (j - M/2) / M cycles per pixel, stored as [S][2] (ky, kx) radians per pixel. */
void ora_golden_angle2(size_t spokes, size_t readout, double* traj);

/* Route trajectory normalisation (noncartesian_rss_hdf5.cpp:776-834): raw [S][kx, ky] in
 * units 1 = cycles/FOV, 2 = rad/px, 3 = encoded matrix index (0 = unspecified) ->
 * [S][ky, kx] rad/px.  Returns 0 or the first failure: 1 non-finite, 2 outside the
 * Nyquist interval, 3 unspecified units, 4 encoded dimension < 2. */
int ora_normalize_trajectory2(size_t samples, const float* raw, int units, uint32_t encoded_rows,
                              uint32_t encoded_cols, float* out);

/* ---- 3-D: coordinates [S][3] in (z, y, x) order, radians per pixel; images [z][y][x]. */

/* x(z,y,x) = sum_s y_s exp(+i(w_z (z-oz) + w_y (y-oy) + w_x (x-ox))), no normalisation. */
void ora_nudft3_adjoint(size_t nz, size_t ny, size_t nx, double oz, double oy, double ox, size_t samples,
                        const double* traj, const double* values, double* image);
/* |k|^2 per sample. */
/* Exact 3-D forward model: y_s = sum_v x_v exp(-i w_s . (v - o)), the transpose of
 * ora_nudft3_adjoint; image [nz][ny][nx] and values [S] interleaved complex128. */
void ora_nudft3_forward(size_t nz, size_t ny, size_t nx, double oz, double oy, double ox, size_t samples,
                        const double* traj, const double* image, double* values);
void ora_ramp_sq_dcf3(size_t samples, const double* traj, double* dcf);
/* Bin (cz * gy + cy) * gx + cx per sample from the float coordinates, stable counting sort. */
int ora_bin3(size_t samples, const float* traj, uint32_t gz, uint32_t gy, uint32_t gx, uint32_t* bins,
             uint32_t* cell_start, uint32_t* perm);
/* KB gridding adjoint on a (gz, gy, gx) grid, centred crop to (nz, ny, nx), deapodized. */
int ora_kb_adjoint3(size_t nz, size_t ny, size_t nx, size_t gz, size_t gy, size_t gx, int width, double beta,
                    double oz, double oy, double ox, size_t samples, const double* traj, const double* values,
                    const double* dcf, double* image);
/* ora_kb_adjoint3 evaluated only at the output voxels of the lattice zs x ys x xs (indices
 * into the cropped volume), from samples [s_begin, s_end): the DFT of the KB grid is summed
 * directly at those voxels, per sample as a product of three one-axis sums over its taps,
 * and divided by the deapodization.  `out` [nzs][nys][nxs] (interleaved) is ADDED to, so
 * sample ranges can be split across threads.  The grid-cell rule, taps, weight and origin
 * phase are those of ora_kb_adjoint3.  Returns -1 on an unmappable coordinate. */
int ora_kb_adjoint3_voxels(size_t nz, size_t ny, size_t nx, size_t gz, size_t gy, size_t gx, int width, double beta,
                           double oz, double oy, double ox, size_t s_begin, size_t s_end, const double* traj,
                           const double* values, const double* dcf, size_t nzs, const uint32_t* zs, size_t nys,
                           const uint32_t* ys, size_t nxs, const uint32_t* xs, double* out);
/* Golden-means kooshball: spoke p along (frac(p g1), sqrt(1-c^2) sin a, sqrt(1-c^2) cos a),
 * a = 2 pi frac(p g2), g1 = 0.4656, g2 = 0.6823; readout as ora_golden_angle2; [S][3]. */
void ora_golden_means3(size_t spokes, size_t readout, double* traj);

/* ---- k-space conditioning (providers/kspacejet-kspace-conditioning/src/operators/). */

/* y[o][s] = sum_i w[o][i] x[i][s], complex, FP64 accumulation in input order: the
 * channel-major matmul W[out x in] X[in x samples] of noise_prewhiten.cpp,
 * coil_compress.cpp and their non-Cartesian forms (samples = rows * cols for a frame). */
void ora_channel_mix(size_t out_channels, size_t in_channels, size_t samples, const double* w, const double* x,
                     double* y);
/* y[c][r][k] = x[c][r][k] * p[c][k] in float, (ac - bd, ad + bc) with every product and
 * sum rounded to float (phase_correct.cpp:127-143; std::complex<float> multiply of finite
 * values). */
void ora_phase_correct(size_t channels, size_t rows, size_t cols, const float* x, const float* p, float* y);
/* y[l][k] = x[l][offset + k] for every line l of a channel-major frame
 * (readout_oversampling_remove.cpp:165-181). */
void ora_readout_crop(size_t lines, size_t in_cols, size_t out_cols, size_t offset, const float* x, float* y);

/* ---- calibration estimators (providers/kspacejet-calibration/src/operators/). */

/* Hermitian covariance [C][C] (interleaved complex128) of float samples x[C][S] in FP64.
 * centred = 1: sum_s conj(x_r - m_r)(x_c - m_c) / S with per-channel means m
 * (kspacejet-stats moments.hpp:99-140, population normalisation as noise_model_estimate.cpp:182);
 * centred = 0: sum_s x_r conj(x_c) / S (coil_compression_basis_estimate.cpp:207-227). */
void ora_channel_covariance(size_t channels, size_t samples, const float* x, int centred, double* cov);
/* The reference's largest-pivot Jacobi eigensolver (kspacejet-linalg decompositions.hpp:533-626)
 * restated in FP64: eigenvalues ascending, eigenvector k in column k of vectors [n][n]
 * (interleaved).  Returns 0, or -1 on a non-finite entry or no convergence. */
int ora_hermitian_eigen(size_t n, const double* matrix, double* values, double* vectors);
/* Whitening W = V diag(1 / sqrt(max(lambda, floor))) V^H of the centred covariance, [C][C]
 * (whitening.hpp:24-50 over noise_model_estimate.cpp:182-190).  Returns 0 or -1. */
int ora_noise_model_estimate(size_t channels, size_t samples, const float* x, double floor_value, double* whitening);
/* B[v][p] = conj(V[p][P-1-v]) of the uncentred covariance, [V][P]
 * (coil_compression_basis_estimate.cpp:207-238).  Returns 0 or -1. */
int ora_coil_compression_basis_estimate(size_t physical, size_t virtual_channels, size_t samples, const float* x,
                                        double* basis);
/* p[c][k] = conj(s) / |s|, s = float sum over lines of ref[c][line][k] in line order,
 * |s| = (float) sqrt((double) re^2 + (double) im^2) (the computation of glibc's hypotf
 * behind std::abs of std::complex<float>); 1 + 0i where |s| = 0
 * (phase_correction_estimate.cpp:141-163). */
void ora_phase_correction_estimate(size_t channels, size_t lines, size_t readout, const float* ref, float* phase);

/* ---- sensitivity maps (planned interface providers/interfaces/org.kspacejet.calibration/
 * sensitivity_map_estimate.json; no reference implementation). */

/* Walsh's adaptive sensitivity estimate: at each pixel, the dominant eigenvector of
 * R = sum over a window x window neighbourhood (clipped at the image edges) of I I^H, by
 * power iteration from R e_0 (normalised after every step, `iterations` steps), rotated
 * so that coil 0 is real and >= 0; unit norm, or zero where R e_0 vanishes.
 * images and maps [rows][cols][coils], interleaved complex128. */
void ora_walsh_maps(size_t rows, size_t cols, size_t coils, size_t window, size_t iterations, const double* images,
                    double* maps);

#ifdef __cplusplus
}
#endif

#endif
