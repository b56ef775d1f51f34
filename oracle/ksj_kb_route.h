/* FP32 CPU restatements of the GPU routes (test infrastructure: the like-for-like CPU
 * baselines of the benchmark; see ksj_kb_route.c for the definitions). */
#ifndef KSJ_KB_ROUTE_H_
#define KSJ_KB_ROUTE_H_

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* RSS image [n][n] of one scan (values: [coils][samples] interleaved complex float,
 * traj: [samples][ky, kx] rad/px) through ramp DCF, KB width `width` (4/6/8) with the
 * Beatty beta on the 2n x 2n grid (n a power of two), inverse FFT, deapodization and
 * crop; the scan is reconstructed `scans` times on `threads` threads (the timing
 * sample) and the result written once.  `scan_seconds` (may be NULL) receives the wall
 * time of the scans, after the per-trajectory tables are made.  Returns 0, or -1 on a
 * bad argument or coordinate. */
int ora_kb_route_rss_f32(size_t n, size_t samples, size_t coils, const float* traj, const float* values, int width,
                         size_t scans, int threads, float* rss, double* scan_seconds);

/* RSS image [rows][cols] of a Cartesian scan (kspace: [coils][rows][cols] interleaved
 * complex float, rows and cols powers of two): per coil the inverse 2-D FFT with
 * Normalization::inverse, centred (fftshift(ifft2(ifftshift(x)))) or not, then RSS in
 * coil order; `scans` repeats on `threads` threads, timed as above.  Returns 0 or -1. */
int ora_cart_route_rss_f32(size_t rows, size_t cols, size_t coils, const float* kspace, int centred, size_t scans,
                           int threads, float* rss, double* scan_seconds);

/* RSS volume [n][n][n] of a 3-D kooshball scan (values [coils][samples], traj
 * [samples][z, y, x] rad/px; n >= 16, 2n a power of two): the KB adjoint of
 * ora_kb_adjoint3 with the |k|^2 DCF and origin (n-1)/2, width `width` (4/6/8), FP32;
 * threads take whole coils, each on its own (2n)^3 grid.  `coil_seconds` (may be NULL)
 * receives the wall time of the coils after the per-trajectory tables.  Returns 0, or -1
 * on a bad argument, an unmappable coordinate or an allocation failure. */
int ora_kb_route3_rss_f32(size_t n, size_t samples, size_t coils, const float* traj, const float* values, int width,
                          int threads, float* rss, double* coil_seconds);

/* The route's replay copy loop (noncartesian_rss_hdf5.cpp:1077-1083) in C, for the CPU
 * baseline of the acquisition assembly: acquisitions packed one after another, each
 * [channels][counts[a]] complex float (ISMRMRD samples[sample + channel * ns]), -> frame
 * [channels][S] with one memcpy per acquisition and channel, repeated `frames` times on one
 * thread.  `seconds` (may be NULL) receives the wall time of the repeats.  Returns 0, or -1
 * on the route's rejections (no acquisition, one without samples, channels outside [1, 64]). */
int ora_assemble_acquisitions_f32(const float* packed, const unsigned* counts, size_t acquisitions, size_t channels,
                                  size_t frames, float* frame, double* seconds);

#ifdef __cplusplus
}
#endif

#endif
