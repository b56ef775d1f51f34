/* FP32 CPU restatement of the GPU radial route (test infrastructure: the like-for-like
 * CPU baseline of the benchmark; see ksj_kb_route.c for the definition). */
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

#ifdef __cplusplus
}
#endif

#endif
