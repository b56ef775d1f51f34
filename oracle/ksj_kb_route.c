/* CPU restatement of the GPU radial reconstruction route, FP32 (test infrastructure:
 * the like-for-like CPU baseline the benchmark times beside the GPU path, and nothing
 * else; no product path links it).
 *
 * Definition (the one ora_kb_adjoint2 states in FP64, which the tests hold it to):
 *   c_s = y_s * hypot(w_r, w_c) * exp(+i(w_r (N/2 - r0) + w_c (N/2 - c0))),  r0 = c0 = (N-1)/2
 *   spread c_s with phi(d_r) phi(d_c), Kaiser-Bessel width W, Beatty beta for sigma = 2,
 *   onto the 2N x 2N grid (taps at cell - W/2 + 1 .. cell + W/2, wrapped),
 *   unscaled +i inverse FFT, crop j in [-N/2, N/2), divide by phi_hat(j_r) phi_hat(j_c),
 *   RSS over the coils.
 * The per-trajectory part (cells, taps, DCF and phase factors, deapodization) is made
 * once, as the GPU's resident plan is; each scan then grids every coil on its own grid
 * (2N x 2N complex, 2 MiB at N = 256, cache resident), runs radix-2 row FFTs and the
 * column FFTs of the N columns the crop keeps, and accumulates the RSS.  Scans run in
 * parallel, one per thread. */
#define _POSIX_C_SOURCE 200809L
#include "ksj_kb_route.h"
#include "ksj_oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

typedef struct {
  float re, im;
} cf;

typedef struct {
  size_t n, g, samples, coils, width;
  uint32_t* r0;       /* [S] first tap row (wrapped) */
  uint32_t* c0;       /* [S] first tap column */
  float* taps;        /* [S][2W]: row taps, then column taps */
  cf* factor;         /* [S] dcf * exp(+i phase) */
  float* inv_deapod;  /* [N][N] */
  cf* twiddle;        /* [G/2] exp(+2 pi i k / G) */
  uint32_t* bitrev;   /* [G] */
} route;

static int log2_exact(size_t g) {
  int b = 0;
  while (((size_t)1 << b) < g) ++b;
  return ((size_t)1 << b) == g ? b : -1;
}

static void fft_inplace(const route* p, cf* a, size_t stride) {
  const size_t g = p->g;
  for (size_t i = 0; i < g; ++i) {
    const size_t j = p->bitrev[i];
    if (j > i) {
      const cf t = a[i * stride];
      a[i * stride] = a[j * stride];
      a[j * stride] = t;
    }
  }
  for (size_t len = 2; len <= g; len <<= 1) {
    const size_t half = len / 2, step = g / len;
    for (size_t i = 0; i < g; i += len)
      for (size_t j = 0; j < half; ++j) {
        const cf w = p->twiddle[j * step];
        cf* x = &a[(i + j) * stride];
        cf* y = &a[(i + j + half) * stride];
        const cf v = {y->re * w.re - y->im * w.im, y->re * w.im + y->im * w.re};
        const cf u = *x;
        x->re = u.re + v.re;
        x->im = u.im + v.im;
        y->re = u.re - v.re;
        y->im = u.im - v.im;
      }
  }
}

static void route_free(route* p) {
  free(p->r0);
  free(p->c0);
  free(p->taps);
  free(p->factor);
  free(p->inv_deapod);
  free(p->twiddle);
  free(p->bitrev);
}

static int route_make(route* p, size_t n, size_t samples, size_t coils, const float* traj, int width) {
  memset(p, 0, sizeof(*p));
  p->n = n;
  p->g = 2 * n;
  p->samples = samples;
  p->coils = coils;
  p->width = (size_t)width;
  const size_t g = p->g, W = (size_t)width;
  const int bits = log2_exact(g);
  if (bits < 1 || (width != 4 && width != 6 && width != 8)) return -1;
  p->r0 = (uint32_t*)malloc(samples * sizeof(uint32_t));
  p->c0 = (uint32_t*)malloc(samples * sizeof(uint32_t));
  p->taps = (float*)malloc(samples * 2 * W * sizeof(float));
  p->factor = (cf*)malloc(samples * sizeof(cf));
  p->inv_deapod = (float*)malloc(n * n * sizeof(float));
  p->twiddle = (cf*)malloc(g / 2 * sizeof(cf));
  p->bitrev = (uint32_t*)malloc(g * sizeof(uint32_t));
  if (!p->r0 || !p->c0 || !p->taps || !p->factor || !p->inv_deapod || !p->twiddle || !p->bitrev) {
    route_free(p);
    return -1;
  }
  const double beta = ora_kb_beta(width, 2.0);
  const int lead = width / 2 - 1;
  const double shift = (double)(n / 2) - ((double)n - 1.0) / 2.0;
  const double scale = (double)g / (2.0 * 3.14159265358979323846);
  for (size_t s = 0; s < samples; ++s) {
    const float wr = traj[2 * s], wc = traj[2 * s + 1];
    uint32_t cr, cc;
    float fr, fc;
    if (ora_grid_coordinate(wr, (uint32_t)g, &cr, &fr) || ora_grid_coordinate(wc, (uint32_t)g, &cc, &fc)) {
      route_free(p);
      return -1;
    }
    /* cells from the float coordinate (as the GPU bins its float payload), tap
     * distances, DCF and phase from the exact coordinate, as the GPU's FP64 plan */
    double ur = wr * scale, uc = wc * scale;
    if (ur < 0.0) ur += (double)g;
    if (ur >= (double)g) ur -= (double)g;
    if (uc < 0.0) uc += (double)g;
    if (uc >= (double)g) uc -= (double)g;
    const double dr = ur - (double)cr, dc = uc - (double)cc;
    for (size_t t = 0; t < W; ++t) {
      p->taps[s * 2 * W + t] = (float)ora_kb_kernel((double)((int)t - lead) - dr, width, beta);
      p->taps[s * 2 * W + W + t] = (float)ora_kb_kernel((double)((int)t - lead) - dc, width, beta);
    }
    p->r0[s] = (uint32_t)((cr + g - (size_t)lead) % g);
    p->c0[s] = (uint32_t)((cc + g - (size_t)lead) % g);
    const double dcf = hypot((double)wr, (double)wc);
    const double phase = (double)wr * shift + (double)wc * shift;
    p->factor[s].re = (float)(dcf * cos(phase));
    p->factor[s].im = (float)(dcf * sin(phase));
  }
  for (size_t r = 0; r < n; ++r) {
    const double dr = ora_kb_ft((double)r - (double)(n / 2), g, width, beta);
    for (size_t c = 0; c < n; ++c)
      p->inv_deapod[r * n + c] = (float)(1.0 / (dr * ora_kb_ft((double)c - (double)(n / 2), g, width, beta)));
  }
  for (size_t k = 0; k < g / 2; ++k) {
    const double a = 2.0 * 3.14159265358979323846 * (double)k / (double)g;
    p->twiddle[k].re = (float)cos(a);
    p->twiddle[k].im = (float)sin(a);
  }
  for (size_t i = 0; i < g; ++i) {
    uint32_t r = 0;
    for (int b = 0; b < bits; ++b) r |= (uint32_t)((i >> b) & 1u) << (bits - 1 - b);
    p->bitrev[i] = r;
  }
  return 0;
}

/* One scan: values [coil][sample] interleaved complex float -> rss [N][N]. */
static void route_scan(const route* p, const float* values, cf* grid, cf* column, float* rss) {
  const size_t n = p->n, g = p->g, W = p->width, S = p->samples, mask = g - 1;
  for (size_t i = 0; i < n * n; ++i) rss[i] = 0.0f;
  for (size_t coil = 0; coil < p->coils; ++coil) {
    memset(grid, 0, g * g * sizeof(cf));
    const cf* y = (const cf*)values + coil * S;
    for (size_t s = 0; s < S; ++s) {
      const cf f = p->factor[s];
      const cf z = {y[s].re * f.re - y[s].im * f.im, y[s].re * f.im + y[s].im * f.re};
      const float* tr = p->taps + s * 2 * W;
      const float* tc = tr + W;
      const size_t r0 = p->r0[s], c0 = p->c0[s];
      for (size_t ty = 0; ty < W; ++ty) {
        cf* row = grid + ((r0 + ty) & mask) * g;
        const float zr = z.re * tr[ty], zi = z.im * tr[ty];
        for (size_t tx = 0; tx < W; ++tx) {
          cf* cell = row + ((c0 + tx) & mask);
          cell->re += zr * tc[tx];
          cell->im += zi * tc[tx];
        }
      }
    }
    for (size_t r = 0; r < g; ++r) fft_inplace(p, grid + r * g, 1);
    for (size_t c = 0; c < n; ++c) {
      const size_t ic = (c + g - n / 2) % g;
      for (size_t r = 0; r < g; ++r) column[r] = grid[r * g + ic];
      fft_inplace(p, column, 1);
      for (size_t r = 0; r < n; ++r) {
        const cf v = column[(r + g - n / 2) % g];
        const float d = p->inv_deapod[r * n + c];
        const float re = v.re * d, im = v.im * d;
        rss[r * n + c] += re * re + im * im;
      }
    }
  }
  for (size_t i = 0; i < n * n; ++i) rss[i] = sqrtf(rss[i]);
}

typedef struct {
  const route* p;
  const float* values;
  size_t first, scans, step;
  float* rss;
  int failed;
} worker;

static void* run_worker(void* arg) {
  worker* w = (worker*)arg;
  cf* grid = (cf*)malloc(w->p->g * w->p->g * sizeof(cf));
  cf* column = (cf*)malloc(w->p->g * sizeof(cf));
  if (!grid || !column) {
    w->failed = 1;
  } else {
    for (size_t s = w->first; s < w->scans; s += w->step) route_scan(w->p, w->values, grid, column, w->rss);
  }
  free(grid);
  free(column);
  return NULL;
}

static double now_seconds(void) {
  struct timespec t;
  clock_gettime(CLOCK_MONOTONIC, &t);
  return (double)t.tv_sec + 1e-9 * (double)t.tv_nsec;
}

int ora_kb_route_rss_f32(size_t n, size_t samples, size_t coils, const float* traj, const float* values, int width,
                         size_t scans, int threads, float* rss, double* scan_seconds) {
  route p;
  if (n < 2 || samples == 0 || coils == 0 || scans == 0 || threads < 1) return -1;
  if (route_make(&p, n, samples, coils, traj, width)) return -1;
  const double t0 = now_seconds();
  const size_t t = (size_t)threads < scans ? (size_t)threads : scans;
  worker* ws = (worker*)calloc(t, sizeof(worker));
  pthread_t* ids = (pthread_t*)calloc(t, sizeof(pthread_t));
  float* outs = (float*)malloc(t * n * n * sizeof(float));
  int rc = (ws && ids && outs) ? 0 : -1;
  size_t started = 0;
  for (size_t i = 0; rc == 0 && i < t; ++i) {
    ws[i] = (worker){&p, values, i, scans, t, outs + i * n * n, 0};
    if (pthread_create(&ids[i], NULL, run_worker, &ws[i]) != 0) rc = -1;
    else ++started;
  }
  for (size_t i = 0; i < started; ++i) {
    pthread_join(ids[i], NULL);
    if (ws[i].failed) rc = -1;
  }
  if (scan_seconds) *scan_seconds = now_seconds() - t0;
  if (rc == 0) memcpy(rss, outs, n * n * sizeof(float)); /* every scan's input is the same */
  free(ws);
  free(ids);
  free(outs);
  route_free(&p);
  return rc;
}
