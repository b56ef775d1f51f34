/* CPU restatements of the GPU reconstruction routes, FP32 (test infrastructure: the
 * like-for-like CPU baselines the benchmark times beside the GPU path, and nothing
 * else; no product path links it): the 2-D radial route below, then the Cartesian
 * inverse FFT + RSS and the 3-D kooshball route at the end of the file.
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

/* Radix-2 inverse-direction (+i) FFT of one power-of-two length. */
typedef struct {
  size_t g;
  cf* twiddle;        /* [g/2] exp(+2 pi i k / g) */
  uint32_t* bitrev;   /* [g] */
} fftplan;

typedef struct {
  size_t n, g, samples, coils, width;
  uint32_t* r0;       /* [S] first tap row (wrapped) */
  uint32_t* c0;       /* [S] first tap column */
  float* taps;        /* [S][2W]: row taps, then column taps */
  cf* factor;         /* [S] dcf * exp(+i phase) */
  float* inv_deapod;  /* [N][N] */
  fftplan f;
} route;

static int log2_exact(size_t g) {
  int b = 0;
  while (((size_t)1 << b) < g) ++b;
  return ((size_t)1 << b) == g ? b : -1;
}

static void fftplan_free(fftplan* f) {
  free(f->twiddle);
  free(f->bitrev);
  f->twiddle = NULL;
  f->bitrev = NULL;
}

static int fftplan_make(fftplan* f, size_t g) {
  f->g = g;
  f->twiddle = NULL;
  f->bitrev = NULL;
  const int bits = log2_exact(g);
  if (bits < 1) return -1;
  f->twiddle = (cf*)malloc(g / 2 * sizeof(cf));
  f->bitrev = (uint32_t*)malloc(g * sizeof(uint32_t));
  if (!f->twiddle || !f->bitrev) {
    fftplan_free(f);
    return -1;
  }
  for (size_t k = 0; k < g / 2; ++k) {
    const double a = 2.0 * 3.14159265358979323846 * (double)k / (double)g;
    f->twiddle[k].re = (float)cos(a);
    f->twiddle[k].im = (float)sin(a);
  }
  for (size_t i = 0; i < g; ++i) {
    uint32_t r = 0;
    for (int b = 0; b < bits; ++b) r |= (uint32_t)((i >> b) & 1u) << (bits - 1 - b);
    f->bitrev[i] = r;
  }
  return 0;
}

static void fft_inplace(const fftplan* f, cf* a, size_t stride) {
  const size_t g = f->g;
  for (size_t i = 0; i < g; ++i) {
    const size_t j = f->bitrev[i];
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
        const cf w = f->twiddle[j * step];
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
  fftplan_free(&p->f);
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
  if (!p->r0 || !p->c0 || !p->taps || !p->factor || !p->inv_deapod || fftplan_make(&p->f, g)) {
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
    for (size_t r = 0; r < g; ++r) fft_inplace(&p->f, grid + r * g, 1);
    for (size_t c = 0; c < n; ++c) {
      const size_t ic = (c + g - n / 2) % g;
      for (size_t r = 0; r < g; ++r) column[r] = grid[r * g + ic];
      fft_inplace(&p->f, column, 1);
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

/* ------------------------------------------------------------------ Cartesian (cfg1)
 * Centred (or uncentred) inverse 2-D FFT per coil with Normalization::inverse (1/cols
 * after the row pass, 1/rows after the column pass), then RSS over the coils: the
 * definition of ora_ifft2 + ora_rss, in FP32 on radix-2 FFTs.  Coil planes [coil][row][col]. */
typedef struct {
  size_t rows, cols, coils;
  int centred;
  fftplan fr, fc;
  const float* kspace;
  size_t first, scans, step;
  float* rss;
  int failed;
} cart_worker;

static void cart_scan(const cart_worker* w, cf* work, cf* column) {
  const size_t R = w->rows, Cn = w->cols, hr = w->centred ? R / 2 : 0, hc = w->centred ? Cn / 2 : 0;
  const float sr = 1.0f / (float)Cn, sc = 1.0f / (float)R;
  for (size_t i = 0; i < R * Cn; ++i) w->rss[i] = 0.0f;
  for (size_t coil = 0; coil < w->coils; ++coil) {
    const cf* k = (const cf*)w->kspace + coil * R * Cn;
    for (size_t r = 0; r < R; ++r) {  /* input shift (ifftshift for even sizes) */
      const cf* src = k + ((r + hr) % R) * Cn;
      cf* dst = work + r * Cn;
      for (size_t c = 0; c < Cn; ++c) dst[c] = src[(c + hc) % Cn];
      fft_inplace(&w->fc, dst, 1);
      for (size_t c = 0; c < Cn; ++c) {
        dst[c].re *= sr;
        dst[c].im *= sr;
      }
    }
    for (size_t c = 0; c < Cn; ++c) {
      for (size_t r = 0; r < R; ++r) column[r] = work[r * Cn + c];
      fft_inplace(&w->fr, column, 1);
      const size_t oc = (c + hc) % Cn;  /* output shift */
      for (size_t r = 0; r < R; ++r) {
        const float re = column[r].re * sc, im = column[r].im * sc;
        w->rss[((r + hr) % R) * Cn + oc] += re * re + im * im;
      }
    }
  }
  for (size_t i = 0; i < R * Cn; ++i) w->rss[i] = sqrtf(w->rss[i]);
}

static void* cart_run(void* arg) {
  cart_worker* w = (cart_worker*)arg;
  cf* work = (cf*)malloc(w->rows * w->cols * sizeof(cf));
  cf* column = (cf*)malloc(w->rows * sizeof(cf));
  if (!work || !column) {
    w->failed = 1;
  } else {
    for (size_t s = w->first; s < w->scans; s += w->step) cart_scan(w, work, column);
  }
  free(work);
  free(column);
  return NULL;
}

int ora_cart_route_rss_f32(size_t rows, size_t cols, size_t coils, const float* kspace, int centred, size_t scans,
                           int threads, float* rss, double* scan_seconds) {
  if (coils == 0 || scans == 0 || threads < 1) return -1;
  fftplan fr, fc;
  if (fftplan_make(&fr, rows)) return -1;
  if (fftplan_make(&fc, cols)) {
    fftplan_free(&fr);
    return -1;
  }
  const size_t t = (size_t)threads < scans ? (size_t)threads : scans;
  cart_worker* ws = (cart_worker*)calloc(t, sizeof(cart_worker));
  pthread_t* ids = (pthread_t*)calloc(t, sizeof(pthread_t));
  float* outs = (float*)malloc(t * rows * cols * sizeof(float));
  int rc = (ws && ids && outs) ? 0 : -1;
  const double t0 = now_seconds();
  size_t started = 0;
  for (size_t i = 0; rc == 0 && i < t; ++i) {
    ws[i] = (cart_worker){rows, cols, coils, centred, fr, fc, kspace, i, scans, t, outs + i * rows * cols, 0};
    if (pthread_create(&ids[i], NULL, cart_run, &ws[i]) != 0) rc = -1;
    else ++started;
  }
  for (size_t i = 0; i < started; ++i) {
    pthread_join(ids[i], NULL);
    if (ws[i].failed) rc = -1;
  }
  if (scan_seconds) *scan_seconds = now_seconds() - t0;
  if (rc == 0) memcpy(rss, outs, rows * cols * sizeof(float)); /* every scan's input is the same */
  free(ws);
  free(ids);
  free(outs);
  fftplan_free(&fr);
  fftplan_free(&fc);
  return rc;
}

/* ------------------------------------------------------------------ 3-D kooshball (cfg4)
 * The definition of ora_kb_adjoint3 (origin (n-1)/2, |k|^2 DCF from ora_ramp_sq_dcf3) in
 * FP32 on the n^3 volume's 2n^3 grid: per-sample tables (cells from the float coordinate,
 * taps and the DCF-phase factor in FP64 rounded to float, samples in cell order so the
 * scatter walks the grid in order) made once; per coil, KB scatter, radix-2 inverse FFT
 * along x of every grid row, along y of the n kept x columns and along z of the n^2 kept
 * (y, x) lines (strided columns gathered eight at a time, a cache line), deapodization,
 * and |x|^2 added into the RSS.  Threads take whole coils. */
typedef struct {
  size_t n, g, samples, width;
  uint32_t* order;    /* [S] samples in (z, y, x) cell order */
  uint32_t* first;    /* [S][3] first tap index per axis (wrapped), in `order` */
  float* taps;        /* [S][3W] z, y, x taps, in `order` */
  cf* factor;         /* [S] dcf * exp(+i phase), in `order` */
  float* inv_deapod;  /* [n] per axis: 1 / phi_hat(j - n/2) (the grid is cubic) */
  fftplan f;
} route3;

static void route3_free(route3* p) {
  free(p->order);
  free(p->first);
  free(p->taps);
  free(p->factor);
  free(p->inv_deapod);
  fftplan_free(&p->f);
}

/* Per-sample tables of samples [begin, end) of `order` (one thread's share). */
typedef struct {
  route3* p;
  const float* traj;
  const uint32_t* cells;
  size_t begin, end;
} fill3;

static void* fill3_run(void* arg) {
  fill3* f = (fill3*)arg;
  route3* p = f->p;
  const size_t g = p->g, W = p->width, n = p->n;
  const int width = (int)W;
  const double beta = ora_kb_beta(width, 2.0);
  const int lead = width / 2 - 1;
  const double shift = (double)(n / 2) - ((double)n - 1.0) / 2.0;
  const double scale = (double)g / (2.0 * 3.14159265358979323846);
  for (size_t i = f->begin; i < f->end; ++i) {
    const size_t s = p->order[i];
    double phase = 0.0, k2 = 0.0;
    for (int a = 0; a < 3; ++a) {
      const double w = f->traj[3 * s + a];
      double u = w * scale;
      if (u < 0.0) u += (double)g;
      if (u >= (double)g) u -= (double)g;
      const double d = u - (double)f->cells[3 * s + a];
      for (size_t t = 0; t < W; ++t)
        p->taps[(i * 3 + (size_t)a) * W + t] = (float)ora_kb_kernel((double)((int)t - lead) - d, width, beta);
      p->first[3 * i + (size_t)a] = (uint32_t)((f->cells[3 * s + a] + g - (size_t)lead) % g);
      phase += w * shift;
      k2 += w * w;
    }
    p->factor[i].re = (float)(k2 * cos(phase));
    p->factor[i].im = (float)(k2 * sin(phase));
  }
  return NULL;
}

static int route3_make(route3* p, size_t n, size_t samples, const float* traj, int width, int threads) {
  memset(p, 0, sizeof(*p));
  p->n = n;
  p->g = 2 * n;
  p->samples = samples;
  p->width = (size_t)width;
  const size_t g = p->g, W = (size_t)width;
  if (log2_exact(g) < 1 || (width != 4 && width != 6 && width != 8) || samples >= ((size_t)1 << 32)) return -1;
  p->order = (uint32_t*)malloc(samples * sizeof(uint32_t));
  p->first = (uint32_t*)malloc(samples * 3 * sizeof(uint32_t));
  p->taps = (float*)malloc(samples * 3 * W * sizeof(float));
  p->factor = (cf*)malloc(samples * sizeof(cf));
  p->inv_deapod = (float*)malloc(n * sizeof(float));
  uint32_t* cells = (uint32_t*)malloc(samples * 3 * sizeof(uint32_t));
  size_t* start = (size_t*)calloc(g * g + 1, sizeof(size_t));
  int rc = (p->order && p->first && p->taps && p->factor && p->inv_deapod && cells && start && !fftplan_make(&p->f, g))
               ? 0
               : -1;
  /* cells from the float coordinate; samples in grid-row (z, y) order by a counting sort
   * (stable, so each row keeps the acquisition order) */
  for (size_t s = 0; rc == 0 && s < samples; ++s) {
    float f;
    for (int a = 0; a < 3; ++a)
      if (ora_grid_coordinate(traj[3 * s + a], (uint32_t)g, &cells[3 * s + a], &f)) rc = -1;
    if (rc == 0) ++start[(size_t)cells[3 * s] * g + cells[3 * s + 1] + 1];
  }
  if (rc == 0) {
    for (size_t r = 0; r < g * g; ++r) start[r + 1] += start[r];
    for (size_t s = 0; s < samples; ++s) p->order[start[(size_t)cells[3 * s] * g + cells[3 * s + 1]]++] = (uint32_t)s;
    const size_t t = threads < 1 ? 1 : (size_t)threads;
    fill3* fs = (fill3*)calloc(t, sizeof(fill3));
    pthread_t* ids = (pthread_t*)calloc(t, sizeof(pthread_t));
    size_t started = 0;
    if (!fs || !ids) rc = -1;
    for (size_t i = 0; rc == 0 && i < t; ++i) {
      fs[i] = (fill3){p, traj, cells, samples * i / t, samples * (i + 1) / t};
      if (pthread_create(&ids[i], NULL, fill3_run, &fs[i]) != 0) rc = -1;
      else ++started;
    }
    for (size_t i = 0; i < started; ++i) pthread_join(ids[i], NULL);
    free(fs);
    free(ids);
  }
  if (rc == 0) {
    const double beta = ora_kb_beta(width, 2.0);
    for (size_t j = 0; j < n; ++j)
      p->inv_deapod[j] = (float)(1.0 / ora_kb_ft((double)j - (double)(n / 2), g, width, beta));
  }
  free(cells);
  free(start);
  if (rc != 0) route3_free(p);
  return rc;
}

/* FFT of the 8 columns at `base` (element k of column q at base[k * stride + q]) through
 * `buf` [g][8], written back in place. */
static void fft_columns8(const fftplan* f, cf* base, size_t stride, size_t count, cf* buf) {
  const size_t g = f->g;
  for (size_t k = 0; k < g; ++k) memcpy(buf + 8 * k, base + k * stride, count * sizeof(cf));
  for (size_t q = 0; q < count; ++q) fft_inplace(f, buf + q, 8);
  for (size_t k = 0; k < g; ++k) memcpy(base + k * stride, buf + 8 * k, count * sizeof(cf));
}

typedef struct {
  const route3* p;
  const float* values;  /* [coils][S] */
  size_t first, coils, step;
  float* sumsq;         /* [n^3] this thread's sum of squares */
  int failed;
} worker3;

static void* run3(void* arg) {
  worker3* w = (worker3*)arg;
  const route3* p = w->p;
  const size_t n = p->n, g = p->g, W = p->width, S = p->samples, mask = g - 1;
  cf* grid = (cf*)malloc(g * g * g * sizeof(cf));
  cf* buf = (cf*)malloc(8 * g * sizeof(cf));
  if (!grid || !buf) {
    w->failed = 1;
    free(grid);
    free(buf);
    return NULL;
  }
  for (size_t i = 0; i < n * n * n; ++i) w->sumsq[i] = 0.0f;
  for (size_t coil = w->first; coil < w->coils; coil += w->step) {
    memset(grid, 0, g * g * g * sizeof(cf));
    const cf* y = (const cf*)w->values + coil * S;
    for (size_t i = 0; i < S; ++i) {
      const cf v = y[p->order[i]], f = p->factor[i];
      const cf z = {v.re * f.re - v.im * f.im, v.re * f.im + v.im * f.re};
      const float* tz = p->taps + i * 3 * W;
      const float* ty = tz + W;
      const float* tx = ty + W;
      const size_t z0 = p->first[3 * i], y0 = p->first[3 * i + 1], x0 = p->first[3 * i + 2];
      for (size_t a = 0; a < W; ++a) {
        const size_t mz = (z0 + a) & mask;
        for (size_t b = 0; b < W; ++b) {
          cf* row = grid + (mz * g + ((y0 + b) & mask)) * g;
          const float wzy = tz[a] * ty[b];
          const float zr = z.re * wzy, zi = z.im * wzy;
          for (size_t c = 0; c < W; ++c) {
            cf* cell = row + ((x0 + c) & mask);
            cell->re += zr * tx[c];
            cell->im += zi * tx[c];
          }
        }
      }
    }
    /* x along every grid row */
    for (size_t r = 0; r < g * g; ++r) fft_inplace(&p->f, grid + r * g, 1);
    /* y along the n kept x columns of every z plane, eight columns at a time: kept x
     * index j reads grid column (j - n/2) mod g, i.e. columns [g - n/2, g) then [0, n/2) */
    for (size_t z = 0; z < g; ++z)
      for (size_t part = 0; part < 2; ++part) {
        const size_t x0 = part ? 0 : g - n / 2;
        for (size_t x = 0; x < n / 2; x += 8)
          fft_columns8(&p->f, grid + z * g * g + x0 + x, g, n / 2 - x < 8 ? n / 2 - x : 8, buf);
      }
    /* z along the n^2 kept (y, x) lines; output (z, y, x) reads grid ((z, y, x) - n/2) mod g */
    for (size_t yy = 0; yy < n; ++yy) {
      const size_t iy = (yy + g - n / 2) % g;
      for (size_t part = 0; part < 2; ++part) {
        const size_t xb = part ? 0 : g - n / 2, ob = part ? n / 2 : 0;
        for (size_t x = 0; x < n / 2; x += 8) {
          const size_t cnt = n / 2 - x < 8 ? n / 2 - x : 8;
          fft_columns8(&p->f, grid + iy * g + xb + x, g * g, cnt, buf);
          for (size_t zz = 0; zz < n; ++zz) {
            const size_t iz = (zz + g - n / 2) % g;
            const cf* src = grid + (iz * g + iy) * g + xb + x;
            const float dzy = p->inv_deapod[zz] * p->inv_deapod[yy];
            float* out = w->sumsq + (zz * n + yy) * n + ob + x;
            for (size_t q = 0; q < cnt; ++q) {
              const float d = dzy * p->inv_deapod[ob + x + q];
              const float re = src[q].re * d, im = src[q].im * d;
              out[q] += re * re + im * im;
            }
          }
        }
      }
    }
  }
  free(grid);
  free(buf);
  return NULL;
}

int ora_kb_route3_rss_f32(size_t n, size_t samples, size_t coils, const float* traj, const float* values, int width,
                          int threads, float* rss, double* coil_seconds) {
  route3 p;
  if (n < 16 || coils == 0 || threads < 1) return -1;
  if (route3_make(&p, n, samples, traj, width, threads)) return -1;
  const size_t t = (size_t)threads < coils ? (size_t)threads : coils, v = n * n * n;
  worker3* ws = (worker3*)calloc(t, sizeof(worker3));
  pthread_t* ids = (pthread_t*)calloc(t, sizeof(pthread_t));
  float* sums = (float*)malloc(t * v * sizeof(float));
  int rc = (ws && ids && sums) ? 0 : -1;
  const double t0 = now_seconds();
  size_t started = 0;
  for (size_t i = 0; rc == 0 && i < t; ++i) {
    ws[i] = (worker3){&p, values, i, coils, t, sums + i * v, 0};
    if (pthread_create(&ids[i], NULL, run3, &ws[i]) != 0) rc = -1;
    else ++started;
  }
  for (size_t i = 0; i < started; ++i) {
    pthread_join(ids[i], NULL);
    if (ws[i].failed) rc = -1;
  }
  if (coil_seconds) *coil_seconds = now_seconds() - t0;
  if (rc == 0)
    for (size_t i = 0; i < v; ++i) {
      float s = 0.0f;
      for (size_t k = 0; k < t; ++k) s += sums[k * v + i];
      rss[i] = sqrtf(s);
    }
  free(ws);
  free(ids);
  free(sums);
  route3_free(&p);
  return rc;
}

int ora_assemble_acquisitions_f32(const float* packed, const unsigned* counts, size_t acquisitions, size_t channels,
                                  size_t frames, float* frame, double* seconds) {
  if (acquisitions == 0 || channels == 0 || channels > 64 || !packed || !counts || !frame) return -1;
  size_t total = 0;
  for (size_t a = 0; a < acquisitions; ++a) {
    if (counts[a] == 0) return -1;
    total += counts[a];
  }
  const double t0 = now_seconds();
  for (size_t f = 0; f < frames; ++f) {
    size_t written = 0;        /* samples of the acquisitions before this one */
    const float* src = packed; /* this acquisition's block */
    for (size_t a = 0; a < acquisitions; ++a) {
      const size_t n = counts[a];
      for (size_t c = 0; c < channels; ++c)
        memcpy(frame + 2 * (c * total + written), src + 2 * c * n, n * 2 * sizeof(float));
      src += 2 * channels * n;
      written += n;
    }
  }
  if (seconds) *seconds = now_seconds() - t0;
  return 0;
}
