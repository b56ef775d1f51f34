/*
 * FP64 CPU oracle for the KSpaceJet reconstruction hot path.
 *
 * TEST INFRASTRUCTURE ONLY: a plain restatement of the reference algorithms
 * used to check the HIP path.  The product never links or calls it.
 * Built with -ffp-contract=off so the FP64 binning arithmetic is the exact
 * IEEE sequence the GPU binning kernel must reproduce.
 */
#include "ksj_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define ORA_PI 3.14159265358979323846
#define ORA_TWO_PI 6.28318530717958647692

/* ---------------------------------------------------------------- direct NUDFT */

void ora_nudft2_forward(size_t rows, size_t cols, double row_origin, double col_origin, size_t samples,
                        const double* traj, const double* image, double* out) {
  /* eigen_nufft_direct_nudft.cpp:52-74: y_s = sum_{r,c} x(r,c) exp(-i(w_r (r-r0) + w_c (c-c0))) */
  for (size_t s = 0; s < samples; ++s) {
    const double wr = traj[2 * s], wc = traj[2 * s + 1];
    double acc_re = 0.0, acc_im = 0.0;
    for (size_t r = 0; r < rows; ++r) {
      const double rp = (double)r - row_origin;
      for (size_t c = 0; c < cols; ++c) {
        const double cp = (double)c - col_origin;
        const double phase = -(wr * rp + wc * cp);
        const double cs = cos(phase), sn = sin(phase);
        const double xr = image[2 * (r * cols + c)], xi = image[2 * (r * cols + c) + 1];
        acc_re += xr * cs - xi * sn;
        acc_im += xr * sn + xi * cs;
      }
    }
    out[2 * s] = acc_re;
    out[2 * s + 1] = acc_im;
  }
}

void ora_nudft2_adjoint(size_t rows, size_t cols, double row_origin, double col_origin, size_t samples,
                        const double* traj, const double* values, double* image) {
  /* eigen_nufft_direct_nudft.cpp:77-97: x(r,c) = sum_s y_s exp(+i(w_r (r-r0) + w_c (c-c0))) */
  for (size_t r = 0; r < rows; ++r) {
    const double rp = (double)r - row_origin;
    for (size_t c = 0; c < cols; ++c) {
      const double cp = (double)c - col_origin;
      double acc_re = 0.0, acc_im = 0.0;
      for (size_t s = 0; s < samples; ++s) {
        const double phase = traj[2 * s] * rp + traj[2 * s + 1] * cp;
        const double cs = cos(phase), sn = sin(phase);
        const double yr = values[2 * s], yi = values[2 * s + 1];
        acc_re += yr * cs - yi * sn;
        acc_im += yr * sn + yi * cs;
      }
      image[2 * (r * cols + c)] = acc_re;
      image[2 * (r * cols + c) + 1] = acc_im;
    }
  }
}

void ora_ramp_dcf2(size_t samples, const double* traj, double* dcf) {
  /* eigen_radial_gridding.cpp:279-295 */
  for (size_t s = 0; s < samples; ++s) dcf[s] = hypot(traj[2 * s], traj[2 * s + 1]);
}

/* ---------------------------------------------------------------- FFT */

static size_t smallest_factor(size_t n) {
  static const size_t preferred[] = {4, 2, 3, 5};
  for (size_t i = 0; i < 4; ++i)
    if (n % preferred[i] == 0) return preferred[i];
  for (size_t p = 7; p * p <= n; p += 2)
    if (n % p == 0) return p;
  return n;
}

void ora_fft1(size_t n, double* data, int sign) {
  if (n <= 1) return;
  double* a = data;
  double* b = (double*)malloc(2 * n * sizeof(double));
  double* v = (double*)malloc(2 * n * sizeof(double));
  double* w = (double*)malloc(2 * n * sizeof(double));
  /* roots of unity e^{sign 2 pi i m / n}; every twiddle below is one of them */
  double* tw = (double*)malloc(2 * n * sizeof(double));
  for (size_t m = 0; m < n; ++m) {
    const double angle = sign * ORA_TWO_PI * (double)m / (double)n;
    tw[2 * m] = cos(angle);
    tw[2 * m + 1] = sin(angle);
  }
  size_t ns = 1;
  size_t rest = n;
  while (rest > 1) {
    const size_t radix = smallest_factor(rest);
    rest /= radix;
    const size_t stride = n / radix;
    const size_t tw_step = n / (ns * radix);
    for (size_t j = 0; j < stride; ++j) {
      const size_t k = j % ns;
      for (size_t r = 0; r < radix; ++r) {
        const size_t m = k * r * tw_step;
        const double cs = tw[2 * m], sn = tw[2 * m + 1];
        const double xr = a[2 * (j + r * stride)], xi = a[2 * (j + r * stride) + 1];
        v[2 * r] = xr * cs - xi * sn;
        v[2 * r + 1] = xr * sn + xi * cs;
      }
      for (size_t q = 0; q < radix; ++q) {
        double sr = 0.0, si = 0.0;
        for (size_t r = 0; r < radix; ++r) {
          const size_t m = ((r * q) % radix) * stride;
          const double cs = tw[2 * m], sn = tw[2 * m + 1];
          sr += v[2 * r] * cs - v[2 * r + 1] * sn;
          si += v[2 * r] * sn + v[2 * r + 1] * cs;
        }
        w[2 * q] = sr;
        w[2 * q + 1] = si;
      }
      const size_t base = (j / ns) * ns * radix + k;
      for (size_t q = 0; q < radix; ++q) {
        b[2 * (base + q * ns)] = w[2 * q];
        b[2 * (base + q * ns) + 1] = w[2 * q + 1];
      }
    }
    double* t = a;
    a = b;
    b = t;
    ns *= radix;
  }
  if (a != data) {
    memcpy(data, a, 2 * n * sizeof(double));
    free(a);
  } else {
    free(b);
  }
  free(v);
  free(w);
  free(tw);
}

/* Unscaled +i inverse DFT over rows then columns of a rows x cols interleaved matrix. */
static void inverse_dft2_unscaled(size_t rows, size_t cols, double* data) {
  for (size_t r = 0; r < rows; ++r) ora_fft1(cols, data + 2 * r * cols, +1);
  double* line = (double*)malloc(2 * rows * sizeof(double));
  for (size_t c = 0; c < cols; ++c) {
    for (size_t r = 0; r < rows; ++r) {
      line[2 * r] = data[2 * (r * cols + c)];
      line[2 * r + 1] = data[2 * (r * cols + c) + 1];
    }
    ora_fft1(rows, line, +1);
    for (size_t r = 0; r < rows; ++r) {
      data[2 * (r * cols + c)] = line[2 * r];
      data[2 * (r * cols + c) + 1] = line[2 * r + 1];
    }
  }
  free(line);
}

static void circular_shift2(size_t rows, size_t cols, double* data, size_t shift_r, size_t shift_c) {
  double* tmp = (double*)malloc(2 * rows * cols * sizeof(double));
  for (size_t r = 0; r < rows; ++r)
    for (size_t c = 0; c < cols; ++c) {
      const size_t rr = (r + shift_r) % rows, cc = (c + shift_c) % cols;
      tmp[2 * (rr * cols + cc)] = data[2 * (r * cols + c)];
      tmp[2 * (rr * cols + cc) + 1] = data[2 * (r * cols + c) + 1];
    }
  memcpy(data, tmp, 2 * rows * cols * sizeof(double));
  free(tmp);
}

void ora_ifft2(size_t rows, size_t cols, double* data, int centred) {
  /* eigen_fft_transforms.cpp:148-205 with Normalization::inverse (:14-27): 1/cols then 1/rows */
  if (centred) circular_shift2(rows, cols, data, rows - rows / 2, cols - cols / 2); /* ifftshift */
  inverse_dft2_unscaled(rows, cols, data);
  const double scale = 1.0 / ((double)rows * (double)cols);
  for (size_t i = 0; i < 2 * rows * cols; ++i) data[i] *= scale;
  if (centred) circular_shift2(rows, cols, data, rows / 2, cols / 2); /* fftshift */
}

/* ---------------------------------------------------------------- radial bilinear route */

static int reference_grid_coordinate(double w, size_t extent, size_t* lower, double* fraction) {
  /* eigen_radial_gridding.cpp:262-277 */
  const double e = (double)extent;
  double u = w * e / (2.0 * ORA_PI);
  if (u < 0.0) u += e;
  if (u >= e) u -= e;
  if (!isfinite(u) || u < 0.0 || u >= e) return -1;
  *lower = (size_t)floor(u);
  *fraction = u - (double)*lower;
  return 0;
}

int ora_radial_linear_gridding2(size_t rows, size_t cols, double row_origin, double col_origin, size_t samples,
                                const double* values, const double* traj, const double* dcf, double* image) {
  /* The payload is float32, checked against float pi (radial_gridding_reconstruct.cpp:217-226). */
  const double pi_f = (double)(float)ORA_PI;
  for (size_t s = 0; s < samples; ++s)
    if (!(traj[2 * s] >= -pi_f && traj[2 * s] <= pi_f && traj[2 * s + 1] >= -pi_f && traj[2 * s + 1] <= pi_f))
      return -1;
  memset(image, 0, 2 * rows * cols * sizeof(double));
  for (size_t s = 0; s < samples; ++s) {
    const double wr = traj[2 * s], wc = traj[2 * s + 1];
    size_t rl, cl;
    double rf, cf;
    if (reference_grid_coordinate(wr, rows, &rl, &rf) || reference_grid_coordinate(wc, cols, &cl, &cf)) return -1;
    const size_t ru = (rl + 1) % rows, cu = (cl + 1) % cols;
    const double phase = -(wr * row_origin + wc * col_origin);
    const double pr = cos(phase), pi = sin(phase);
    const double yr = values[2 * s] * dcf[s], yi = values[2 * s + 1] * dcf[s];
    const double zr = yr * pr - yi * pi, zi = yr * pi + yi * pr;
    const double wts[4] = {(1.0 - rf) * (1.0 - cf), (1.0 - rf) * cf, rf * (1.0 - cf), rf * cf};
    const size_t idx[4] = {rl * cols + cl, rl * cols + cu, ru * cols + cl, ru * cols + cu};
    for (int t = 0; t < 4; ++t) {
      image[2 * idx[t]] += zr * wts[t];
      image[2 * idx[t] + 1] += zi * wts[t];
    }
  }
  inverse_dft2_unscaled(rows, cols, image);
  return 0;
}

/* ---------------------------------------------------------------- coil combine */

void ora_rss(size_t pixels, size_t coils, const double* coil_last, double* out) {
  for (size_t p = 0; p < pixels; ++p) {
    double acc = 0.0;
    for (size_t c = 0; c < coils; ++c) {
      const double re = coil_last[2 * (p * coils + c)], im = coil_last[2 * (p * coils + c) + 1];
      acc += re * re + im * im;
    }
    out[p] = sqrt(acc);
  }
}

void ora_sense_combine(size_t pixels, size_t coils, const double* images, const double* sens, double* out) {
  for (size_t p = 0; p < pixels; ++p) {
    double nr = 0.0, ni = 0.0, den = 0.0;
    for (size_t c = 0; c < coils; ++c) {
      const double ir = images[2 * (p * coils + c)], ii = images[2 * (p * coils + c) + 1];
      const double sr = sens[2 * (p * coils + c)], si = sens[2 * (p * coils + c) + 1];
      nr += sr * ir + si * ii; /* conj(S) * I */
      ni += sr * ii - si * ir;
      den += sr * sr + si * si;
    }
    out[2 * p] = den > 0.0 ? nr / den : 0.0;
    out[2 * p + 1] = den > 0.0 ? ni / den : 0.0;
  }
}

/* ---------------------------------------------------------------- Kaiser-Bessel */

double ora_i0(double x) {
  /* Power series sum_k ((x/2)^k / k!)^2, summed until terms stop changing the sum. */
  const double q = 0.25 * x * x;
  double term = 1.0, sum = 1.0;
  for (int k = 1; k < 500; ++k) {
    term *= q / ((double)k * (double)k);
    sum += term;
    if (term < sum * 1e-17) break;
  }
  return sum;
}

double ora_kb_beta(int width, double sigma) {
  /* Beatty, Nishimura, Pauly, IEEE TMI 24(6) 2005, eq. 5 */
  const double a = (double)width / sigma * (sigma - 0.5);
  return ORA_PI * sqrt(a * a - 0.8);
}

double ora_kb_kernel(double distance, int width, double beta) {
  const double half = 0.5 * (double)width;
  if (!(fabs(distance) < half)) return 0.0;
  const double t = distance / half;
  return ora_i0(beta * sqrt(1.0 - t * t)) / ora_i0(beta);
}

double ora_kb_ft(double index, size_t grid, int width, double beta) {
  /* Continuous FT of ora_kb_kernel at angular frequency 2 pi index / grid:
   * (W / I0(beta)) sinh(z)/z with z^2 = beta^2 - (pi W index / grid)^2 (sin for z^2 < 0). */
  const double a = ORA_PI * (double)width * index / (double)grid;
  const double z2 = beta * beta - a * a;
  double shape;
  if (z2 > 0.0) {
    const double z = sqrt(z2);
    shape = sinh(z) / z;
  } else if (z2 < 0.0) {
    const double z = sqrt(-z2);
    shape = sin(z) / z;
  } else {
    shape = 1.0;
  }
  return (double)width * shape / ora_i0(beta);
}

int ora_grid_coordinate(float w, uint32_t extent, uint32_t* cell, float* frac) {
  const double e = (double)extent;
  double u = ((double)w * e) / ORA_TWO_PI;
  if (u < 0.0) u += e;
  if (u >= e) u -= e;
  if (!(u >= 0.0 && u < e)) return -1;
  const double f = floor(u);
  *cell = (uint32_t)f;
  *frac = (float)(u - f);
  return 0;
}

int ora_bin2(size_t samples, const float* traj, uint32_t grid_rows, uint32_t grid_cols, uint32_t* bins,
             float* frac, uint32_t* cell_start, uint32_t* perm) {
  const size_t nbins = (size_t)grid_rows * grid_cols;
  memset(cell_start, 0, (nbins + 1) * sizeof(uint32_t));
  for (size_t s = 0; s < samples; ++s) {
    uint32_t cr, cc;
    if (ora_grid_coordinate(traj[2 * s], grid_rows, &cr, &frac[2 * s]) ||
        ora_grid_coordinate(traj[2 * s + 1], grid_cols, &cc, &frac[2 * s + 1]))
      return -1;
    bins[s] = cr * grid_cols + cc;
    cell_start[bins[s] + 1] += 1;
  }
  for (size_t b = 0; b < nbins; ++b) cell_start[b + 1] += cell_start[b];
  uint32_t* cursor = (uint32_t*)malloc(nbins * sizeof(uint32_t));
  memcpy(cursor, cell_start, nbins * sizeof(uint32_t));
  for (size_t s = 0; s < samples; ++s) perm[cursor[bins[s]]++] = (uint32_t)s;
  free(cursor);
  return 0;
}

int ora_kb_adjoint2(size_t n_rows, size_t n_cols, size_t g_rows, size_t g_cols, int width, double beta,
                    double row_origin, double col_origin, size_t samples, const double* traj, const double* values,
                    const double* dcf, double* image) {
  double* grid = (double*)calloc(2 * g_rows * g_cols, sizeof(double));
  const double shift_r = (double)(n_rows / 2) - row_origin;
  const double shift_c = (double)(n_cols / 2) - col_origin;
  const int lead = width / 2 - 1;
  double* wr_tap = (double*)malloc((size_t)width * sizeof(double));
  double* wc_tap = (double*)malloc((size_t)width * sizeof(double));
  for (size_t s = 0; s < samples; ++s) {
    const double wr = traj[2 * s], wc = traj[2 * s + 1];
    const double weight = dcf ? dcf[s] : 1.0;
    const double phase = wr * shift_r + wc * shift_c;
    const double pr = cos(phase), pi = sin(phase);
    const double yr = values[2 * s] * weight, yi = values[2 * s + 1] * weight;
    const double zr = yr * pr - yi * pi, zi = yr * pi + yi * pr;
    uint32_t cr, cc;
    float fr, fc;
    if (ora_grid_coordinate((float)wr, (uint32_t)g_rows, &cr, &fr) ||
        ora_grid_coordinate((float)wc, (uint32_t)g_cols, &cc, &fc)) {
      free(grid);
      free(wr_tap);
      free(wc_tap);
      return -1;
    }
    /* The cell is taken from the float-rounded coordinate rule above (the GPU bins the
     * float32 payload); the tap distances use the exact FP64 coordinate. */
    double ur = wr * (double)g_rows / ORA_TWO_PI, uc = wc * (double)g_cols / ORA_TWO_PI;
    if (ur < 0.0) ur += (double)g_rows;
    if (ur >= (double)g_rows) ur -= (double)g_rows;
    if (uc < 0.0) uc += (double)g_cols;
    if (uc >= (double)g_cols) uc -= (double)g_cols;
    const double dr0 = ur - (double)cr, dc0 = uc - (double)cc;
    for (int t = 0; t < width; ++t) {
      wr_tap[t] = ora_kb_kernel((double)(t - lead) - dr0, width, beta);
      wc_tap[t] = ora_kb_kernel((double)(t - lead) - dc0, width, beta);
    }
    for (int ty = 0; ty < width; ++ty) {
      const size_t mr = ((size_t)cr + g_rows + (size_t)ty - (size_t)lead) % g_rows;
      for (int tx = 0; tx < width; ++tx) {
        const size_t mc = ((size_t)cc + g_cols + (size_t)tx - (size_t)lead) % g_cols;
        const double w = wr_tap[ty] * wc_tap[tx];
        grid[2 * (mr * g_cols + mc)] += zr * w;
        grid[2 * (mr * g_cols + mc) + 1] += zi * w;
      }
    }
  }
  inverse_dft2_unscaled(g_rows, g_cols, grid);
  for (size_t r = 0; r < n_rows; ++r) {
    const long jr = (long)r - (long)(n_rows / 2);
    const size_t ir = (size_t)((jr + (long)g_rows) % (long)g_rows);
    const double dr = ora_kb_ft((double)jr, g_rows, width, beta);
    for (size_t c = 0; c < n_cols; ++c) {
      const long jc = (long)c - (long)(n_cols / 2);
      const size_t ic = (size_t)((jc + (long)g_cols) % (long)g_cols);
      const double d = dr * ora_kb_ft((double)jc, g_cols, width, beta);
      image[2 * (r * n_cols + c)] = grid[2 * (ir * g_cols + ic)] / d;
      image[2 * (r * n_cols + c) + 1] = grid[2 * (ir * g_cols + ic) + 1] / d;
    }
  }
  free(grid);
  free(wr_tap);
  free(wc_tap);
  return 0;
}

void ora_golden_angle2(size_t spokes, size_t readout, double* traj) {
  const double golden = ORA_PI * (sqrt(5.0) - 1.0) / 2.0; /* pi / phi = 111.246 degrees */
  for (size_t p = 0; p < spokes; ++p) {
    const double angle = (double)p * golden;
    const double cs = cos(angle), sn = sin(angle);
    for (size_t j = 0; j < readout; ++j) {
      const double k = ORA_TWO_PI * ((double)j - (double)(readout / 2)) / (double)readout;
      traj[2 * (p * readout + j)] = k * sn;
      traj[2 * (p * readout + j) + 1] = k * cs;
    }
  }
}

/* ---------------------------------------------------------------- 3-D (kooshball) */
/* The reference has no 3-D path; these extend its 2-D definitions axis by axis:
 * coordinates [S][3] in (z, y, x) order, radians per pixel, adjoint sign +i
 * (eigen_nufft_direct_nudft.cpp:77-97 with a third axis), origins per axis. */

void ora_nudft3_adjoint(size_t nz, size_t ny, size_t nx, double oz, double oy, double ox, size_t samples,
                        const double* traj, const double* values, double* image) {
  for (size_t z = 0; z < nz; ++z)
    for (size_t y = 0; y < ny; ++y)
      for (size_t x = 0; x < nx; ++x) {
        const double pz = (double)z - oz, py = (double)y - oy, px = (double)x - ox;
        double acc_re = 0.0, acc_im = 0.0;
        for (size_t s = 0; s < samples; ++s) {
          const double phase = traj[3 * s] * pz + traj[3 * s + 1] * py + traj[3 * s + 2] * px;
          const double cs = cos(phase), sn = sin(phase);
          const double yr = values[2 * s], yi = values[2 * s + 1];
          acc_re += yr * cs - yi * sn;
          acc_im += yr * sn + yi * cs;
        }
        image[2 * ((z * ny + y) * nx + x)] = acc_re;
        image[2 * ((z * ny + y) * nx + x) + 1] = acc_im;
      }
}

void ora_nudft3_forward(size_t nz, size_t ny, size_t nx, double oz, double oy, double ox, size_t samples,
                        const double* traj, const double* image, double* values) {
  /* y_s = sum_v x_v exp(-i w_s . (v - o)): the exact forward model the 3-D data follow */
  for (size_t s = 0; s < samples; ++s) {
    double acc_re = 0.0, acc_im = 0.0;
    for (size_t z = 0; z < nz; ++z)
      for (size_t y = 0; y < ny; ++y)
        for (size_t x = 0; x < nx; ++x) {
          const double phase = -(traj[3 * s] * ((double)z - oz) + traj[3 * s + 1] * ((double)y - oy) +
                                 traj[3 * s + 2] * ((double)x - ox));
          const double cs = cos(phase), sn = sin(phase);
          const double xr = image[2 * ((z * ny + y) * nx + x)], xi = image[2 * ((z * ny + y) * nx + x) + 1];
          acc_re += xr * cs - xi * sn;
          acc_im += xr * sn + xi * cs;
        }
    values[2 * s] = acc_re;
    values[2 * s + 1] = acc_im;
  }
}

void ora_ramp_sq_dcf3(size_t samples, const double* traj, double* dcf) {
  /* |k|^2 in (radians per pixel)^2: the density of a 3-D radial trajectory falls as 1/|k|^2 */
  for (size_t s = 0; s < samples; ++s)
    dcf[s] = traj[3 * s] * traj[3 * s] + traj[3 * s + 1] * traj[3 * s + 1] + traj[3 * s + 2] * traj[3 * s + 2];
}

int ora_bin3(size_t samples, const float* traj, uint32_t gz, uint32_t gy, uint32_t gx, uint32_t* bins,
             uint32_t* cell_start, uint32_t* perm) {
  const size_t nbins = (size_t)gz * gy * gx;
  memset(cell_start, 0, (nbins + 1) * sizeof(uint32_t));
  for (size_t s = 0; s < samples; ++s) {
    uint32_t cz, cy, cx;
    float f;
    if (ora_grid_coordinate(traj[3 * s], gz, &cz, &f) || ora_grid_coordinate(traj[3 * s + 1], gy, &cy, &f) ||
        ora_grid_coordinate(traj[3 * s + 2], gx, &cx, &f))
      return -1;
    bins[s] = (cz * gy + cy) * gx + cx;
    cell_start[bins[s] + 1] += 1;
  }
  for (size_t b = 0; b < nbins; ++b) cell_start[b + 1] += cell_start[b];
  uint32_t* cursor = (uint32_t*)malloc(nbins * sizeof(uint32_t));
  memcpy(cursor, cell_start, nbins * sizeof(uint32_t));
  for (size_t s = 0; s < samples; ++s) perm[cursor[bins[s]]++] = (uint32_t)s;
  free(cursor);
  return 0;
}

/* Unscaled +i inverse DFT along each axis of a gz x gy x gx interleaved volume. */
static void inverse_dft3_unscaled(size_t gz, size_t gy, size_t gx, double* data) {
  const size_t nmax = gz > gy ? (gz > gx ? gz : gx) : (gy > gx ? gy : gx);
  double* line = (double*)malloc(2 * nmax * sizeof(double));
  for (size_t z = 0; z < gz; ++z)
    for (size_t y = 0; y < gy; ++y) ora_fft1(gx, data + 2 * (z * gy + y) * gx, +1);
  for (size_t z = 0; z < gz; ++z)
    for (size_t x = 0; x < gx; ++x) {
      for (size_t y = 0; y < gy; ++y) {
        line[2 * y] = data[2 * ((z * gy + y) * gx + x)];
        line[2 * y + 1] = data[2 * ((z * gy + y) * gx + x) + 1];
      }
      ora_fft1(gy, line, +1);
      for (size_t y = 0; y < gy; ++y) {
        data[2 * ((z * gy + y) * gx + x)] = line[2 * y];
        data[2 * ((z * gy + y) * gx + x) + 1] = line[2 * y + 1];
      }
    }
  for (size_t y = 0; y < gy; ++y)
    for (size_t x = 0; x < gx; ++x) {
      for (size_t z = 0; z < gz; ++z) {
        line[2 * z] = data[2 * ((z * gy + y) * gx + x)];
        line[2 * z + 1] = data[2 * ((z * gy + y) * gx + x) + 1];
      }
      ora_fft1(gz, line, +1);
      for (size_t z = 0; z < gz; ++z) {
        data[2 * ((z * gy + y) * gx + x)] = line[2 * z];
        data[2 * ((z * gy + y) * gx + x) + 1] = line[2 * z + 1];
      }
    }
  free(line);
}

/* Grid coordinate of an FP64 frequency on an axis of `extent` points (same wrap as the binning). */
static double wrapped_coordinate(double w, size_t extent) {
  double u = w * (double)extent / ORA_TWO_PI;
  if (u < 0.0) u += (double)extent;
  if (u >= (double)extent) u -= (double)extent;
  return u;
}

int ora_kb_adjoint3(size_t nz, size_t ny, size_t nx, size_t gz, size_t gy, size_t gx, int width, double beta,
                    double oz, double oy, double ox, size_t samples, const double* traj, const double* values,
                    const double* dcf, double* image) {
  /* KB gridding onto the (gz, gy, gx) grid, unscaled inverse DFT, crop of the centred
   * n-cube (output index i reads grid frequency i - n/2), division by the separable
   * KB transform: the 2-D route of ora_kb_adjoint2 with a third axis. */
  double* grid = (double*)calloc(2 * gz * gy * gx, sizeof(double));
  const double sz = (double)(nz / 2) - oz, sy = (double)(ny / 2) - oy, sx = (double)(nx / 2) - ox;
  const int lead = width / 2 - 1;
  double tz[16], ty[16], tx[16];
  for (size_t s = 0; s < samples; ++s) {
    const double wz = traj[3 * s], wy = traj[3 * s + 1], wx = traj[3 * s + 2];
    const double weight = dcf ? dcf[s] : 1.0;
    const double phase = wz * sz + wy * sy + wx * sx;
    const double pr = cos(phase), pi = sin(phase);
    const double yr = values[2 * s] * weight, yi = values[2 * s + 1] * weight;
    const double zr = yr * pr - yi * pi, zi = yr * pi + yi * pr;
    uint32_t cz, cy, cx;
    float f;
    if (ora_grid_coordinate((float)wz, (uint32_t)gz, &cz, &f) || ora_grid_coordinate((float)wy, (uint32_t)gy, &cy, &f) ||
        ora_grid_coordinate((float)wx, (uint32_t)gx, &cx, &f)) {
      free(grid);
      return -1;
    }
    const double dz = wrapped_coordinate(wz, gz) - (double)cz, dy = wrapped_coordinate(wy, gy) - (double)cy;
    const double dx = wrapped_coordinate(wx, gx) - (double)cx;
    for (int t = 0; t < width; ++t) {
      tz[t] = ora_kb_kernel((double)(t - lead) - dz, width, beta);
      ty[t] = ora_kb_kernel((double)(t - lead) - dy, width, beta);
      tx[t] = ora_kb_kernel((double)(t - lead) - dx, width, beta);
    }
    for (int a = 0; a < width; ++a) {
      const size_t mz = ((size_t)cz + gz + (size_t)a - (size_t)lead) % gz;
      for (int b = 0; b < width; ++b) {
        const size_t my = ((size_t)cy + gy + (size_t)b - (size_t)lead) % gy;
        const double wzy = tz[a] * ty[b];
        for (int c = 0; c < width; ++c) {
          const size_t mx = ((size_t)cx + gx + (size_t)c - (size_t)lead) % gx;
          const double w = wzy * tx[c];
          const size_t i = (mz * gy + my) * gx + mx;
          grid[2 * i] += zr * w;
          grid[2 * i + 1] += zi * w;
        }
      }
    }
  }
  inverse_dft3_unscaled(gz, gy, gx, grid);
  for (size_t z = 0; z < nz; ++z) {
    const long jz = (long)z - (long)(nz / 2);
    const size_t iz = (size_t)((jz + (long)gz) % (long)gz);
    const double dz = ora_kb_ft((double)jz, gz, width, beta);
    for (size_t y = 0; y < ny; ++y) {
      const long jy = (long)y - (long)(ny / 2);
      const size_t iy = (size_t)((jy + (long)gy) % (long)gy);
      const double dzy = dz * ora_kb_ft((double)jy, gy, width, beta);
      for (size_t x = 0; x < nx; ++x) {
        const long jx = (long)x - (long)(nx / 2);
        const size_t ix = (size_t)((jx + (long)gx) % (long)gx);
        const double d = dzy * ora_kb_ft((double)jx, gx, width, beta);
        const size_t o = (z * ny + y) * nx + x, i = (iz * gy + iy) * gx + ix;
        image[2 * o] = grid[2 * i] / d;
        image[2 * o + 1] = grid[2 * i + 1] / d;
      }
    }
  }
  free(grid);
  return 0;
}

/* One axis of the voxel evaluation: e^{+2 pi i m j / g} for m = 0 .. g-1, j = the crop index
 * minus n/2 (the grid frequency ora_kb_adjoint3 reads for it), indexed exactly by (m j) mod g. */
static double* axis_phasors(size_t g, size_t n, size_t count, const uint32_t* idx) {
  double* e = (double*)malloc(2 * count * g * sizeof(double));
  for (size_t i = 0; i < count; ++i) {
    const long j = (long)idx[i] - (long)(n / 2);
    const size_t jm = (size_t)((j % (long)g + (long)g) % (long)g);
    for (size_t m = 0; m < g; ++m) {
      const double a = ORA_TWO_PI * (double)((m * jm) % g) / (double)g;
      e[2 * (i * g + m)] = cos(a);
      e[2 * (i * g + m) + 1] = sin(a);
    }
  }
  return e;
}

/* sum_a t[a] e^{2 pi i (c + a - lead) j / g} for each requested index of one axis. */
static void axis_sums(const double* e, size_t g, size_t count, int width, int lead, uint32_t c, const double* t,
                      double* out) {
  for (size_t i = 0; i < count; ++i) {
    double sr = 0.0, si = 0.0;
    for (int a = 0; a < width; ++a) {
      const size_t m = ((size_t)c + g + (size_t)a - (size_t)lead) % g;
      sr += t[a] * e[2 * (i * g + m)];
      si += t[a] * e[2 * (i * g + m) + 1];
    }
    out[2 * i] = sr;
    out[2 * i + 1] = si;
  }
}

int ora_kb_adjoint3_voxels(size_t nz, size_t ny, size_t nx, size_t gz, size_t gy, size_t gx, int width, double beta,
                           double oz, double oy, double ox, size_t s_begin, size_t s_end, const double* traj,
                           const double* values, const double* dcf, size_t nzs, const uint32_t* zs, size_t nys,
                           const uint32_t* ys, size_t nxs, const uint32_t* xs, double* out) {
  const double sz = (double)(nz / 2) - oz, sy = (double)(ny / 2) - oy, sx = (double)(nx / 2) - ox;
  const int lead = width / 2 - 1;
  double* ez = axis_phasors(gz, nz, nzs, zs);
  double* ey = axis_phasors(gy, ny, nys, ys);
  double* ex = axis_phasors(gx, nx, nxs, xs);
  double* az = (double*)malloc(2 * nzs * sizeof(double));
  double* ay = (double*)malloc(2 * nys * sizeof(double));
  double* ax = (double*)malloc(2 * nxs * sizeof(double));
  double* acc = (double*)calloc(2 * nzs * nys * nxs, sizeof(double));
  double tz[16], ty[16], tx[16];
  int rc = 0;
  for (size_t s = s_begin; s < s_end; ++s) {
    const double wz = traj[3 * s], wy = traj[3 * s + 1], wx = traj[3 * s + 2];
    const double weight = dcf ? dcf[s] : 1.0;
    const double phase = wz * sz + wy * sy + wx * sx;
    const double pr = cos(phase), pi = sin(phase);
    const double yr = values[2 * s] * weight, yi = values[2 * s + 1] * weight;
    const double zr = yr * pr - yi * pi, zi = yr * pi + yi * pr;
    uint32_t cz, cy, cx;
    float f;
    if (ora_grid_coordinate((float)wz, (uint32_t)gz, &cz, &f) || ora_grid_coordinate((float)wy, (uint32_t)gy, &cy, &f) ||
        ora_grid_coordinate((float)wx, (uint32_t)gx, &cx, &f)) {
      rc = -1;
      break;
    }
    const double dz = wrapped_coordinate(wz, gz) - (double)cz, dy = wrapped_coordinate(wy, gy) - (double)cy;
    const double dx = wrapped_coordinate(wx, gx) - (double)cx;
    for (int t = 0; t < width; ++t) {
      tz[t] = ora_kb_kernel((double)(t - lead) - dz, width, beta);
      ty[t] = ora_kb_kernel((double)(t - lead) - dy, width, beta);
      tx[t] = ora_kb_kernel((double)(t - lead) - dx, width, beta);
    }
    axis_sums(ez, gz, nzs, width, lead, cz, tz, az);
    axis_sums(ey, gy, nys, width, lead, cy, ty, ay);
    axis_sums(ex, gx, nxs, width, lead, cx, tx, ax);
    for (size_t a = 0; a < nzs; ++a) {
      const double qr = zr * az[2 * a] - zi * az[2 * a + 1], qi = zr * az[2 * a + 1] + zi * az[2 * a];
      for (size_t b = 0; b < nys; ++b) {
        const double rr = qr * ay[2 * b] - qi * ay[2 * b + 1], ri = qr * ay[2 * b + 1] + qi * ay[2 * b];
        double* o = acc + 2 * (a * nys + b) * nxs;
        for (size_t c = 0; c < nxs; ++c) {
          o[2 * c] += rr * ax[2 * c] - ri * ax[2 * c + 1];
          o[2 * c + 1] += rr * ax[2 * c + 1] + ri * ax[2 * c];
        }
      }
    }
  }
  if (rc == 0) {
    for (size_t a = 0; a < nzs; ++a) {
      const double dza = ora_kb_ft((double)((long)zs[a] - (long)(nz / 2)), gz, width, beta);
      for (size_t b = 0; b < nys; ++b) {
        const double dzy = dza * ora_kb_ft((double)((long)ys[b] - (long)(ny / 2)), gy, width, beta);
        for (size_t c = 0; c < nxs; ++c) {
          const double d = dzy * ora_kb_ft((double)((long)xs[c] - (long)(nx / 2)), gx, width, beta);
          const size_t i = (a * nys + b) * nxs + c;
          out[2 * i] += acc[2 * i] / d;
          out[2 * i + 1] += acc[2 * i + 1] / d;
        }
      }
    }
  }
  free(ez);
  free(ey);
  free(ex);
  free(az);
  free(ay);
  free(ax);
  free(acc);
  return rc;
}

void ora_golden_means3(size_t spokes, size_t readout, double* traj) {
  /* 3-D radial "kooshball" with multidimensional golden means (Chan et al., MRM 61
   * 2009): spoke p has polar cosine frac(p g1) and azimuth 2 pi frac(p g2), with
   * g1 = 0.4656, g2 = 0.6823 the golden means of the 3x3 Fibonacci recurrence; each
   * spoke runs through the centre, k_j = (j - M/2)/M cycles per pixel, written
   * (kz, ky, kx) in radians per pixel. */
  const double g1 = 0.465571231876768, g2 = 0.682327803828019;
  for (size_t p = 0; p < spokes; ++p) {
    const double cz = fmod((double)p * g1, 1.0);
    const double az = ORA_TWO_PI * fmod((double)p * g2, 1.0);
    const double sz = sqrt(1.0 - cz * cz);
    const double dz = cz, dy = sz * sin(az), dx = sz * cos(az);
    for (size_t j = 0; j < readout; ++j) {
      const double k = ORA_TWO_PI * ((double)j - (double)(readout / 2)) / (double)readout;
      traj[3 * (p * readout + j)] = k * dz;
      traj[3 * (p * readout + j) + 1] = k * dy;
      traj[3 * (p * readout + j) + 2] = k * dx;
    }
  }
}

/* ---------------------------------------------------------------- trajectory units */

static int normalize_coordinate(float c, int units, uint32_t encoded, float* out) {
  /* noncartesian_rss_hdf5.cpp:776-808, float arithmetic as there. */
  const float pi_f = 3.14159265358979323846f;
  if (!isfinite(c)) return 1;
  switch (units) {
    case 1: /* cycles_per_fov */
      if (fabsf(c) > 0.5f) return 2;
      *out = c * (2.0f * pi_f);
      return 0;
    case 2: /* radians_per_pixel */
      if (fabsf(c) > pi_f) return 2;
      *out = c;
      return 0;
    case 3: { /* encoded_matrix_index */
      if (encoded < 2) return 4;
      const float half = (float)encoded / 2.0f;
      if (fabsf(c) > half) return 2;
      *out = c * (2.0f * pi_f / (float)encoded);
      return 0;
    }
    default:
      return 3;
  }
}

int ora_normalize_trajectory2(size_t samples, const float* raw, int units, uint32_t encoded_rows,
                              uint32_t encoded_cols, float* out) {
  /* noncartesian_rss_hdf5.cpp:815-834: raw [kx, ky] per sample, each axis normalised with
   * its own encoded dimension (kx first), written [row = ky, column = kx].  Returns 0 or
   * the code of the first failure: 1 non-finite, 2 outside the Nyquist interval,
   * 3 units unspecified, 4 encoded dimension < 2. */
  for (size_t s = 0; s < samples; ++s) {
    float col, row;
    int rc = normalize_coordinate(raw[2 * s], units, encoded_cols, &col);
    if (rc) return rc;
    rc = normalize_coordinate(raw[2 * s + 1], units, encoded_rows, &row);
    if (rc) return rc;
    out[2 * s] = row;
    out[2 * s + 1] = col;
  }
  return 0;
}

/* ------------------------------------------------------------------ k-space conditioning */

void ora_channel_mix(size_t out_channels, size_t in_channels, size_t samples, const double* w, const double* x,
                     double* y) {
  /* noncartesian_noise_prewhiten.cpp:126-133 (noise_prewhiten.cpp:141, coil_compress.cpp:159,
   * noncartesian_coil_compress.cpp:148):
   * ksj::linalg::matmul(W, X, Y) over row-major [out][in] and channel-major [in][samples]. */
  for (size_t o = 0; o < out_channels; ++o) {
    for (size_t s = 0; s < samples; ++s) {
      double re = 0.0, im = 0.0;
      for (size_t i = 0; i < in_channels; ++i) {
        const double wr = w[2 * (o * in_channels + i)], wi = w[2 * (o * in_channels + i) + 1];
        const double xr = x[2 * (i * samples + s)], xi = x[2 * (i * samples + s) + 1];
        re += wr * xr - wi * xi;
        im += wr * xi + wi * xr;
      }
      y[2 * (o * samples + s)] = re;
      y[2 * (o * samples + s) + 1] = im;
    }
  }
}

void ora_phase_correct(size_t channels, size_t rows, size_t cols, const float* x, const float* p, float* y) {
  /* phase_correct.cpp:140-142: value * phase(channel, sample % cols); built with
   * -ffp-contract=off so each product and sum rounds to float as the reference's does. */
  for (size_t c = 0; c < channels; ++c) {
    for (size_t r = 0; r < rows; ++r) {
      for (size_t k = 0; k < cols; ++k) {
        const size_t e = (c * rows + r) * cols + k;
        const float a = x[2 * e], b = x[2 * e + 1];
        const float pc = p[2 * (c * cols + k)], pd = p[2 * (c * cols + k) + 1];
        const float ac = a * pc, bd = b * pd, ad = a * pd, bc = b * pc;
        y[2 * e] = ac - bd;
        y[2 * e + 1] = ad + bc;
      }
    }
  }
}

void ora_readout_crop(size_t lines, size_t in_cols, size_t out_cols, size_t offset, const float* x, float* y) {
  /* readout_oversampling_remove.cpp:174-180: copy_n of out_cols values from offset per line. */
  for (size_t l = 0; l < lines; ++l)
    for (size_t k = 0; k < out_cols; ++k) {
      y[2 * (l * out_cols + k)] = x[2 * (l * in_cols + offset + k)];
      y[2 * (l * out_cols + k) + 1] = x[2 * (l * in_cols + offset + k) + 1];
    }
}

/* ---------------------------------------------------------------- calibration */

void ora_channel_covariance(size_t channels, size_t samples, const float* x, int centred, double* cov) {
  double* mr = (double*)calloc(channels, sizeof(double));
  double* mi = (double*)calloc(channels, sizeof(double));
  if (centred) {
    for (size_t c = 0; c < channels; ++c) {
      for (size_t s = 0; s < samples; ++s) {
        mr[c] += x[2 * (c * samples + s)];
        mi[c] += x[2 * (c * samples + s) + 1];
      }
      mr[c] /= (double)samples;
      mi[c] /= (double)samples;
    }
  }
  for (size_t r = 0; r < channels; ++r)
    for (size_t c = 0; c < channels; ++c) {
      double sr = 0.0, si = 0.0;
      for (size_t s = 0; s < samples; ++s) {
        const double ar = x[2 * (r * samples + s)] - mr[r], ai = x[2 * (r * samples + s) + 1] - mi[r];
        const double br = x[2 * (c * samples + s)] - mr[c], bi = x[2 * (c * samples + s) + 1] - mi[c];
        if (centred) { /* conj(a) b */
          sr += ar * br + ai * bi;
          si += ar * bi - ai * br;
        } else { /* a conj(b) */
          sr += ar * br + ai * bi;
          si += ai * br - ar * bi;
        }
      }
      cov[2 * (r * channels + c)] = sr / (double)samples;
      cov[2 * (r * channels + c) + 1] = si / (double)samples;
    }
  free(mr);
  free(mi);
}

int ora_hermitian_eigen(size_t n, const double* matrix, double* values, double* vectors) {
  /* decompositions.hpp:533-626: copy, identity vectors, tolerance from the diagonal scale,
   * largest off-diagonal pivot, rotation (tau, tangent, cosine, sine, pivot phase), sort. */
  double* a = (double*)malloc(2 * n * n * sizeof(double));
  double* v = vectors;
  double scale = 0.0;
  for (size_t r = 0; r < n; ++r) {
    for (size_t c = 0; c < n; ++c) {
      a[2 * (r * n + c)] = matrix[2 * (r * n + c)];
      a[2 * (r * n + c) + 1] = matrix[2 * (r * n + c) + 1];
      v[2 * (r * n + c)] = r == c ? 1.0 : 0.0;
      v[2 * (r * n + c) + 1] = 0.0;
    }
    if (!isfinite(a[2 * (r * n + r)])) {
      free(a);
      return -1;
    }
    scale += fabs(a[2 * (r * n + r)]);
  }
  const double tol = 2.220446049250313e-16 * (scale > 1.0 ? scale : 1.0) * (double)n * 8.0;
  const size_t max_rot = n * n * 64;
  for (size_t rot = 0; rot < max_rot; ++rot) {
    size_t p = 0, q = 0;
    double mag = 0.0;
    for (size_t r = 0; r < n; ++r)
      for (size_t c = r + 1; c < n; ++c) {
        const double m = hypot(a[2 * (r * n + c)], a[2 * (r * n + c) + 1]);
        if (!isfinite(m)) {
          free(a);
          return -1;
        }
        if (m > mag) {
          mag = m;
          p = r;
          q = c;
        }
      }
    if (mag <= tol) {
      for (size_t i = 0; i < n; ++i) values[i] = a[2 * (i * n + i)];
      for (size_t first = 0; first < n; ++first) {
        size_t mn = first;
        for (size_t cand = first + 1; cand < n; ++cand)
          if (values[cand] < values[mn]) mn = cand;
        if (mn == first) continue;
        const double t = values[first];
        values[first] = values[mn];
        values[mn] = t;
        for (size_t r = 0; r < n; ++r)
          for (int h = 0; h < 2; ++h) {
            const double u = v[2 * (r * n + first) + h];
            v[2 * (r * n + first) + h] = v[2 * (r * n + mn) + h];
            v[2 * (r * n + mn) + h] = u;
          }
      }
      free(a);
      return 0;
    }
    const double app = a[2 * (p * n + p)], aqq = a[2 * (q * n + q)];
    const double tau = (aqq - app) / (2.0 * mag);
    const double tg = tau >= 0.0 ? 1.0 / (tau + sqrt(1.0 + tau * tau)) : -1.0 / (-tau + sqrt(1.0 + tau * tau));
    const double cs = 1.0 / sqrt(1.0 + tg * tg), sn = tg * cs;
    const double phr = a[2 * (p * n + q)] / mag, phi = a[2 * (p * n + q) + 1] / mag; /* phase; conj = (phr, -phi) */
    for (size_t i = 0; i < n; ++i) {
      if (i == p || i == q) continue;
      const double rr = a[2 * (i * n + p)], ri = a[2 * (i * n + p) + 1];
      const double qr0 = a[2 * (i * n + q)], qi0 = a[2 * (i * n + q) + 1];
      const double cr = qr0 * phr + qi0 * phi, ci = qi0 * phr - qr0 * phi; /* times conj(phase) */
      const double ur = cs * rr - sn * cr, ui = cs * ri - sn * ci;
      const double wr = sn * rr + cs * cr, wi = sn * ri + cs * ci;
      a[2 * (i * n + p)] = ur;
      a[2 * (i * n + p) + 1] = ui;
      a[2 * (p * n + i)] = ur;
      a[2 * (p * n + i) + 1] = -ui;
      a[2 * (i * n + q)] = wr;
      a[2 * (i * n + q) + 1] = wi;
      a[2 * (q * n + i)] = wr;
      a[2 * (q * n + i) + 1] = -wi;
    }
    a[2 * (p * n + p)] = cs * cs * app - 2.0 * cs * sn * mag + sn * sn * aqq;
    a[2 * (p * n + p) + 1] = 0.0;
    a[2 * (q * n + q)] = sn * sn * app + 2.0 * cs * sn * mag + cs * cs * aqq;
    a[2 * (q * n + q) + 1] = 0.0;
    a[2 * (p * n + q)] = a[2 * (p * n + q) + 1] = 0.0;
    a[2 * (q * n + p)] = a[2 * (q * n + p) + 1] = 0.0;
    for (size_t r = 0; r < n; ++r) {
      const double xr = v[2 * (r * n + p)], xi = v[2 * (r * n + p) + 1];
      const double yr0 = v[2 * (r * n + q)], yi0 = v[2 * (r * n + q) + 1];
      const double yr = yr0 * phr + yi0 * phi, yi = yi0 * phr - yr0 * phi;
      v[2 * (r * n + p)] = cs * xr - sn * yr;
      v[2 * (r * n + p) + 1] = cs * xi - sn * yi;
      v[2 * (r * n + q)] = sn * xr + cs * yr;
      v[2 * (r * n + q) + 1] = sn * xi + cs * yi;
    }
  }
  free(a);
  return -1;
}

int ora_noise_model_estimate(size_t channels, size_t samples, const float* x, double floor_value, double* whitening) {
  double* cov = (double*)malloc(2 * channels * channels * sizeof(double));
  double* vec = (double*)malloc(2 * channels * channels * sizeof(double));
  double* val = (double*)malloc(channels * sizeof(double));
  ora_channel_covariance(channels, samples, x, 1, cov);
  const int rc = ora_hermitian_eigen(channels, cov, val, vec);
  if (rc == 0) {
    for (size_t r = 0; r < channels; ++r)
      for (size_t c = 0; c < channels; ++c) {
        double sr = 0.0, si = 0.0;
        for (size_t k = 0; k < channels; ++k) {
          const double inv = 1.0 / sqrt(val[k] > floor_value ? val[k] : floor_value);
          const double ar = vec[2 * (r * channels + k)], ai = vec[2 * (r * channels + k) + 1];
          const double br = vec[2 * (c * channels + k)], bi = vec[2 * (c * channels + k) + 1];
          sr += (ar * br + ai * bi) * inv;
          si += (ai * br - ar * bi) * inv;
        }
        whitening[2 * (r * channels + c)] = sr;
        whitening[2 * (r * channels + c) + 1] = si;
      }
  }
  free(cov);
  free(vec);
  free(val);
  return rc;
}

int ora_coil_compression_basis_estimate(size_t physical, size_t virtual_channels, size_t samples, const float* x,
                                        double* basis) {
  double* cov = (double*)malloc(2 * physical * physical * sizeof(double));
  double* vec = (double*)malloc(2 * physical * physical * sizeof(double));
  double* val = (double*)malloc(physical * sizeof(double));
  ora_channel_covariance(physical, samples, x, 0, cov);
  const int rc = ora_hermitian_eigen(physical, cov, val, vec);
  if (rc == 0) {
    for (size_t vc = 0; vc < virtual_channels; ++vc)
      for (size_t p = 0; p < physical; ++p) {
        const size_t col = physical - 1 - vc;
        basis[2 * (vc * physical + p)] = vec[2 * (p * physical + col)];
        basis[2 * (vc * physical + p) + 1] = -vec[2 * (p * physical + col) + 1];
      }
  }
  free(cov);
  free(vec);
  free(val);
  return rc;
}

void ora_phase_correction_estimate(size_t channels, size_t lines, size_t readout, const float* ref, float* phase) {
  for (size_t c = 0; c < channels; ++c)
    for (size_t k = 0; k < readout; ++k) {
      float sr = 0.f, si = 0.f;
      for (size_t l = 0; l < lines; ++l) {
        sr += ref[2 * ((c * lines + l) * readout + k)];
        si += ref[2 * ((c * lines + l) * readout + k) + 1];
      }
      const float mag = (float)sqrt((double)sr * (double)sr + (double)si * (double)si);
      float* out = phase + 2 * (c * readout + k);
      if (mag == 0.f) {
        out[0] = 1.f;
        out[1] = 0.f;
      } else {
        out[0] = sr / mag;
        out[1] = -si / mag;
      }
    }
}

/* ---------------------------------------------------------------- sensitivity maps */

/* Walsh, Gmitro & Marcellin, "Adaptive reconstruction of phased array MR imagery",
 * MRM 43:682-690 (2000): at each pixel the coil sensitivity vector is the dominant
 * eigenvector of the local coil covariance R = sum over the window of I I^H.  Restated
 * with the local covariance formed explicitly and the power iteration written out. */
static void walsh_normalise(size_t coils, double* v) {
  double n2 = 0.0;
  for (size_t c = 0; c < coils; ++c) n2 += v[2 * c] * v[2 * c] + v[2 * c + 1] * v[2 * c + 1];
  const double inv = n2 > 0.0 ? 1.0 / sqrt(n2) : 0.0;
  for (size_t c = 0; c < 2 * coils; ++c) v[c] *= inv;
}

void ora_walsh_maps(size_t rows, size_t cols, size_t coils, size_t window, size_t iterations, const double* images,
                    double* maps) {
  const long r = (long)(window / 2);
  double* R = (double*)calloc(2 * coils * coils, sizeof(double));
  double* v = (double*)calloc(2 * coils, sizeof(double));
  double* u = (double*)calloc(2 * coils, sizeof(double));
  for (size_t y = 0; y < rows; ++y)
    for (size_t x = 0; x < cols; ++x) {
      memset(R, 0, 2 * coils * coils * sizeof(double));
      for (long dy = -r; dy <= r; ++dy)
        for (long dx = -r; dx <= r; ++dx) {
          const long yy = (long)y + dy, xx = (long)x + dx;
          if (yy < 0 || yy >= (long)rows || xx < 0 || xx >= (long)cols) continue;
          const double* a = images + 2 * (((size_t)yy * cols + (size_t)xx) * coils);
          for (size_t i = 0; i < coils; ++i)
            for (size_t j = 0; j < coils; ++j) { /* R[i][j] += a_i conj(a_j) */
              R[2 * (i * coils + j)] += a[2 * i] * a[2 * j] + a[2 * i + 1] * a[2 * j + 1];
              R[2 * (i * coils + j) + 1] += a[2 * i + 1] * a[2 * j] - a[2 * i] * a[2 * j + 1];
            }
        }
      for (size_t i = 0; i < coils; ++i) { /* v = R e_0 */
        v[2 * i] = R[2 * (i * coils)];
        v[2 * i + 1] = R[2 * (i * coils) + 1];
      }
      walsh_normalise(coils, v);
      for (size_t it = 0; it < iterations; ++it) {
        for (size_t i = 0; i < coils; ++i) {
          double ur = 0.0, ui = 0.0;
          for (size_t j = 0; j < coils; ++j) {
            const double rr = R[2 * (i * coils + j)], ri = R[2 * (i * coils + j) + 1];
            ur += rr * v[2 * j] - ri * v[2 * j + 1];
            ui += rr * v[2 * j + 1] + ri * v[2 * j];
          }
          u[2 * i] = ur;
          u[2 * i + 1] = ui;
        }
        memcpy(v, u, 2 * coils * sizeof(double));
        walsh_normalise(coils, v);
      }
      const double m = sqrt(v[0] * v[0] + v[1] * v[1]); /* coil 0 real and >= 0 */
      double* out = maps + 2 * ((y * cols + x) * coils);
      for (size_t c = 0; c < coils; ++c) {
        if (m > 0.0) {
          out[2 * c] = (v[2 * c] * v[0] + v[2 * c + 1] * v[1]) / m;
          out[2 * c + 1] = (v[2 * c + 1] * v[0] - v[2 * c] * v[1]) / m;
        } else {
          out[2 * c] = v[2 * c];
          out[2 * c + 1] = v[2 * c + 1];
        }
      }
    }
  free(R);
  free(v);
  free(u);
}
