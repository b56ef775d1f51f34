// FP64 kernel-function helpers shared by the 2-D and 3-D plans: the modified Bessel
// function I0, the Kaiser-Bessel (or linear) kernel value, and the grid-coordinate
// rule of the binning.
#pragma once

#include "plan.hpp"

namespace ksj::gpu {

constexpr uint32_t kErrBadCoordinate = 1;

// Modified Bessel function I0 by its power series (FP64), the same definition as
// the oracle's; beta <= ~40 converges in < 60 terms.
__device__ inline double bessel_i0(double x) {
  const double q = 0.25 * x * x;
  double term = 1.0, sum = 1.0;
  for (int k = 1; k < 500; ++k) {
    term *= q / (static_cast<double>(k) * static_cast<double>(k));
    sum += term;
    if (term < sum * 1e-17) break;
  }
  return sum;
}

__device__ inline double kernel_value(double d, uint32_t kernel, int width, double beta, double i0_beta) {
  if (kernel == kKernelLinear) {
    const double a = fabs(d);
    return a < 1.0 ? 1.0 - a : 0.0;
  }
  const double half = 0.5 * width;
  if (!(fabs(d) < half)) return 0.0;
  const double t = d / half;
  return bessel_i0(beta * sqrt(1.0 - t * t)) / i0_beta;
}

// FP64 grid coordinate; the IEEE sequence must match the CPU binning exactly.
__device__ inline bool grid_coordinate(float w, uint32_t extent, uint32_t* cell, double* frac) {
#pragma clang fp contract(off)
  const double e = static_cast<double>(extent);
  double u = (static_cast<double>(w) * e) / kTwoPi;
  if (u < 0.0) u += e;
  if (u >= e) u -= e;
  if (!(u >= 0.0 && u < e)) return false;
  const double f = floor(u);
  *cell = static_cast<uint32_t>(f);
  *frac = u - f;
  return true;
}

// I0 on the host (same series).
inline double host_i0(double x) {
  const double q = 0.25 * x * x;
  double term = 1.0, sum = 1.0;
  for (int k = 1; k < 500; ++k) {
    term *= q / (static_cast<double>(k) * static_cast<double>(k));
    sum += term;
    if (term < sum * 1e-17) break;
  }
  return sum;
}

}  // namespace ksj::gpu
