// C binding over the reference's own Eigen-free NUFFT sources.
//
// TEST INFRASTRUCTURE ONLY.  oracle/Makefile compiles this file together with
//   libs/numerics/kspacejet-nufft/src/eigen/eigen_radial_gridding.cpp
//   libs/numerics/kspacejet-nufft/src/eigen/eigen_nufft_direct_nudft.cpp
// where they lie in the reference tree, into oracle/_ref/libksj_ref.so.  It is
// used to pin the FP64 oracle against the reference's compiled code and as the
// reference-route CPU baseline.  No reference source is copied here; this file
// only adapts raw pointers to the reference's view types.

#include "kspacejet/nufft/detail/eigen/eigen_nufft_direct_nudft.hpp"
#include "kspacejet/nufft/detail/eigen/eigen_radial_gridding.hpp"

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstddef>
#include <cstdint>
#include <thread>
#include <vector>

namespace {

namespace ref = ksj::nufft::detail::eigen;
using ksj::array::MatrixView;
using ksj::array::VectorView;

template <typename T>
int gridding(std::size_t rows, std::size_t cols, double r0, double c0, std::size_t samples,
             const std::complex<T>* values, const T* traj, const T* dcf, std::complex<T>* image) {
  try {
    const std::size_t line = std::max(rows, cols);
    std::vector<std::complex<T>> inter(rows * cols), src(line), dst(line);
    ref::radial_linear_gridding2_adjoint(
      ksj::nufft::Grid2D{rows, cols, r0, c0}, VectorView<const std::complex<T>>(values, samples),
      MatrixView<const T>(traj, samples, 2U), VectorView<const T>(dcf, samples), MatrixView<std::complex<T>>(image, rows, cols),
      ksj::nufft::RadialGridding2Workspace<T>{MatrixView<std::complex<T>>(inter.data(), rows, cols),
                                              VectorView<std::complex<T>>(src.data(), line),
                                              VectorView<std::complex<T>>(dst.data(), line)});
    return 0;
  } catch (...) {
    return -1;
  }
}

}  // namespace

extern "C" {

int ref_ramp_dcf2_f32(std::size_t samples, const float* traj, float* dcf) {
  try {
    ref::radial_analytic_ramp_dcf2(MatrixView<const float>(traj, samples, 2U), VectorView<float>(dcf, samples));
    return 0;
  } catch (...) {
    return -1;
  }
}

int ref_ramp_dcf2_f64(std::size_t samples, const double* traj, double* dcf) {
  try {
    ref::radial_analytic_ramp_dcf2(MatrixView<const double>(traj, samples, 2U), VectorView<double>(dcf, samples));
    return 0;
  } catch (...) {
    return -1;
  }
}

int ref_radial_linear_gridding2_f32(std::size_t rows, std::size_t cols, double r0, double c0, std::size_t samples,
                                    const float* values, const float* traj, const float* dcf, float* image) {
  return gridding<float>(rows, cols, r0, c0, samples, reinterpret_cast<const std::complex<float>*>(values), traj, dcf,
                         reinterpret_cast<std::complex<float>*>(image));
}

int ref_radial_linear_gridding2_f64(std::size_t rows, std::size_t cols, double r0, double c0, std::size_t samples,
                                    const double* values, const double* traj, const double* dcf, double* image) {
  return gridding<double>(rows, cols, r0, c0, samples, reinterpret_cast<const std::complex<double>*>(values), traj,
                          dcf, reinterpret_cast<std::complex<double>*>(image));
}

int ref_direct_adjoint2_f64(std::size_t rows, std::size_t cols, double r0, double c0, std::size_t samples,
                            const double* traj, const double* values, double* image) {
  try {
    ref::direct_nudft2_adjoint(ksj::nufft::Grid2D{rows, cols, r0, c0},
                               VectorView<const std::complex<double>>(
                                 reinterpret_cast<const std::complex<double>*>(values), samples),
                               MatrixView<const double>(traj, samples, 2U),
                               MatrixView<std::complex<double>>(reinterpret_cast<std::complex<double>*>(image), rows,
                                                                cols));
    return 0;
  } catch (...) {
    return -1;
  }
}

int ref_direct_forward2_f64(std::size_t rows, std::size_t cols, double r0, double c0, std::size_t samples,
                            const double* traj, const double* image, double* values) {
  try {
    ref::direct_nudft2_forward(ksj::nufft::Grid2D{rows, cols, r0, c0},
                               MatrixView<const std::complex<double>>(
                                 reinterpret_cast<const std::complex<double>*>(image), rows, cols),
                               MatrixView<const double>(traj, samples, 2U),
                               VectorView<std::complex<double>>(reinterpret_cast<std::complex<double>*>(values),
                                                                samples));
    return 0;
  } catch (...) {
    return -1;
  }
}

// The reference radial route of one scan (radial_gridding_reconstruct.cpp:391-413 and
// coil_combine_rss.cpp:308): ramp DCF once, per channel bilinear gridding + unscaled
// inverse transform, then root-sum-of-squares over channels.  Scans are spread over
// `threads` host threads, one scan per thread at a time (scan-level parallelism).
int ref_radial_route_rss_batch(std::size_t scans, std::size_t rows, std::size_t cols, std::size_t channels,
                               std::size_t samples, const float* values, const float* traj, float* rss,
                               unsigned threads) {
  const double r0 = static_cast<double>(rows - 1U) / 2.0;
  const double c0 = static_cast<double>(cols - 1U) / 2.0;
  std::vector<int> status(scans, 0);
  auto work = [&](std::size_t first, std::size_t step) {
    std::vector<float> dcf(samples);
    std::vector<std::complex<float>> image(rows * cols);
    std::vector<float> acc(rows * cols);
    for (std::size_t scan = first; scan < scans; scan += step) {
      const float* scan_traj = traj + scan * samples * 2U;
      if (ref_ramp_dcf2_f32(samples, scan_traj, dcf.data()) != 0) {
        status[scan] = -1;
        continue;
      }
      std::fill(acc.begin(), acc.end(), 0.0F);
      for (std::size_t ch = 0; ch < channels; ++ch) {
        const float* v = values + (scan * channels + ch) * samples * 2U;
        if (gridding<float>(rows, cols, r0, c0, samples, reinterpret_cast<const std::complex<float>*>(v), scan_traj,
                            dcf.data(), image.data()) != 0) {
          status[scan] = -1;
          break;
        }
        for (std::size_t p = 0; p < rows * cols; ++p) acc[p] += std::norm(image[p]);
      }
      for (std::size_t p = 0; p < rows * cols; ++p) rss[scan * rows * cols + p] = std::sqrt(acc[p]);
    }
  };
  threads = std::max(1U, threads);
  std::vector<std::thread> pool;
  for (unsigned t = 1; t < threads; ++t) pool.emplace_back(work, t, threads);
  work(0U, threads);
  for (auto& th : pool) th.join();
  for (int s : status)
    if (s != 0) return -1;
  return 0;
}

// Timing form of the route above: `scans` repetitions of one scan's data, each
// thread writing its own RSS image (the last one is copied to `rss`).
int ref_radial_route_rss_repeat(std::size_t scans, std::size_t rows, std::size_t cols, std::size_t channels,
                                std::size_t samples, const float* values, const float* traj, float* rss,
                                unsigned threads) {
  threads = std::max(1U, threads);
  std::vector<std::vector<float>> out(threads, std::vector<float>(rows * cols));
  std::vector<int> status(threads, 0);
  auto work = [&](unsigned t) {
    for (std::size_t scan = t; scan < scans; scan += threads)
      if (ref_radial_route_rss_batch(1U, rows, cols, channels, samples, values, traj, out[t].data(), 1U) != 0)
        status[t] = -1;
  };
  std::vector<std::thread> pool;
  for (unsigned t = 1; t < threads; ++t) pool.emplace_back(work, t);
  work(0U);
  for (auto& th : pool) th.join();
  std::copy(out[0].begin(), out[0].end(), rss);
  for (int s : status)
    if (s != 0) return -1;
  return 0;
}

}  // extern "C"
