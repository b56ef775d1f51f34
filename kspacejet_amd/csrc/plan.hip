// Trajectory plan kernels: FP64 coordinates, binning, taps, heavy cells, deapodization.
#include "plan.hpp"
#include "sort.hpp"

#include <cmath>
#include <string>

namespace ksj::gpu {
namespace {

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

struct CoordArgs {
  const float* traj;
  const float* dcf;
  uint32_t samples, grid_rows, grid_cols, density;
  double shift_r, shift_c;
  uint32_t* bins;
  double* frac;
  float2* weight;
  uint32_t* err;
};

__global__ void coords_kernel(CoordArgs a) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= a.samples) return;
  const float wr = a.traj[2 * s], wc = a.traj[2 * s + 1];
  constexpr float kPiF = 3.14159265358979323846f;
  uint32_t cr = 0, cc = 0;
  double fr = 0.0, fc = 0.0;
  const bool finite = isfinite(wr) && isfinite(wc);
  const bool in_range = finite && wr >= -kPiF && wr <= kPiF && wc >= -kPiF && wc <= kPiF;
  if (!in_range || !grid_coordinate(wr, a.grid_rows, &cr, &fr) || !grid_coordinate(wc, a.grid_cols, &cc, &fc)) {
    atomicOr(a.err, kErrBadCoordinate);
    cr = cc = 0;
    fr = fc = 0.0;
  }
  a.bins[s] = cr * a.grid_cols + cc;
  a.frac[2 * s] = fr;
  a.frac[2 * s + 1] = fc;
  double dcf = 1.0;
  if (a.density == kDensityRamp) dcf = hypot(static_cast<double>(wr), static_cast<double>(wc));
  else if (a.density == kDensityGiven) dcf = static_cast<double>(a.dcf[s]);
  double sn, cs;
  sincos(static_cast<double>(wr) * a.shift_r + static_cast<double>(wc) * a.shift_c, &sn, &cs);
  a.weight[s] = make_float2(static_cast<float>(dcf * cs), static_cast<float>(dcf * sn));
}

__global__ void cell_start_kernel(const uint32_t* sorted_keys, uint32_t samples, uint64_t nbins, uint32_t* start) {
  const uint64_t k = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (k > samples) return;
  const int64_t prev = k == 0 ? -1 : static_cast<int64_t>(sorted_keys[k - 1]);
  const int64_t cur = k == samples ? static_cast<int64_t>(nbins) : static_cast<int64_t>(sorted_keys[k]);
  for (int64_t b = prev + 1; b <= cur; ++b) start[b] = static_cast<uint32_t>(k);
}

struct TapArgs {
  const uint32_t* src;
  const double* frac;
  uint32_t samples, kernel;
  int width;
  double beta, i0_beta;
  uint32_t* rank;
  float* taps;
};

__global__ void taps_kernel(TapArgs a) {
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= a.samples) return;
  const uint32_t s = a.src[k];
  a.rank[s] = k;
  const double fr = a.frac[2 * s], fc = a.frac[2 * s + 1];
  const int lead = a.width / 2 - 1;
  float* out = a.taps + static_cast<uint64_t>(k) * 2 * a.width;
  for (int t = 0; t < a.width; ++t) {
    out[t] = static_cast<float>(kernel_value(static_cast<double>(t - lead) - fr, a.kernel, a.width, a.beta, a.i0_beta));
    out[a.width + t] =
      static_cast<float>(kernel_value(static_cast<double>(t - lead) - fc, a.kernel, a.width, a.beta, a.i0_beta));
  }
}

// Number of samples a cell gathers (sum over its W x W contributing bins).
__global__ void heavy_kernel(const uint32_t* start, uint32_t grid_rows, uint32_t grid_cols, int width,
                             uint32_t threshold, uint8_t* flag, uint32_t* list, uint32_t capacity,
                             uint32_t* counter) {
  const uint64_t m = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (m >= static_cast<uint64_t>(grid_rows) * grid_cols) return;
  const int y = static_cast<int>(m / grid_cols), x = static_cast<int>(m % grid_cols);
  const int lead = width / 2 - 1;
  uint32_t total = 0;
  for (int ty = 0; ty < width; ++ty) {
    const int by = (y - ty + lead + static_cast<int>(grid_rows)) % static_cast<int>(grid_rows);
    for (int tx = 0; tx < width; ++tx) {
      const int bx = (x - tx + lead + static_cast<int>(grid_cols)) % static_cast<int>(grid_cols);
      const uint64_t b = static_cast<uint64_t>(by) * grid_cols + bx;
      total += start[b + 1] - start[b];
    }
  }
  const bool heavy = total > threshold;
  flag[m] = heavy ? 1 : 0;
  if (heavy) {
    const uint32_t i = atomicAdd(counter, 1u);
    if (i < capacity) list[i] = static_cast<uint32_t>(m);
  }
}

__global__ void deapod_kernel(float* table, uint32_t n, uint32_t grid, int width, double beta, double i0_beta) {
  const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  // 1 / phi_hat(j) with phi_hat the continuous FT of the normalised KB kernel
  // at angular frequency 2 pi j / G: (W / I0(beta)) sinh(z)/z, z^2 = beta^2 - (pi W j / G)^2.
  const double j = static_cast<double>(static_cast<int64_t>(r) - static_cast<int64_t>(n / 2));
  const double a = kPi * width * j / static_cast<double>(grid);
  const double z2 = beta * beta - a * a;
  double shape = 1.0;
  if (z2 > 0.0) {
    const double z = sqrt(z2);
    shape = sinh(z) / z;
  } else if (z2 < 0.0) {
    const double z = sqrt(-z2);
    shape = sin(z) / z;
  }
  table[r] = static_cast<float>(i0_beta / (width * shape));
}

__global__ void golden_angle_kernel(float* traj, uint32_t spokes, uint32_t readout) {
  const uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= static_cast<uint64_t>(spokes) * readout) return;
  const uint32_t p = static_cast<uint32_t>(i / readout), j = static_cast<uint32_t>(i % readout);
  const double golden = kPi * (sqrt(5.0) - 1.0) / 2.0;
  double sn, cs;
  sincos(static_cast<double>(p) * golden, &sn, &cs);
  const double k = kTwoPi * (static_cast<double>(j) - static_cast<double>(readout / 2)) / static_cast<double>(readout);
  traj[2 * i] = static_cast<float>(k * sn);
  traj[2 * i + 1] = static_cast<float>(k * cs);
}

__global__ void ramp_dcf_kernel(const float* traj, float* dcf, uint64_t samples) {
  const uint64_t s = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (s >= samples) return;
  dcf[s] = hypotf(traj[2 * s], traj[2 * s + 1]);
}

double host_i0(double x) {
  const double q = 0.25 * x * x;
  double term = 1.0, sum = 1.0;
  for (int k = 1; k < 500; ++k) {
    term *= q / (static_cast<double>(k) * static_cast<double>(k));
    sum += term;
    if (term < sum * 1e-17) break;
  }
  return sum;
}

template <typename T> T* alloc(size_t count, size_t* total) {
  T* p = nullptr;
  const size_t bytes = count * sizeof(T);
  KSJ_HIP(hipMalloc(&p, bytes == 0 ? 1 : bytes));
  *total += bytes;
  return p;
}

}  // namespace

Plan2* plan2_create(const Plan2Desc& desc) {
  require(desc.image_rows >= 1 && desc.image_cols >= 1, "plan image dimensions must be positive");
  require(desc.grid_rows >= desc.image_rows && desc.grid_cols >= desc.image_cols, "grid must not be smaller than the image");
  require(desc.grid_rows <= 8192 && desc.grid_cols <= 8192, "grid side must be <= 8192");
  require(desc.width == 2 || desc.width == 4 || desc.width == 6 || desc.width == 8, "kernel width must be 2, 4, 6 or 8");
  require(desc.kernel == kKernelKaiserBessel || (desc.kernel == kKernelLinear && desc.width == 2),
          "linear gridding uses width 2");
  require(desc.centred == 1 || (desc.grid_rows == desc.image_rows && desc.grid_cols == desc.image_cols),
          "uncentred output requires an unoversampled grid");
  require(desc.samples >= 1, "plan needs at least one sample");
  require(desc.heavy_threshold >= 1, "heavy threshold must be positive");
  auto* p = new Plan2();
  p->d = desc;
  if (p->d.kernel == kKernelKaiserBessel && p->d.beta <= 0.0) {
    const double sigma = static_cast<double>(desc.grid_rows) / desc.image_rows;
    const double a = static_cast<double>(desc.width) / sigma * (sigma - 0.5);
    p->d.beta = kPi * std::sqrt(a * a - 0.8);
  }
  try {
    KSJ_HIP(hipGetDevice(&p->device));
    const uint64_t S = desc.samples;
    p->nbins = static_cast<uint64_t>(desc.grid_rows) * desc.grid_cols;
    p->key_bits = 1;
    while ((1ull << p->key_bits) < p->nbins) ++p->key_bits;
    const uint64_t cap = std::min<uint64_t>(p->nbins, S * desc.width * desc.width / desc.heavy_threshold + 1);
    p->heavy_capacity = static_cast<uint32_t>(cap);
    size_t total = 0;
    p->weight = alloc<float2>(S, &total);
    p->rank = alloc<uint32_t>(S, &total);
    p->src = alloc<uint32_t>(S, &total);
    p->bins = alloc<uint32_t>(S, &total);
    p->cell_start = alloc<uint32_t>(p->nbins + 1, &total);
    p->taps = alloc<float>(S * 2 * desc.width, &total);
    p->heavy_flag = alloc<uint8_t>(p->nbins, &total);
    p->heavy_list = alloc<uint32_t>(cap, &total);
    p->frac = alloc<double>(2 * S, &total);
    p->sorted_keys = alloc<uint32_t>(S, &total);
    p->counters = alloc<uint32_t>(4, &total);
    p->sort_scratch = alloc<char>(radix_sort_scratch_bytes(S), &total);
    if (p->d.kernel == kKernelKaiserBessel) {
      p->deapod_rows = alloc<float>(desc.image_rows, &total);
      p->deapod_cols = alloc<float>(desc.image_cols, &total);
      const double i0b = host_i0(p->d.beta);
      hipLaunchKernelGGL(deapod_kernel, dim3(ceil_div(desc.image_rows, 256)), dim3(256), 0, 0, p->deapod_rows,
                         desc.image_rows, desc.grid_rows, static_cast<int>(desc.width), p->d.beta, i0b);
      KSJ_LAUNCH_CHECK();
      hipLaunchKernelGGL(deapod_kernel, dim3(ceil_div(desc.image_cols, 256)), dim3(256), 0, 0, p->deapod_cols,
                         desc.image_cols, desc.grid_cols, static_cast<int>(desc.width), p->d.beta, i0b);
      KSJ_LAUNCH_CHECK();
      KSJ_HIP(hipStreamSynchronize(0));
    }
    p->device_bytes = total;
  } catch (...) {
    plan2_destroy(p);
    throw;
  }
  return p;
}

void plan2_destroy(Plan2* p) {
  if (!p) return;
  void* ptrs[] = {p->weight, p->rank, p->src, p->bins, p->cell_start, p->taps, p->heavy_flag, p->heavy_list,
                  p->frac, p->sorted_keys, p->counters, p->sort_scratch, p->deapod_rows, p->deapod_cols};
  for (void* q : ptrs)
    if (q) (void)hipFree(q);
  delete p;
}

void plan2_set_trajectory(Plan2* p, const float* d_traj, const float* d_dcf, hipStream_t stream) {
  require(p != nullptr && d_traj != nullptr, "plan and trajectory must be non-null");
  require(p->d.density != kDensityGiven || d_dcf != nullptr, "density weights must be given");
  const uint32_t S = p->d.samples;
  p->has_trajectory = false;
  KSJ_HIP(hipMemsetAsync(p->counters, 0, 4 * sizeof(uint32_t), stream));
  CoordArgs ca{};
  ca.traj = d_traj;
  ca.dcf = d_dcf;
  ca.samples = S;
  ca.grid_rows = p->d.grid_rows;
  ca.grid_cols = p->d.grid_cols;
  ca.density = p->d.density;
  ca.shift_r = (p->d.centred ? static_cast<double>(p->d.image_rows / 2) : 0.0) - p->d.row_origin;
  ca.shift_c = (p->d.centred ? static_cast<double>(p->d.image_cols / 2) : 0.0) - p->d.col_origin;
  ca.bins = p->bins;
  ca.frac = p->frac;
  ca.weight = p->weight;
  ca.err = p->counters + 1;
  hipLaunchKernelGGL(coords_kernel, dim3(ceil_div(S, 256)), dim3(256), 0, stream, ca);
  KSJ_LAUNCH_CHECK();
  radix_sort_pairs(p->bins, nullptr, p->sorted_keys, p->src, S, p->key_bits, p->sort_scratch, stream);
  hipLaunchKernelGGL(cell_start_kernel, dim3(ceil_div(static_cast<uint64_t>(S) + 1, 256)), dim3(256), 0, stream,
                     p->sorted_keys, S, p->nbins, p->cell_start);
  KSJ_LAUNCH_CHECK();
  TapArgs ta{};
  ta.src = p->src;
  ta.frac = p->frac;
  ta.samples = S;
  ta.kernel = p->d.kernel;
  ta.width = static_cast<int>(p->d.width);
  ta.beta = p->d.beta;
  ta.i0_beta = p->d.kernel == kKernelKaiserBessel ? host_i0(p->d.beta) : 1.0;
  ta.rank = p->rank;
  ta.taps = p->taps;
  hipLaunchKernelGGL(taps_kernel, dim3(ceil_div(S, 256)), dim3(256), 0, stream, ta);
  KSJ_LAUNCH_CHECK();
  hipLaunchKernelGGL(heavy_kernel, dim3(ceil_div(p->nbins, 256)), dim3(256), 0, stream, p->cell_start,
                     p->d.grid_rows, p->d.grid_cols, static_cast<int>(p->d.width), p->d.heavy_threshold,
                     p->heavy_flag, p->heavy_list, p->heavy_capacity, p->counters);
  KSJ_LAUNCH_CHECK();
  uint32_t host_counters[4] = {0, 0, 0, 0};
  KSJ_HIP(hipMemcpyAsync(host_counters, p->counters, sizeof(host_counters), hipMemcpyDeviceToHost, stream));
  KSJ_HIP(hipStreamSynchronize(stream));
  if (host_counters[1] & kErrBadCoordinate)
    throw Error(1, "trajectory must be finite and within [-pi, pi] radians per pixel");
  require(host_counters[0] <= p->heavy_capacity, "heavy-cell list overflow");
  p->heavy_count = host_counters[0];
  p->has_trajectory = true;
}

void golden_angle2(float* d_traj, uint32_t spokes, uint32_t readout, hipStream_t stream) {
  const uint64_t n = static_cast<uint64_t>(spokes) * readout;
  if (n == 0) return;
  hipLaunchKernelGGL(golden_angle_kernel, dim3(ceil_div(n, 256)), dim3(256), 0, stream, d_traj, spokes, readout);
  KSJ_LAUNCH_CHECK();
}

void ramp_dcf2(const float* d_traj, float* d_dcf, uint64_t samples, hipStream_t stream) {
  if (samples == 0) return;
  hipLaunchKernelGGL(ramp_dcf_kernel, dim3(ceil_div(samples, 256)), dim3(256), 0, stream, d_traj, d_dcf, samples);
  KSJ_LAUNCH_CHECK();
}

}  // namespace ksj::gpu
