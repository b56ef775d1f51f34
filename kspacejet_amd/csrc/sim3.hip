// FP64 Kaiser-Bessel forward NUFFT (type 2) of a 3-D volume, for simulating the k-space
// a kooshball acquisition would record:
//   y_s = sum_v x_v exp(-i w_s . (v - o))       (the exact forward model, w in rad/px)
// approximated as
//   y_s = exp(-i w_s . (N/2 - o)) sum_m phi(m - u_s) X[m mod G],
//   X = FFT(grid), grid[(v - N/2) mod G] = x_v / phi_hat(v - N/2), u_s = w_s G / (2 pi),
// with phi the normalised KB kernel of width W (Beatty beta for the oversampling G/N) and
// phi_hat its continuous Fourier transform -- the same kernel, taps and deapodization the
// adjoint uses, so the two are adjoint up to the density weights.  Everything in FP64:
// this is the data model, not the reconstruction.
//
// The FFT is an in-place iterative radix-2 decimation in time per line in LDS (bit-reversed
// load, log2 G butterfly stages, FP64 twiddles from sincospi), L lines per block taken
// either along memory (x lines: consecutive threads on consecutive elements) or across it
// (y and z lines: consecutive threads on the same element of consecutive lines).
#include "sim3.hpp"

#include "kb.hpp"

#include <algorithm>

namespace ksj::gpu {
namespace {

constexpr int kFftThreads = 256;

struct Axis3 {
  uint32_t n[3];  // image (z, y, x)
  uint32_t g[3];  // grid
  double shift[3];  // N/2 - origin per axis
};

__device__ inline double2 zmul(double2 a, double2 b) { return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }

__device__ inline double kb_ft(double j, uint32_t grid, int width, double beta) {
  const double a = kPi * width * j / static_cast<double>(grid);
  const double z2 = beta * beta - a * a;
  if (z2 > 0.0) {
    const double z = sqrt(z2);
    return sinh(z) / z;
  }
  if (z2 < 0.0) {
    const double z = sqrt(-z2);
    return sin(z) / z;
  }
  return 1.0;
}

__global__ void zero_kernel(double2* grid, uint64_t count) {
  for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < count;
       i += static_cast<uint64_t>(gridDim.x) * blockDim.x)
    grid[i] = make_double2(0.0, 0.0);
}

// grid[(v - N/2) mod G] = x_v / (phi_hat_z phi_hat_y phi_hat_x), phi_hat = (W / I0(beta)) shape.
__global__ void scatter_kernel(const double2* image, Axis3 a, int width, double beta, double i0_beta, double2* grid) {
  const uint64_t total = static_cast<uint64_t>(a.n[0]) * a.n[1] * a.n[2];
  for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < total;
       i += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
    const uint32_t x = static_cast<uint32_t>(i % a.n[2]);
    const uint64_t r = i / a.n[2];
    const uint32_t y = static_cast<uint32_t>(r % a.n[1]), z = static_cast<uint32_t>(r / a.n[1]);
    const uint32_t v[3] = {z, y, x};
    double scale = 1.0;
    uint32_t m[3];
    for (int d = 0; d < 3; ++d) {
      const int64_t t = static_cast<int64_t>(v[d]) - static_cast<int64_t>(a.n[d] / 2);
      scale *= i0_beta / (width * kb_ft(static_cast<double>(t), a.g[d], width, beta));
      m[d] = static_cast<uint32_t>((t + a.g[d]) % a.g[d]);
    }
    const double2 xv = image[i];
    grid[(static_cast<uint64_t>(m[0]) * a.g[1] + m[1]) * a.g[2] + m[2]] = make_double2(xv.x * scale, xv.y * scale);
  }
}

__device__ inline uint32_t bit_reverse(uint32_t e, int bits) { return __brev(e) >> (32 - bits); }

// Forward FFT of `lines` lines of length n = 2^bits: element e of line l at
// base(l) + e * elem_stride, base(l) = (l / per_plane) * plane_stride + (l % per_plane) * line_stride.
struct Fft64Args {
  double2* data;
  uint64_t lines, per_plane, plane_stride, line_stride, elem_stride;
  uint32_t n;
  int bits;
  int lines_fast;  // 1: consecutive threads take the same element of consecutive lines
  int lpb;         // lines per block
};

__global__ __launch_bounds__(kFftThreads) void fft64_kernel(Fft64Args a) {
  extern __shared__ double2 lds[];  // [lpb][n] then the twiddles [n / 2]
  const uint32_t n = a.n;
  double2* tw = lds + static_cast<uint64_t>(a.lpb) * n;
  for (uint32_t k = threadIdx.x; k < n / 2; k += blockDim.x) {
    double s, c;
    sincospi(-2.0 * static_cast<double>(k) / static_cast<double>(n), &s, &c);
    tw[k] = make_double2(c, s);
  }
  const uint64_t l0 = static_cast<uint64_t>(blockIdx.x) * a.lpb;
  const uint32_t count = static_cast<uint32_t>(a.lpb) * n;
  auto base = [&](uint64_t l) { return (l / a.per_plane) * a.plane_stride + (l % a.per_plane) * a.line_stride; };
  for (uint32_t i = threadIdx.x; i < count; i += blockDim.x) {
    uint32_t ly, e;
    if (a.lines_fast) {
      ly = i % a.lpb;
      e = i / a.lpb;
    } else {
      ly = i / n;
      e = i % n;
    }
    const uint64_t l = l0 + ly;
    const double2 v = l < a.lines ? a.data[base(l) + e * a.elem_stride] : make_double2(0.0, 0.0);
    lds[ly * n + bit_reverse(e, a.bits)] = v;
  }
  __syncthreads();
  for (uint32_t h = 1; h < n; h <<= 1) {
    const uint32_t butterflies = static_cast<uint32_t>(a.lpb) * (n / 2);
    for (uint32_t b = threadIdx.x; b < butterflies; b += blockDim.x) {
      const uint32_t ly = b / (n / 2), q = b % (n / 2);
      const uint32_t k = q % h, i = (q / h) * 2 * h + k, j = i + h;
      double2* line = lds + ly * n;
      const double2 w = tw[k * (n / (2 * h))];
      const double2 u = line[i], t = zmul(line[j], w);
      line[i] = make_double2(u.x + t.x, u.y + t.y);
      line[j] = make_double2(u.x - t.x, u.y - t.y);
    }
    __syncthreads();
  }
  for (uint32_t i = threadIdx.x; i < count; i += blockDim.x) {
    uint32_t ly, e;
    if (a.lines_fast) {
      ly = i % a.lpb;
      e = i / a.lpb;
    } else {
      ly = i / n;
      e = i % n;
    }
    const uint64_t l = l0 + ly;
    if (l < a.lines) a.data[base(l) + e * a.elem_stride] = lds[ly * n + e];
  }
}

// y_s = exp(-i w . shift) sum over the W^3 cells m around u_s of phi(m - u_s) X[m mod G].
struct InterpArgs {
  const double2* grid;
  const float* traj;  // [S][3] (z, y, x) rad/px
  double2* out;
  uint64_t samples;
  Axis3 a;
  int width;
  double beta, i0_beta;
};

template <int W>
__global__ __launch_bounds__(256) void interp_kernel(InterpArgs p) {
  const uint64_t s = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (s >= p.samples) return;
  constexpr int lead = W / 2 - 1;
  double taps[3][W];
  uint32_t cells[3][W];
  double phase = 0.0;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const double w = static_cast<double>(p.traj[3 * s + d]);
    phase += w * p.a.shift[d];
    const double g = static_cast<double>(p.a.g[d]);
    double u = w * g / kTwoPi;
    const double f = floor(u);
    const int64_t c = static_cast<int64_t>(f);
    const double frac = u - f;
#pragma unroll
    for (int t = 0; t < W; ++t) {
      taps[d][t] = kernel_value(static_cast<double>(t - lead) - frac, kKernelKaiserBessel, W, p.beta, p.i0_beta);
      const int64_t m = c - lead + t;
      cells[d][t] = static_cast<uint32_t>(((m % static_cast<int64_t>(g)) + static_cast<int64_t>(g)) %
                                          static_cast<int64_t>(g));
    }
  }
  // Rows (tz, ty) in a rolled loop, taps picked by unrolled selects (static register
  // indices), one row of W loads in flight at a time.
  auto pick = [](const double (&v)[W], int i) {
    double r = 0.0;
#pragma unroll
    for (int k = 0; k < W; ++k) r = k == i ? v[k] : r;
    return r;
  };
  auto pick_cell = [](const uint32_t (&v)[W], int i) {
    uint32_t r = 0;
#pragma unroll
    for (int k = 0; k < W; ++k) r = k == i ? v[k] : r;
    return r;
  };
  double re = 0.0, im = 0.0;
#pragma unroll 1
  for (int r = 0; r < W * W; ++r) {
    const int tz = r / W, ty = r % W;
    const double wzy = pick(taps[0], tz) * pick(taps[1], ty);
    const uint64_t row = (static_cast<uint64_t>(pick_cell(cells[0], tz)) * p.a.g[1] + pick_cell(cells[1], ty)) *
                         p.a.g[2];
#pragma unroll
    for (int tx = 0; tx < W; ++tx) {
      const double2 v = p.grid[row + cells[2][tx]];
      const double w = wzy * taps[2][tx];
      re += w * v.x;
      im += w * v.y;
    }
  }
  double sn, cs;
  sincos(-phase, &sn, &cs);
  p.out[s] = make_double2(re * cs - im * sn, re * sn + im * cs);
}

int log2_exact(uint32_t n) {
  int b = 0;
  while ((1u << b) < n) ++b;
  return (1u << b) == n ? b : -1;
}

void fft64_axis(double2* grid, const uint32_t g[3], int axis, hipStream_t stream) {
  Fft64Args f{};
  f.data = grid;
  f.n = g[axis];
  f.bits = log2_exact(f.n);
  const uint64_t gz = g[0], gy = g[1], gx = g[2];
  if (axis == 2) {  // x lines: contiguous elements, one line after another
    f.lines = gz * gy;
    f.per_plane = f.lines;
    f.plane_stride = 0;
    f.line_stride = gx;
    f.elem_stride = 1;
    f.lines_fast = 0;
  } else if (axis == 1) {  // y lines (z, x): element stride G_x
    f.lines = gz * gx;
    f.per_plane = gx;
    f.plane_stride = gy * gx;
    f.line_stride = 1;
    f.elem_stride = gx;
    f.lines_fast = 1;
  } else {  // z lines (y, x): element stride G_y G_x
    f.lines = gy * gx;
    f.per_plane = f.lines;
    f.plane_stride = 0;
    f.line_stride = 1;
    f.elem_stride = gy * gx;
    f.lines_fast = 1;
  }
  // Lines per block: 64 KiB of LDS for the lines (128-byte segments for 512-point y and z
  // lines), next to the twiddles.
  f.lpb = static_cast<int>(std::max<uint32_t>(1, std::min<uint32_t>(16, 4096 / f.n)));
  const size_t shmem = (static_cast<size_t>(f.lpb) * f.n + f.n / 2) * sizeof(double2);
  if (shmem > 65536)
    KSJ_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(&fft64_kernel),
                                hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(shmem)));
  const unsigned blocks = ceil_div(f.lines, static_cast<uint64_t>(f.lpb));
  hipLaunchKernelGGL(fft64_kernel, dim3(blocks), dim3(kFftThreads), shmem, stream, f);
  KSJ_LAUNCH_CHECK();
}

}  // namespace

size_t nufft3_forward_f64_workspace_bytes(const uint32_t image[3]) {
  size_t cells = 1;
  for (int d = 0; d < 3; ++d) cells *= 2ull * image[d];
  return cells * sizeof(double2);
}

void nufft3_forward_f64(const double2* d_image, const uint32_t image[3], const double origin[3], const float* d_traj,
                        uint64_t samples, int width, double2* d_out, void* d_ws, size_t ws_bytes,
                        hipStream_t stream) {
  require(d_image != nullptr && d_traj != nullptr && d_out != nullptr && d_ws != nullptr, "null argument");
  require(width == 4 || width == 6 || width == 8, "forward NUFFT width must be 4, 6 or 8");
  Axis3 a{};
  for (int d = 0; d < 3; ++d) {
    require(image[d] >= 1 && image[d] <= 512, "forward NUFFT volume sides must be 1..512");
    a.n[d] = image[d];
    a.g[d] = 2 * image[d];
    require(log2_exact(a.g[d]) >= 1, "forward NUFFT grid sides (twice the volume) must be powers of two");
    a.shift[d] = static_cast<double>(image[d] / 2) - origin[d];
  }
  require(ws_bytes >= nufft3_forward_f64_workspace_bytes(image), "forward NUFFT workspace too small");
  auto* grid = static_cast<double2*>(d_ws);
  const uint64_t cells = static_cast<uint64_t>(a.g[0]) * a.g[1] * a.g[2];
  const double sigma = 2.0;
  const double ab = static_cast<double>(width) / sigma * (sigma - 0.5);
  const double beta = kPi * std::sqrt(ab * ab - 0.8);
  const double i0b = host_i0(beta);
  hipLaunchKernelGGL(zero_kernel, dim3(4096), dim3(256), 0, stream, grid, cells);
  KSJ_LAUNCH_CHECK();
  hipLaunchKernelGGL(scatter_kernel, dim3(4096), dim3(256), 0, stream, d_image, a, width, beta, i0b, grid);
  KSJ_LAUNCH_CHECK();
  for (int axis = 2; axis >= 0; --axis) fft64_axis(grid, a.g, axis, stream);
  if (samples == 0) return;
  InterpArgs p{};
  p.grid = grid;
  p.traj = d_traj;
  p.out = d_out;
  p.samples = samples;
  p.a = a;
  p.width = width;
  p.beta = beta;
  p.i0_beta = i0b;
  const dim3 grid_dim(ceil_div(samples, 256)), block(256);
  if (width == 4)
    hipLaunchKernelGGL(interp_kernel<4>, grid_dim, block, 0, stream, p);
  else if (width == 6)
    hipLaunchKernelGGL(interp_kernel<6>, grid_dim, block, 0, stream, p);
  else
    hipLaunchKernelGGL(interp_kernel<8>, grid_dim, block, 0, stream, p);
  KSJ_LAUNCH_CHECK();
}

}  // namespace ksj::gpu
