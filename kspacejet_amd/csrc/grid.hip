// Gather gridding kernels (no float atomics, fixed summation order).
#include "grid.hpp"

namespace ksj::gpu {
namespace {

struct PrepArgs {
  const float2* kspace;  // [C][S]
  const float2* weight;  // [S]
  const uint32_t* rank;  // [S]
  float2* sorted;        // [S][Cp]
  uint32_t samples, coils, cp;
};

__global__ __launch_bounds__(256) void prep_kernel(PrepArgs a) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= a.samples) return;
  const float2 w = a.weight[s];
  float4* dst = reinterpret_cast<float4*>(a.sorted + static_cast<uint64_t>(a.rank[s]) * a.cp);
  for (uint32_t g = 0; g < a.cp; g += kCoilGroup) {
    float2 v[kCoilGroup];
#pragma unroll
    for (uint32_t c = 0; c < kCoilGroup; ++c) {
      const uint32_t cc = g + c;
      v[c] = cc < a.coils ? cmul(a.kspace[static_cast<uint64_t>(cc) * a.samples + s], w) : make_float2(0.f, 0.f);
    }
#pragma unroll
    for (uint32_t c = 0; c < kCoilGroup; c += 2) dst[(g + c) / 2] = make_float4(v[c].x, v[c].y, v[c + 1].x, v[c + 1].y);
  }
}

struct GatherArgs {
  const float2* sorted;        // [S][Cp]
  const float* taps;           // [S][2W]
  const uint32_t* start;       // [nbins + 1]
  const uint8_t* heavy_flag;   // [nbins]
  const uint32_t* heavy_list;  // [heavy_count]
  float2* grid;                // [C][nbins]
  uint64_t nbins;
  uint32_t grid_rows, grid_cols, coils, cp;
};

__device__ inline void accumulate(float2* acc, float w, const float2* row) {
  const float4* d = reinterpret_cast<const float4*>(row);
#pragma unroll
  for (uint32_t c = 0; c < kCoilGroup / 2; ++c) {
    const float4 v = d[c];
    acc[2 * c].x += w * v.x;
    acc[2 * c].y += w * v.y;
    acc[2 * c + 1].x += w * v.z;
    acc[2 * c + 1].y += w * v.w;
  }
}

template <int W>
__global__ __launch_bounds__(256) void gather_kernel(GatherArgs a) {
  const uint64_t m = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (m >= a.nbins || a.heavy_flag[m]) return;
  const uint32_t g = blockIdx.y * kCoilGroup;
  const int gr = static_cast<int>(a.grid_rows), gc = static_cast<int>(a.grid_cols);
  const int y = static_cast<int>(m / a.grid_cols), x = static_cast<int>(m % a.grid_cols);
  constexpr int lead = W / 2 - 1;
  float2 acc[kCoilGroup];
#pragma unroll
  for (uint32_t c = 0; c < kCoilGroup; ++c) acc[c] = make_float2(0.f, 0.f);
#pragma unroll
  for (int ty = 0; ty < W; ++ty) {
    int by = y - ty + lead;
    by = by < 0 ? by + gr : (by >= gr ? by - gr : by);
    const uint32_t* row_start = a.start + static_cast<uint64_t>(by) * a.grid_cols;
#pragma unroll
    for (int tx = 0; tx < W; ++tx) {
      int bx = x - tx + lead;
      bx = bx < 0 ? bx + gc : (bx >= gc ? bx - gc : bx);
      const uint32_t k0 = row_start[bx], k1 = row_start[bx + 1];
      for (uint32_t k = k0; k < k1; ++k) {
        const float* tp = a.taps + static_cast<uint64_t>(k) * (2 * W);
        const float w = tp[ty] * tp[W + tx];
        accumulate(acc, w, a.sorted + static_cast<uint64_t>(k) * a.cp + g);
      }
    }
  }
#pragma unroll
  for (uint32_t c = 0; c < kCoilGroup; ++c)
    if (g + c < a.coils) a.grid[static_cast<uint64_t>(g + c) * a.nbins + m] = acc[c];
}

template <int W>
__global__ __launch_bounds__(256) void gather_heavy_kernel(GatherArgs a) {
  __shared__ float2 red[256][kCoilGroup];
  const uint32_t m = a.heavy_list[blockIdx.x];
  const int t = threadIdx.x;
  const int gr = static_cast<int>(a.grid_rows), gc = static_cast<int>(a.grid_cols);
  const int y = static_cast<int>(m / a.grid_cols), x = static_cast<int>(m % a.grid_cols);
  constexpr int lead = W / 2 - 1;
  for (uint32_t g = 0; g < a.cp; g += kCoilGroup) {
    float2 acc[kCoilGroup];
#pragma unroll
    for (uint32_t c = 0; c < kCoilGroup; ++c) acc[c] = make_float2(0.f, 0.f);
    for (int ty = 0; ty < W; ++ty) {
      int by = y - ty + lead;
      by = by < 0 ? by + gr : (by >= gr ? by - gr : by);
      const uint32_t* row_start = a.start + static_cast<uint64_t>(by) * a.grid_cols;
      for (int tx = 0; tx < W; ++tx) {
        int bx = x - tx + lead;
        bx = bx < 0 ? bx + gc : (bx >= gc ? bx - gc : bx);
        const uint32_t k0 = row_start[bx], k1 = row_start[bx + 1];
        for (uint32_t k = k0 + t; k < k1; k += blockDim.x) {
          const float* tp = a.taps + static_cast<uint64_t>(k) * (2 * W);
          accumulate(acc, tp[ty] * tp[W + tx], a.sorted + static_cast<uint64_t>(k) * a.cp + g);
        }
      }
    }
#pragma unroll
    for (uint32_t c = 0; c < kCoilGroup; ++c) red[t][c] = acc[c];
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
      if (t < s) {
#pragma unroll
        for (uint32_t c = 0; c < kCoilGroup; ++c) red[t][c] = cadd(red[t][c], red[t + s][c]);
      }
      __syncthreads();
    }
    if (t < static_cast<int>(kCoilGroup) && g + t < a.coils) a.grid[static_cast<uint64_t>(g + t) * a.nbins + m] = red[0][t];
    __syncthreads();
  }
}

template <int W>
void launch_gather(const GatherArgs& a, uint32_t heavy_count, hipStream_t stream) {
  hipLaunchKernelGGL((gather_kernel<W>), dim3(ceil_div(a.nbins, 256), a.cp / kCoilGroup), dim3(256), 0, stream, a);
  KSJ_LAUNCH_CHECK();
  if (heavy_count > 0) {
    hipLaunchKernelGGL((gather_heavy_kernel<W>), dim3(heavy_count), dim3(256), 0, stream, a);
    KSJ_LAUNCH_CHECK();
  }
}

}  // namespace

void grid_prep(const Plan2& plan, const float2* d_kspace, uint32_t coils, float2* d_sorted, hipStream_t stream) {
  require(plan.has_trajectory, "plan has no trajectory");
  PrepArgs a{};
  a.kspace = d_kspace;
  a.weight = plan.weight;
  a.rank = plan.rank;
  a.sorted = d_sorted;
  a.samples = plan.d.samples;
  a.coils = coils;
  a.cp = padded_coils(coils);
  hipLaunchKernelGGL(prep_kernel, dim3(ceil_div(a.samples, 256)), dim3(256), 0, stream, a);
  KSJ_LAUNCH_CHECK();
}

void grid_gather(const Plan2& plan, const float2* d_sorted, uint32_t coils, float2* d_grid, hipStream_t stream) {
  require(plan.has_trajectory, "plan has no trajectory");
  GatherArgs a{};
  a.sorted = d_sorted;
  a.taps = plan.taps;
  a.start = plan.cell_start;
  a.heavy_flag = plan.heavy_flag;
  a.heavy_list = plan.heavy_list;
  a.grid = d_grid;
  a.nbins = plan.nbins;
  a.grid_rows = plan.d.grid_rows;
  a.grid_cols = plan.d.grid_cols;
  a.coils = coils;
  a.cp = padded_coils(coils);
  switch (plan.d.width) {
    case 2: return launch_gather<2>(a, plan.heavy_count, stream);
    case 4: return launch_gather<4>(a, plan.heavy_count, stream);
    case 6: return launch_gather<6>(a, plan.heavy_count, stream);
    case 8: return launch_gather<8>(a, plan.heavy_count, stream);
    default: throw Error(1, "unsupported kernel width");
  }
}

}  // namespace ksj::gpu
