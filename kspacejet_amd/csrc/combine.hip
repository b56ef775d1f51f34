// Coil combination kernels.  All HBM-bound: one read of the coil data, one write.
#include "combine.hpp"

#include "fused.hpp"

#include <algorithm>

namespace ksj::gpu {
namespace {

constexpr int kTilePixels = 64;

// One thread per pixel; coil values are read through LDS so a wave's loads are contiguous.
__global__ __launch_bounds__(256) void rss_kernel(const float2* __restrict__ in, float* __restrict__ out,
                                                  uint64_t pixels, uint32_t coils) {
  extern __shared__ float2 tile[];  // [blockDim.x][coils + 1]
  const uint64_t p0 = static_cast<uint64_t>(blockIdx.x) * blockDim.x;
  const uint32_t stride = coils + 1;
  const uint64_t count = (p0 + blockDim.x <= pixels ? blockDim.x : pixels - p0) * coils;
  const float2* src = in + p0 * coils;
  for (uint64_t i = threadIdx.x; i < count; i += blockDim.x) {
    const uint32_t pl = static_cast<uint32_t>(i / coils), c = static_cast<uint32_t>(i % coils);
    tile[pl * stride + c] = src[i];
  }
  __syncthreads();
  const uint64_t p = p0 + threadIdx.x;
  if (p >= pixels) return;
  float acc = 0.f;
  for (uint32_t c = 0; c < coils; ++c) {
    const float2 v = tile[threadIdx.x * stride + c];
    acc += norm2(v);
  }
  out[p] = sqrtf(acc);
}

__global__ __launch_bounds__(256) void major_to_last_kernel(const float2* __restrict__ in, float2* __restrict__ last,
                                                            float* __restrict__ rss, uint64_t pixels, uint32_t coils) {
  extern __shared__ float2 tile[];  // [kTilePixels][coils + 1]
  const uint64_t p0 = static_cast<uint64_t>(blockIdx.x) * kTilePixels;
  const uint32_t stride = coils + 1;
  const uint32_t np = static_cast<uint32_t>(p0 + kTilePixels <= pixels ? kTilePixels : pixels - p0);
  for (uint32_t i = threadIdx.x; i < np * coils; i += blockDim.x) {
    const uint32_t c = i / np, pl = i % np;  // consecutive threads read consecutive pixels of one coil plane
    // Each coil-plane value is read once: a non-temporal load.
    typedef float F2v __attribute__((ext_vector_type(2)));
    const F2v v = __builtin_nontemporal_load(reinterpret_cast<const F2v*>(in + static_cast<uint64_t>(c) * pixels + p0 + pl));
    tile[pl * stride + c] = make_float2(v.x, v.y);
  }
  __syncthreads();
  if (last) {
    float2* dst = last + p0 * coils;
    for (uint32_t i = threadIdx.x; i < np * coils; i += blockDim.x) {
      const uint32_t pl = i / coils, c = i % coils;  // consecutive threads write consecutive addresses
      dst[i] = tile[pl * stride + c];
    }
  }
  if (rss && threadIdx.x < np) {
    float acc = 0.f;
    for (uint32_t c = 0; c < coils; ++c) {
      const float2 v = tile[threadIdx.x * stride + c];
      acc += norm2(v);
    }
    rss[p0 + threadIdx.x] = sqrtf(acc);
  }
}

__global__ __launch_bounds__(256) void major_split_kernel(const float2* __restrict__ in, CoilGroupOutputs out,
                                                          uint64_t pixels) {
  extern __shared__ float2 tile[];  // [kTilePixels][total + 1]
  const uint32_t total = out.groups * out.coils;
  const uint64_t p0 = static_cast<uint64_t>(blockIdx.x) * kTilePixels;
  const uint32_t stride = total + 1;
  const uint32_t np = static_cast<uint32_t>(p0 + kTilePixels <= pixels ? kTilePixels : pixels - p0);
  for (uint32_t i = threadIdx.x; i < np * total; i += blockDim.x) {
    const uint32_t c = i / np, pl = i % np;
    // Each coil-plane value is read once: a non-temporal load.
    typedef float F2v __attribute__((ext_vector_type(2)));
    const F2v v = __builtin_nontemporal_load(reinterpret_cast<const F2v*>(in + static_cast<uint64_t>(c) * pixels + p0 + pl));
    tile[pl * stride + c] = make_float2(v.x, v.y);
  }
  __syncthreads();
  for (uint32_t g = 0; g < out.groups; ++g) {
    if (out.coil_last[g]) {
      float2* dst = out.coil_last[g] + p0 * out.coils;
      for (uint32_t i = threadIdx.x; i < np * out.coils; i += blockDim.x) {
        const uint32_t pl = i / out.coils, c = i % out.coils;
        dst[i] = tile[pl * stride + g * out.coils + c];
      }
    }
  }
  for (uint32_t i = threadIdx.x; i < out.groups * np; i += blockDim.x) {
    const uint32_t g = i / np, pl = i % np;
    if (!out.rss[g]) continue;
    float acc = 0.f;
    for (uint32_t c = 0; c < out.coils; ++c) {
      const float2 v = tile[pl * stride + g * out.coils + c];
      acc += norm2(v);
    }
    out.rss[g][p0 + pl] = sqrtf(acc);
  }
}

// Fused items: one coil group each (coil-major [coils][pixels] -> coil-last and/or
// RSS), the arithmetic of major_split_kernel with one group.
struct SplitItemArgs {
  const float2* in;
  float2* coil_last;
  float* rss;
  uint64_t pixels;
  uint32_t coils, pad;
};

__global__ __launch_bounds__(256) void split_fused_kernel(const FusedTable<SplitItemArgs>* t) {
  extern __shared__ float2 tile[];  // [kTilePixels][coils + 1]
  const uint32_t item = fused_item(t, blockIdx.x);
  const SplitItemArgs a = fused_args(t, item);
  const uint64_t p0 = (blockIdx.x - fused_first(t, item)) * kTilePixels;
  const uint32_t stride = a.coils + 1;
  const uint32_t np = static_cast<uint32_t>(p0 + kTilePixels <= a.pixels ? kTilePixels : a.pixels - p0);
  for (uint32_t i = threadIdx.x; i < np * a.coils; i += blockDim.x) {
    const uint32_t c = i / np, pl = i % np;
    tile[pl * stride + c] = a.in[static_cast<uint64_t>(c) * a.pixels + p0 + pl];
  }
  __syncthreads();
  if (a.coil_last) {
    float2* dst = a.coil_last + p0 * a.coils;
    for (uint32_t i = threadIdx.x; i < np * a.coils; i += blockDim.x) {
      const uint32_t pl = i / a.coils, c = i % a.coils;
      dst[i] = tile[pl * stride + c];
    }
  }
  if (a.rss) {
    for (uint32_t pl = threadIdx.x; pl < np; pl += blockDim.x) {
      float acc = 0.f;
      for (uint32_t c = 0; c < a.coils; ++c) {
        const float2 v = tile[pl * stride + c];
        acc += norm2(v);
      }
      a.rss[p0 + pl] = sqrtf(acc);
    }
  }
}

__global__ __launch_bounds__(256) void sense_kernel(const float2* __restrict__ img, const float2* __restrict__ sens,
                                                    float2* __restrict__ out, uint64_t pixels, uint32_t coils) {
  const uint64_t p = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (p >= pixels) return;
  float nr = 0.f, ni = 0.f, den = 0.f;
  for (uint32_t c = 0; c < coils; ++c) {
    const float2 i = img[p * coils + c], s = sens[p * coils + c];
    nr += s.x * i.x + s.y * i.y;
    ni += s.x * i.y - s.y * i.x;
    den += s.x * s.x + s.y * s.y;
  }
  out[p] = den > 0.f ? make_float2(nr / den, ni / den) : make_float2(0.f, 0.f);
}

}  // namespace

void rss_coil_last(const float2* d_in, float* d_out, uint64_t pixels, uint32_t coils, hipStream_t stream) {
  if (pixels == 0) return;
  require(coils >= 1 && coils <= 256, "RSS supports 1..256 coils");
  // Pixels per block: as many as fit 64 KiB of LDS, a multiple of the wave size.
  uint32_t ppb = static_cast<uint32_t>(65536 / ((coils + 1) * sizeof(float2))) / 64 * 64;
  ppb = ppb < 64 ? 64 : (ppb > 256 ? 256 : ppb);
  const size_t shmem = static_cast<size_t>(ppb) * (coils + 1) * sizeof(float2);
  hipLaunchKernelGGL(rss_kernel, dim3(ceil_div(pixels, ppb)), dim3(ppb), shmem, stream, d_in, d_out, pixels, coils);
  KSJ_LAUNCH_CHECK();
}

void coil_major_to_last(const float2* d_in, float2* d_coil_last, float* d_rss, uint64_t pixels, uint32_t coils,
                        hipStream_t stream) {
  if (pixels == 0 || (!d_coil_last && !d_rss)) return;
  require(coils >= 1 && coils <= 64, "coil transpose supports 1..64 coils");
  const size_t shmem = static_cast<size_t>(kTilePixels) * (coils + 1) * sizeof(float2);
  hipLaunchKernelGGL(major_to_last_kernel, dim3(ceil_div(pixels, kTilePixels)), dim3(256), shmem, stream, d_in,
                     d_coil_last, d_rss, pixels, coils);
  KSJ_LAUNCH_CHECK();
}

void coil_major_split(const float2* d_in, const CoilGroupOutputs& out, uint64_t pixels, hipStream_t stream) {
  if (pixels == 0 || out.groups == 0) return;
  require(out.groups <= kMaxCoilGroups && out.coils >= 1 && out.groups * out.coils <= 64,
          "coil split supports at most 64 stacked coils");
  const size_t shmem = static_cast<size_t>(kTilePixels) * (out.groups * out.coils + 1) * sizeof(float2);
  hipLaunchKernelGGL(major_split_kernel, dim3(ceil_div(pixels, kTilePixels)), dim3(256), shmem, stream, d_in, out,
                     pixels);
  KSJ_LAUNCH_CHECK();
}

void sense_combine(const float2* d_images, const float2* d_sens, float2* d_out, uint64_t pixels, uint32_t coils,
                   hipStream_t stream) {
  if (pixels == 0) return;
  require(coils >= 1, "sensitivity combine needs at least one coil");
  hipLaunchKernelGGL(sense_kernel, dim3(ceil_div(pixels, 256)), dim3(256), 0, stream, d_images, d_sens, d_out, pixels,
                     coils);
  KSJ_LAUNCH_CHECK();
}

size_t coil_split_fused_table_bytes() { return fused_table_bytes<SplitItemArgs>(); }

void coil_split_fused(const float2* const* d_in, float2* const* d_coil_last, float* const* d_rss,
                      const uint64_t* pixels, const uint32_t* coils, uint32_t count, void* d_table,
                      hipStream_t stream) {
  FusedTableBuilder<SplitItemArgs> b;
  uint32_t max_coils = 1;
  for (uint32_t i = 0; i < count; ++i) {
    if (!d_coil_last[i] && !d_rss[i]) continue;
    require(coils[i] >= 1 && coils[i] <= 64, "coil split supports 1..64 coils");
    SplitItemArgs a{};
    a.in = d_in[i];
    a.coil_last = d_coil_last[i];
    a.rss = d_rss[i];
    a.pixels = pixels[i];
    a.coils = coils[i];
    b.add(a, ceil_div(pixels[i], static_cast<uint64_t>(kTilePixels)));
    max_coils = std::max(max_coils, coils[i]);
  }
  if (b.t.count == 0) return;
  b.upload(d_table, stream);
  const size_t shmem = static_cast<size_t>(kTilePixels) * (max_coils + 1) * sizeof(float2);
  hipLaunchKernelGGL(split_fused_kernel, dim3(static_cast<uint32_t>(b.blocks())), dim3(256), shmem, stream,
                     static_cast<const FusedTable<SplitItemArgs>*>(d_table));
  KSJ_LAUNCH_CHECK();
}

}  // namespace ksj::gpu
