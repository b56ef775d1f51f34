// Exclusive scan and stable LSD radix sort for wave64 (gfx950).
#include "sort.hpp"

namespace ksj::gpu {
namespace {

constexpr int kScanThreads = 1024;
constexpr int kScanItems = 4;
constexpr int kScanTile = kScanThreads * kScanItems;

constexpr int kSortThreads = 256;
constexpr int kSortRounds = 8;
constexpr int kSortTile = kSortThreads * kSortRounds;
constexpr int kWaves = kSortThreads / 64;

__device__ inline uint32_t wave_inclusive_scan(uint32_t v) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const uint32_t o = __shfl_up(v, d, 64);
    if (lane >= d) v += o;
  }
  return v;
}

// Exclusive scan of one value per thread over the whole block; returns the block total in *total.
__device__ inline uint32_t block_exclusive_scan(uint32_t v, uint32_t* lds_waves, uint32_t* total) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int nwaves = (blockDim.x + 63) >> 6;
  const uint32_t inc = wave_inclusive_scan(v);
  if (lane == 63) lds_waves[wave] = inc;
  __syncthreads();
  if (wave == 0) {
    const uint32_t w = lane < nwaves ? lds_waves[lane] : 0u;
    const uint32_t wi = wave_inclusive_scan(w);
    if (lane < nwaves) lds_waves[lane] = wi - w;
    if (lane == nwaves - 1) lds_waves[32 + 0] = wi;
  }
  __syncthreads();
  *total = lds_waves[32];
  const uint32_t r = lds_waves[wave] + inc - v;
  __syncthreads();
  return r;
}

__global__ __launch_bounds__(kScanThreads) void scan_reduce(const uint32_t* in, uint64_t n, uint32_t* sums) {
  __shared__ uint32_t w[64];
  const uint64_t base = static_cast<uint64_t>(blockIdx.x) * kScanTile + threadIdx.x * kScanItems;
  uint32_t s = 0;
#pragma unroll
  for (int i = 0; i < kScanItems; ++i)
    if (base + i < n) s += in[base + i];
  uint32_t total;
  (void)block_exclusive_scan(s, w, &total);
  if (threadIdx.x == 0) sums[blockIdx.x] = total;
}

__global__ __launch_bounds__(kScanThreads) void scan_top(uint32_t* sums, uint32_t nb) {
  __shared__ uint32_t w[64];
  uint32_t carry = 0;
  for (uint32_t base = 0; base < nb; base += kScanTile) {
    uint32_t v[kScanItems];
    uint32_t s = 0;
#pragma unroll
    for (int i = 0; i < kScanItems; ++i) {
      const uint32_t idx = base + threadIdx.x * kScanItems + i;
      v[i] = idx < nb ? sums[idx] : 0u;
      s += v[i];
    }
    uint32_t total;
    uint32_t run = block_exclusive_scan(s, w, &total) + carry;
#pragma unroll
    for (int i = 0; i < kScanItems; ++i) {
      const uint32_t idx = base + threadIdx.x * kScanItems + i;
      if (idx < nb) sums[idx] = run;
      run += v[i];
    }
    carry += total;
    __syncthreads();
  }
}

__global__ __launch_bounds__(kScanThreads) void scan_apply(const uint32_t* in, uint32_t* out, uint64_t n,
                                                            const uint32_t* sums) {
  __shared__ uint32_t w[64];
  const uint64_t base = static_cast<uint64_t>(blockIdx.x) * kScanTile + threadIdx.x * kScanItems;
  uint32_t v[kScanItems];
  uint32_t s = 0;
#pragma unroll
  for (int i = 0; i < kScanItems; ++i) {
    v[i] = base + i < n ? in[base + i] : 0u;
    s += v[i];
  }
  uint32_t total;
  uint32_t run = block_exclusive_scan(s, w, &total) + sums[blockIdx.x];
#pragma unroll
  for (int i = 0; i < kScanItems; ++i) {
    if (base + i < n) out[base + i] = run;
    run += v[i];
  }
}

__global__ __launch_bounds__(kSortThreads) void radix_hist(const uint32_t* keys, uint64_t n, int shift,
                                                            uint32_t* counts, uint32_t tiles) {
  __shared__ uint32_t hist[256];
  hist[threadIdx.x] = 0;
  __syncthreads();
  const uint64_t base = static_cast<uint64_t>(blockIdx.x) * kSortTile;
#pragma unroll
  for (int r = 0; r < kSortRounds; ++r) {
    const uint64_t i = base + r * kSortThreads + threadIdx.x;
    if (i < n) atomicAdd(&hist[(keys[i] >> shift) & 255u], 1u);
  }
  __syncthreads();
  counts[static_cast<uint64_t>(threadIdx.x) * tiles + blockIdx.x] = hist[threadIdx.x];
}

__global__ __launch_bounds__(kSortThreads) void radix_scatter(const uint32_t* keys_in, const uint32_t* vals_in,
                                                               uint32_t* keys_out, uint32_t* vals_out, uint64_t n,
                                                               int shift, const uint32_t* offsets, uint32_t tiles) {
  __shared__ uint32_t running[256];
  __shared__ uint32_t wave_cnt[kWaves][256];
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  running[t] = offsets[static_cast<uint64_t>(t) * tiles + blockIdx.x];
#pragma unroll
  for (int w = 0; w < kWaves; ++w) wave_cnt[w][t] = 0;
  __syncthreads();
  const uint64_t base = static_cast<uint64_t>(blockIdx.x) * kSortTile;
  const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  for (int r = 0; r < kSortRounds; ++r) {
    const uint64_t i = base + r * kSortThreads + t;
    const bool valid = i < n;
    const uint32_t key = valid ? keys_in[i] : 0u;
    const uint32_t val = valid ? (vals_in ? vals_in[i] : static_cast<uint32_t>(i)) : 0u;
    const uint32_t d = (key >> shift) & 255u;
    uint64_t peers = __ballot(valid);
#pragma unroll
    for (int b = 0; b < 8; ++b) {
      const uint64_t bits = __ballot((d >> b) & 1u);
      peers &= ((d >> b) & 1u) ? bits : ~bits;
    }
    const uint32_t lower = static_cast<uint32_t>(__popcll(peers & lt_mask));
    if (valid && lower == 0) wave_cnt[wave][d] = static_cast<uint32_t>(__popcll(peers));
    __syncthreads();
    if (valid) {
      uint32_t pos = running[d] + lower;
      for (int w = 0; w < wave; ++w) pos += wave_cnt[w][d];
      keys_out[pos] = key;
      vals_out[pos] = val;
    }
    __syncthreads();
    uint32_t add = 0;
#pragma unroll
    for (int w = 0; w < kWaves; ++w) {
      add += wave_cnt[w][t];
      wave_cnt[w][t] = 0;
    }
    running[t] += add;
    __syncthreads();
  }
}

}  // namespace

size_t scan_scratch_bytes(uint64_t n) {
  const uint64_t nb = (n + kScanTile - 1) / kScanTile;
  return (nb + 64) * sizeof(uint32_t);
}

void exclusive_scan_u32(const uint32_t* in, uint32_t* out, uint64_t n, void* scratch, hipStream_t stream) {
  if (n == 0) return;
  const uint64_t nb = (n + kScanTile - 1) / kScanTile;
  require(nb <= 16u * 1024u * 1024u, "scan too large");
  auto* sums = static_cast<uint32_t*>(scratch);
  hipLaunchKernelGGL(scan_reduce, dim3(static_cast<unsigned>(nb)), dim3(kScanThreads), 0, stream, in, n, sums);
  KSJ_LAUNCH_CHECK();
  hipLaunchKernelGGL(scan_top, dim3(1), dim3(kScanThreads), 0, stream, sums, static_cast<uint32_t>(nb));
  KSJ_LAUNCH_CHECK();
  hipLaunchKernelGGL(scan_apply, dim3(static_cast<unsigned>(nb)), dim3(kScanThreads), 0, stream, in, out, n, sums);
  KSJ_LAUNCH_CHECK();
}

size_t radix_sort_scratch_bytes(uint64_t n) {
  const uint64_t tiles = (n + kSortTile - 1) / kSortTile;
  const uint64_t counts = 256 * tiles;
  auto align = [](size_t b) { return (b + 255) & ~size_t(255); };
  return align(counts * 4) + align(scan_scratch_bytes(counts)) + 2 * align(n * 4);
}

void radix_sort_pairs(const uint32_t* keys_in, const uint32_t* vals_in, uint32_t* keys_out, uint32_t* vals_out,
                      uint64_t n, int key_bits, void* scratch, hipStream_t stream) {
  if (n == 0) return;
  require(n < (1ull << 32), "radix sort supports fewer than 2^32 elements");
  const uint64_t tiles = (n + kSortTile - 1) / kSortTile;
  const uint64_t counts_n = 256 * tiles;
  auto align = [](size_t b) { return (b + 255) & ~size_t(255); };
  char* p = static_cast<char*>(scratch);
  auto* counts = reinterpret_cast<uint32_t*>(p);
  p += align(counts_n * 4);
  void* scan_scratch = p;
  p += align(scan_scratch_bytes(counts_n));
  auto* keys_tmp = reinterpret_cast<uint32_t*>(p);
  p += align(n * 4);
  auto* vals_tmp = reinterpret_cast<uint32_t*>(p);
  const int passes = key_bits <= 0 ? 1 : (key_bits + 7) / 8;
  const uint32_t* kin = keys_in;
  const uint32_t* vin = vals_in;
  for (int pass = 0; pass < passes; ++pass) {
    const bool to_out = ((passes - 1 - pass) % 2) == 0;
    uint32_t* kout = to_out ? keys_out : keys_tmp;
    uint32_t* vout = to_out ? vals_out : vals_tmp;
    const int shift = 8 * pass;
    hipLaunchKernelGGL(radix_hist, dim3(static_cast<unsigned>(tiles)), dim3(kSortThreads), 0, stream, kin, n, shift,
                       counts, static_cast<uint32_t>(tiles));
    KSJ_LAUNCH_CHECK();
    exclusive_scan_u32(counts, counts, counts_n, scan_scratch, stream);
    hipLaunchKernelGGL(radix_scatter, dim3(static_cast<unsigned>(tiles)), dim3(kSortThreads), 0, stream, kin, vin,
                       kout, vout, n, shift, counts, static_cast<uint32_t>(tiles));
    KSJ_LAUNCH_CHECK();
    kin = kout;
    vin = vout;
  }
}

}  // namespace ksj::gpu
