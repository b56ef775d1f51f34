// 3-D trajectory plan kernels: FP64 coordinates, binning, taps, 2x2x2-block gather
// schedule, deapodization tables, golden-means trajectory.
#include "kb.hpp"
#include "plan3.hpp"
#include "sort.hpp"

#include <algorithm>
#include <cmath>

namespace ksj::gpu {
namespace {

struct Coord3Args {
  const float* traj;  // [S][3]
  const float* dcf;
  uint32_t samples, g[3], density;
  double shift[3];
  uint32_t* bins;
  double* frac;
  float2* weight;
  uint32_t* err;
};

__global__ void coords3_kernel(Coord3Args a) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= a.samples) return;
  constexpr float kPiF = 3.14159265358979323846f;
  float w[3];
  uint32_t cell[3] = {0, 0, 0};
  double fr[3] = {0.0, 0.0, 0.0};
  bool ok = true;
  for (int d = 0; d < 3; ++d) {
    w[d] = a.traj[3 * static_cast<uint64_t>(s) + d];
    ok = ok && isfinite(w[d]) && w[d] >= -kPiF && w[d] <= kPiF && grid_coordinate(w[d], a.g[d], &cell[d], &fr[d]);
  }
  if (!ok) {
    atomicOr(a.err, kErrBadCoordinate);
    for (int d = 0; d < 3; ++d) {
      cell[d] = 0;
      fr[d] = 0.0;
    }
  }
  a.bins[s] = (cell[0] * a.g[1] + cell[1]) * a.g[2] + cell[2];
  for (int d = 0; d < 3; ++d) a.frac[3 * static_cast<uint64_t>(s) + d] = fr[d];
  const double wz = w[0], wy = w[1], wx = w[2];
  double dcf = 1.0;
  if (a.density == kDensityRampSquared) dcf = wz * wz + wy * wy + wx * wx;
  else if (a.density == kDensityGiven) dcf = static_cast<double>(a.dcf[s]);
  double sn, cs;
  sincos(wz * a.shift[0] + wy * a.shift[1] + wx * a.shift[2], &sn, &cs);
  a.weight[s] = make_float2(static_cast<float>(dcf * cs), static_cast<float>(dcf * sn));
}

// Bin occupancy counts (integer atomics: the totals do not depend on thread order).
__global__ void bin3_count_kernel(const uint32_t* bins, uint32_t samples, uint32_t* count) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s < samples) atomicAdd(count + bins[s], 1u);
}

__global__ void taps3_kernel(const uint32_t* src, const double* frac, uint32_t samples, int width, double beta,
                             double i0_beta, uint32_t* rank, float* taps) {
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= samples) return;
  const uint32_t s = src[k];
  rank[s] = k;
  const int lead = width / 2 - 1;
  float* out = taps + static_cast<uint64_t>(k) * 3 * width;
  for (int d = 0; d < 3; ++d) {
    const double f = frac[3 * static_cast<uint64_t>(s) + d];
    for (int t = 0; t < width; ++t)
      out[d * width + t] = static_cast<float>(
        kernel_value(static_cast<double>(t - lead) - f, kKernelKaiserBessel, width, beta, i0_beta));
  }
}

// ----------------------------------------------------------------------------- schedule
// Window of block (bz, by, bx): bins 2b - W/2 + r per axis (wrapped), r in [0, W];
// cell 2b + c of an axis takes tap c - r + W - 1 of a sample in window bin r.
struct Sched3Args {
  const uint32_t* start;
  const float* taps;  // [S][3W]
  uint32_t g[3], blocks[3];
  uint64_t nblocks;
  int width;
  uint32_t* off;      // [nblocks + 1]: counts, then offsets
  uint32_t* visit_k;
  float4* visit_w;
};

__device__ inline uint32_t wrap3(int v, int n) { return static_cast<uint32_t>(((v % n) + n) % n); }

__device__ inline void block_origin(const Sched3Args& a, uint64_t blk, int* o) {
  const uint32_t bx = static_cast<uint32_t>(blk % a.blocks[2]);
  const uint64_t r = blk / a.blocks[2];
  const uint32_t by = static_cast<uint32_t>(r % a.blocks[1]), bz = static_cast<uint32_t>(r / a.blocks[1]);
  o[0] = 2 * static_cast<int>(bz) - a.width / 2;
  o[1] = 2 * static_cast<int>(by) - a.width / 2;
  o[2] = 2 * static_cast<int>(bx) - a.width / 2;
}

__global__ void visit3_count_kernel(Sched3Args a) {
  const uint64_t blk = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (blk > a.nblocks) return;
  if (blk == a.nblocks) {
    a.off[blk] = 0;
    return;
  }
  int o[3];
  block_origin(a, blk, o);
  const int W = a.width;
  uint32_t n = 0;
  for (int rz = 0; rz <= W; ++rz) {
    const uint64_t pz = wrap3(o[0] + rz, static_cast<int>(a.g[0]));
    for (int ry = 0; ry <= W; ++ry) {
      const uint64_t row = (pz * a.g[1] + wrap3(o[1] + ry, static_cast<int>(a.g[1]))) * a.g[2];
      for (int rx = 0; rx <= W; ++rx) {
        const uint64_t b = row + wrap3(o[2] + rx, static_cast<int>(a.g[2]));
        n += a.start[b + 1] - a.start[b];
      }
    }
  }
  a.off[blk] = n;
}

__device__ inline float tap_or_zero(const float* t, int i, int W) { return (i >= 0 && i < W) ? t[i] : 0.f; }

__global__ void visit3_fill_kernel(Sched3Args a) {
  const uint64_t blk = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (blk >= a.nblocks) return;
  int o[3];
  block_origin(a, blk, o);
  const int W = a.width;
  uint32_t pos = a.off[blk];
  for (int rz = 0; rz <= W; ++rz) {
    const uint64_t pz = wrap3(o[0] + rz, static_cast<int>(a.g[0]));
    for (int ry = 0; ry <= W; ++ry) {
      const uint64_t row = (pz * a.g[1] + wrap3(o[1] + ry, static_cast<int>(a.g[1]))) * a.g[2];
      for (int rx = 0; rx <= W; ++rx) {
        const uint64_t b = row + wrap3(o[2] + rx, static_cast<int>(a.g[2]));
        for (uint32_t k = a.start[b]; k < a.start[b + 1]; ++k) {
          const float* t = a.taps + static_cast<uint64_t>(k) * 3 * W;
          float wz[2], wy[2], wx[2];
          for (int c = 0; c < 2; ++c) {
            wz[c] = tap_or_zero(t, c - rz + W - 1, W);
            wy[c] = tap_or_zero(t + W, c - ry + W - 1, W);
            wx[c] = tap_or_zero(t + 2 * W, c - rx + W - 1, W);
          }
          a.visit_k[pos] = k;
          a.visit_w[2 * static_cast<uint64_t>(pos)] =
            make_float4(wz[0] * wy[0] * wx[0], wz[0] * wy[0] * wx[1], wz[0] * wy[1] * wx[0], wz[0] * wy[1] * wx[1]);
          a.visit_w[2 * static_cast<uint64_t>(pos) + 1] =
            make_float4(wz[1] * wy[0] * wx[0], wz[1] * wy[0] * wx[1], wz[1] * wy[1] * wx[0], wz[1] * wy[1] * wx[1]);
          ++pos;
        }
      }
    }
  }
}

__global__ void item3_count_kernel(const uint32_t* off, uint64_t nblocks, uint32_t item_visits, uint32_t* first,
                                   uint32_t* multi, uint32_t* counter) {
  const uint64_t b = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (b > nblocks) return;
  if (b == nblocks) {
    first[b] = 0;
    return;
  }
  const uint32_t v = off[b + 1] - off[b];
  const uint32_t n = v == 0 ? 1u : (v + item_visits - 1) / item_visits;
  first[b] = n;
  if (n > 1) multi[atomicAdd(counter, 1u)] = static_cast<uint32_t>(b);
}

__global__ void slot3_count_kernel(const uint32_t* first, uint64_t nblocks, uint32_t* slots) {
  const uint64_t b = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (b > nblocks) return;
  const uint32_t n = b < nblocks ? first[b + 1] - first[b] : 0u;
  slots[b] = n > 1 ? n : 0u;
}

__global__ void item3_fill_kernel(const uint32_t* first, const uint32_t* slot, const uint32_t* off,
                                  uint32_t item_visits, uint64_t nblocks, uint4* info) {
  const uint64_t b = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (b >= nblocks) return;
  const uint32_t vb = off[b], ve = off[b + 1];
  const uint32_t n = first[b + 1] - first[b];
  for (uint32_t j = 0; j < n; ++j) {
    const uint32_t v0 = vb + j * item_visits;
    const uint32_t v1 = v0 + item_visits < ve ? v0 + item_visits : ve;
    info[first[b] + j] = make_uint4(static_cast<uint32_t>(b), v0, v1, n > 1 ? slot[b] + j : 0xFFFFFFFFu);
  }
}

__global__ void golden_means_kernel(float* traj, uint32_t spokes, uint32_t readout) {
  const uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= static_cast<uint64_t>(spokes) * readout) return;
  const uint32_t p = static_cast<uint32_t>(i / readout), j = static_cast<uint32_t>(i % readout);
  const double g1 = 0.465571231876768, g2 = 0.682327803828019;
  const double c = fmod(static_cast<double>(p) * g1, 1.0);
  const double az = kTwoPi * fmod(static_cast<double>(p) * g2, 1.0);
  const double sz = sqrt(1.0 - c * c);
  double sn, cs;
  sincos(az, &sn, &cs);
  const double k = kTwoPi * (static_cast<double>(j) - static_cast<double>(readout / 2)) / static_cast<double>(readout);
  traj[3 * i] = static_cast<float>(k * c);
  traj[3 * i + 1] = static_cast<float>(k * (sz * sn));
  traj[3 * i + 2] = static_cast<float>(k * (sz * cs));
}

template <typename T> T* alloc3(size_t count, size_t* total) {
  T* p = nullptr;
  const size_t bytes = count * sizeof(T);
  KSJ_HIP(hipMalloc(&p, bytes == 0 ? 1 : bytes));
  *total += bytes;
  return p;
}

}  // namespace

Plan3* plan3_create(const Plan3Desc& desc) {
  for (int d = 0; d < 3; ++d) {
    require(desc.image[d] >= 1, "plan image dimensions must be positive");
    require(desc.grid[d] >= desc.image[d], "grid must not be smaller than the image");
    require(desc.grid[d] <= 2048, "3-D grid side must be <= 2048");
  }
  const uint64_t nbins = static_cast<uint64_t>(desc.grid[0]) * desc.grid[1] * desc.grid[2];
  require(nbins < (1ull << 32) - 1, "3-D grid must have fewer than 2^32 points");
  require(desc.width == 2 || desc.width == 4 || desc.width == 6 || desc.width == 8, "kernel width must be 2, 4, 6 or 8");
  require(desc.density == kDensityNone || desc.density == kDensityRampSquared || desc.density == kDensityGiven,
          "3-D density must be none, ramp squared or given");
  require(desc.samples >= 1, "plan needs at least one sample");
  require(desc.item_visits >= 1 && desc.item_visits <= 4096, "item visits must be in [1, 4096]");
  // A sample reaches at most W/2 + 1 blocks per axis; the visit offsets are 32-bit.
  const uint64_t reach = desc.width / 2 + 1;
  require(static_cast<uint64_t>(desc.samples) * reach * reach * reach < (1ull << 32),
          "too many samples for one 3-D plan (visit count must stay below 2^32)");
  auto* p = new Plan3();
  p->d = desc;
  if (p->d.beta <= 0.0) {
    const double sigma = static_cast<double>(desc.grid[2]) / desc.image[2];
    const double a = static_cast<double>(desc.width) / sigma * (sigma - 0.5);
    p->d.beta = kPi * std::sqrt(a * a - 0.8);
  }
  try {
    KSJ_HIP(hipGetDevice(&p->device));
    const uint64_t S = desc.samples;
    p->nbins = nbins;
    p->key_bits = 1;
    while ((1ull << p->key_bits) < nbins) ++p->key_bits;
    for (int d = 0; d < 3; ++d) p->blocks[d] = (desc.grid[d] + 1) / 2;
    p->nblocks = static_cast<uint64_t>(p->blocks[0]) * p->blocks[1] * p->blocks[2];
    size_t total = 0;
    p->weight = alloc3<float2>(S, &total);
    p->rank = alloc3<uint32_t>(S, &total);
    p->src = alloc3<uint32_t>(S, &total);
    p->bins = alloc3<uint32_t>(S, &total);
    p->cell_start = alloc3<uint32_t>(nbins + 1, &total);
    p->taps = alloc3<float>(S * 3 * desc.width, &total);
    p->block_visit_off = alloc3<uint32_t>(p->nblocks + 1, &total);
    p->block_first_item = alloc3<uint32_t>(p->nblocks + 1, &total);
    p->block_first_slot = alloc3<uint32_t>(p->nblocks + 1, &total);
    p->multi_blocks = alloc3<uint32_t>(p->nblocks, &total);
    p->frac = alloc3<double>(3 * S, &total);
    p->sorted_keys = alloc3<uint32_t>(S, &total);
    p->counters = alloc3<uint32_t>(4, &total);
    const uint64_t scan_n = std::max<uint64_t>(std::max<uint64_t>(S, p->nblocks + 1), nbins + 1);
    p->sort_scratch = alloc3<char>(std::max(radix_sort_scratch_bytes(scan_n), scan_scratch_bytes(scan_n)), &total);
    for (int d = 0; d < 3; ++d) {
      p->deapod[d] = alloc3<float>(desc.image[d], &total);
      deapod_table(p->deapod[d], desc.image[d], desc.grid[d], static_cast<int>(desc.width), p->d.beta, 0);
    }
    KSJ_HIP(hipStreamSynchronize(0));
    p->device_bytes = total;
  } catch (...) {
    plan3_destroy(p);
    throw;
  }
  return p;
}

void plan3_destroy(Plan3* p) {
  if (!p) return;
  void* ptrs[] = {p->weight,     p->rank,        p->src,           p->bins,       p->cell_start,
                  p->taps,       p->block_visit_off, p->visit_k,   p->visit_w,    p->block_first_item,
                  p->block_first_slot, p->item_info, p->multi_blocks, p->frac,     p->sorted_keys,
                  p->counters,   p->sort_scratch, p->deapod[0],    p->deapod[1],  p->deapod[2]};
  for (void* q : ptrs)
    if (q) (void)hipFree(q);
  delete p;
}

void plan3_set_trajectory(Plan3* p, const float* d_traj, const float* d_dcf, hipStream_t stream) {
  require(p != nullptr && d_traj != nullptr, "plan and trajectory must be non-null");
  require(p->d.density != kDensityGiven || d_dcf != nullptr, "density weights must be given");
  const uint32_t S = p->d.samples;
  p->has_trajectory = false;
  KSJ_HIP(hipMemsetAsync(p->counters, 0, 4 * sizeof(uint32_t), stream));
  Coord3Args ca{};
  ca.traj = d_traj;
  ca.dcf = d_dcf;
  ca.samples = S;
  ca.density = p->d.density;
  for (int d = 0; d < 3; ++d) {
    ca.g[d] = p->d.grid[d];
    ca.shift[d] = static_cast<double>(p->d.image[d] / 2) - p->d.origin[d];
  }
  ca.bins = p->bins;
  ca.frac = p->frac;
  ca.weight = p->weight;
  ca.err = p->counters + 1;
  hipLaunchKernelGGL(coords3_kernel, dim3(ceil_div(S, 256)), dim3(256), 0, stream, ca);
  KSJ_LAUNCH_CHECK();
  radix_sort_pairs(p->bins, nullptr, p->sorted_keys, p->src, S, p->key_bits, p->sort_scratch, stream);
  // cell_start by counting and scanning: a 3-D radial ball leaves long runs of empty
  // bins (the cube's centre in wrapped index space) that a per-sample fill would walk.
  KSJ_HIP(hipMemsetAsync(p->cell_start, 0, (p->nbins + 1) * sizeof(uint32_t), stream));
  hipLaunchKernelGGL(bin3_count_kernel, dim3(ceil_div(S, 256)), dim3(256), 0, stream, p->bins, S, p->cell_start);
  KSJ_LAUNCH_CHECK();
  exclusive_scan_u32(p->cell_start, p->cell_start, p->nbins + 1, p->sort_scratch, stream);
  hipLaunchKernelGGL(taps3_kernel, dim3(ceil_div(S, 256)), dim3(256), 0, stream, p->src, p->frac, S,
                     static_cast<int>(p->d.width), p->d.beta, host_i0(p->d.beta), p->rank, p->taps);
  KSJ_LAUNCH_CHECK();
  uint32_t host_counters[4] = {0, 0, 0, 0};
  KSJ_HIP(hipMemcpyAsync(host_counters, p->counters, sizeof(host_counters), hipMemcpyDeviceToHost, stream));
  KSJ_HIP(hipStreamSynchronize(stream));
  if (host_counters[1] & kErrBadCoordinate)
    throw Error(1, "trajectory must be finite and within [-pi, pi] radians per pixel");

  Sched3Args sa{};
  sa.start = p->cell_start;
  sa.taps = p->taps;
  for (int d = 0; d < 3; ++d) {
    sa.g[d] = p->d.grid[d];
    sa.blocks[d] = p->blocks[d];
  }
  sa.nblocks = p->nblocks;
  sa.width = static_cast<int>(p->d.width);
  sa.off = p->block_visit_off;
  hipLaunchKernelGGL(visit3_count_kernel, dim3(ceil_div(p->nblocks + 1, 256)), dim3(256), 0, stream, sa);
  KSJ_LAUNCH_CHECK();
  exclusive_scan_u32(p->block_visit_off, p->block_visit_off, p->nblocks + 1, p->sort_scratch, stream);
  uint32_t visits = 0;
  KSJ_HIP(hipMemcpyAsync(&visits, p->block_visit_off + p->nblocks, sizeof(uint32_t), hipMemcpyDeviceToHost, stream));
  KSJ_HIP(hipStreamSynchronize(stream));
  if (visits > p->visit_capacity || p->visit_k == nullptr) {
    if (p->visit_k) KSJ_HIP(hipFree(p->visit_k));
    if (p->visit_w) KSJ_HIP(hipFree(p->visit_w));
    p->visit_k = nullptr;
    p->visit_w = nullptr;
    p->visit_capacity = std::max<uint64_t>(visits, p->visit_capacity);
    KSJ_HIP(hipMalloc(&p->visit_k, std::max<size_t>(1, p->visit_capacity) * sizeof(uint32_t)));
    KSJ_HIP(hipMalloc(&p->visit_w, std::max<size_t>(1, p->visit_capacity) * 2 * sizeof(float4)));
  }
  p->visit_count = visits;
  sa.visit_k = p->visit_k;
  sa.visit_w = p->visit_w;
  hipLaunchKernelGGL(visit3_fill_kernel, dim3(ceil_div(p->nblocks, 256)), dim3(256), 0, stream, sa);
  KSJ_LAUNCH_CHECK();

  KSJ_HIP(hipMemsetAsync(p->counters, 0, sizeof(uint32_t), stream));
  hipLaunchKernelGGL(item3_count_kernel, dim3(ceil_div(p->nblocks + 1, 256)), dim3(256), 0, stream,
                     p->block_visit_off, p->nblocks, p->d.item_visits, p->block_first_item, p->multi_blocks,
                     p->counters);
  KSJ_LAUNCH_CHECK();
  exclusive_scan_u32(p->block_first_item, p->block_first_item, p->nblocks + 1, p->sort_scratch, stream);
  uint32_t items = 0, multi = 0;
  KSJ_HIP(hipMemcpyAsync(&items, p->block_first_item + p->nblocks, sizeof(uint32_t), hipMemcpyDeviceToHost, stream));
  KSJ_HIP(hipMemcpyAsync(&multi, p->counters, sizeof(uint32_t), hipMemcpyDeviceToHost, stream));
  KSJ_HIP(hipStreamSynchronize(stream));
  if (items > p->item_capacity || p->item_info == nullptr) {
    if (p->item_info) KSJ_HIP(hipFree(p->item_info));
    p->item_info = nullptr;
    p->item_capacity = std::max(items, p->item_capacity);
    KSJ_HIP(hipMalloc(&p->item_info, std::max<size_t>(1, p->item_capacity) * sizeof(uint4)));
  }
  p->item_count = items;
  hipLaunchKernelGGL(slot3_count_kernel, dim3(ceil_div(p->nblocks + 1, 256)), dim3(256), 0, stream,
                     p->block_first_item, p->nblocks, p->block_first_slot);
  KSJ_LAUNCH_CHECK();
  exclusive_scan_u32(p->block_first_slot, p->block_first_slot, p->nblocks + 1, p->sort_scratch, stream);
  uint32_t slots = 0;
  KSJ_HIP(hipMemcpyAsync(&slots, p->block_first_slot + p->nblocks, sizeof(uint32_t), hipMemcpyDeviceToHost, stream));
  hipLaunchKernelGGL(item3_fill_kernel, dim3(ceil_div(p->nblocks, 256)), dim3(256), 0, stream, p->block_first_item,
                     p->block_first_slot, p->block_visit_off, p->d.item_visits, p->nblocks, p->item_info);
  KSJ_LAUNCH_CHECK();
  KSJ_HIP(hipStreamSynchronize(stream));
  p->slot_count = slots;
  p->multi_count = multi;
  p->has_trajectory = true;
}

void golden_means3(float* d_traj, uint32_t spokes, uint32_t readout, hipStream_t stream) {
  const uint64_t n = static_cast<uint64_t>(spokes) * readout;
  if (n == 0) return;
  hipLaunchKernelGGL(golden_means_kernel, dim3(ceil_div(n, 256)), dim3(256), 0, stream, d_traj, spokes, readout);
  KSJ_LAUNCH_CHECK();
}

}  // namespace ksj::gpu
