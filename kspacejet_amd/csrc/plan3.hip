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
// Window of block (bz, by, bx) of side s: bins s b - W/2 + r per axis (wrapped),
// r in [0, W + s - 1); cell s b + c of an axis takes tap c - r + W - 1 of a sample in
// window bin r.
struct Sched3Args {
  const uint32_t* start;
  const float* taps;  // [S][3W]
  uint32_t g[3], blocks[3];
  uint64_t nblocks;
  int width, side;
  uint32_t* off;      // [nblocks + 1]: counts, then offsets
  const uint32_t* chunk_off;  // matrix schedule: first group of four of each block (null otherwise)
  uint32_t* visit_k;
  float4* visit_w;    // 2x2x2 schedule
  uint32_t* visit_pos;  // matrix schedule
};

__device__ inline uint32_t wrap3(int v, int n) { return static_cast<uint32_t>(((v % n) + n) % n); }

__device__ inline void block_origin(const Sched3Args& a, uint64_t blk, int* o) {
  const uint32_t bx = static_cast<uint32_t>(blk % a.blocks[2]);
  const uint64_t r = blk / a.blocks[2];
  const uint32_t by = static_cast<uint32_t>(r % a.blocks[1]), bz = static_cast<uint32_t>(r / a.blocks[1]);
  o[0] = a.side * static_cast<int>(bz) - a.width / 2;
  o[1] = a.side * static_cast<int>(by) - a.width / 2;
  o[2] = a.side * static_cast<int>(bx) - a.width / 2;
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
  const int W = a.width, R = W + a.side - 1;
  uint32_t n = 0;
  for (int rz = 0; rz < R; ++rz) {
    const uint64_t pz = wrap3(o[0] + rz, static_cast<int>(a.g[0]));
    for (int ry = 0; ry < R; ++ry) {
      const uint64_t row = (pz * a.g[1] + wrap3(o[1] + ry, static_cast<int>(a.g[1]))) * a.g[2];
      for (int rx = 0; rx < R; ++rx) {
        const uint64_t b = row + wrap3(o[2] + rx, static_cast<int>(a.g[2]));
        n += a.start[b + 1] - a.start[b];
      }
    }
  }
  a.off[blk] = n;
}

__device__ inline float tap_or_zero(const float* t, int i, int W) { return (i >= 0 && i < W) ? t[i] : 0.f; }

// Matrix schedule: visits (sample, window position), the block's list starting on
// its first group of four (padding written by pad3_visits_kernel).
__global__ void visit3_fill_matrix_kernel(Sched3Args a) {
  const uint64_t blk = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (blk >= a.nblocks) return;
  int o[3];
  block_origin(a, blk, o);
  const int R = a.width + a.side - 1;
  uint32_t pos = 4 * a.chunk_off[blk];
  for (int rz = 0; rz < R; ++rz) {
    const uint64_t pz = wrap3(o[0] + rz, static_cast<int>(a.g[0]));
    for (int ry = 0; ry < R; ++ry) {
      const uint64_t row = (pz * a.g[1] + wrap3(o[1] + ry, static_cast<int>(a.g[1]))) * a.g[2];
      for (int rx = 0; rx < R; ++rx) {
        const uint64_t b = row + wrap3(o[2] + rx, static_cast<int>(a.g[2]));
        const uint32_t code = static_cast<uint32_t>(rz) | static_cast<uint32_t>(ry) << 4 | static_cast<uint32_t>(rx) << 8;
        for (uint32_t k = a.start[b]; k < a.start[b + 1]; ++k) {
          a.visit_k[pos] = k;
          a.visit_pos[pos] = code;
          ++pos;
        }
      }
    }
  }
}

// Groups of four per block (chunks[nblocks] = 0 for the scan).
__global__ void chunk3_count_kernel(const uint32_t* off, uint64_t nblocks, uint32_t* chunks) {
  const uint64_t b = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (b > nblocks) return;
  chunks[b] = b < nblocks ? (off[b + 1] - off[b] + 3) / 4 : 0u;
}

// Padding visits: the block's first sample at a position outside every window.
__global__ void pad3_visits_kernel(const uint32_t* off, const uint32_t* chunk_off, uint64_t nblocks,
                                   uint32_t* visit_k, uint32_t* visit_pos) {
  const uint64_t b = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (b >= nblocks) return;
  const uint32_t base = 4 * chunk_off[b], end = 4 * chunk_off[b + 1];
  for (uint32_t p = base + (off[b + 1] - off[b]); p < end; ++p) {
    visit_k[p] = visit_k[base];
    visit_pos[p] = kPadVisitPos;
  }
}

// Item cost (groups of four + 1) for the run split; cost[items] = 0.
__global__ void item3_cost_kernel(const uint4* info, uint32_t items, uint32_t* cost) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < items) cost[i] = (info[i].z - info[i].y) / 4 + 1;
  if (i == items) cost[i] = 0;
}

// first[r] = first item whose cost prefix reaches r * run_cost; first[runs] = items.
__global__ void run3_split_kernel(const uint32_t* prefix, uint32_t items, uint32_t runs, uint32_t run_cost,
                                  uint32_t* first) {
  const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r > runs) return;
  if (r == runs) {
    first[r] = items;
    return;
  }
  const uint64_t target = static_cast<uint64_t>(r) * run_cost;
  uint32_t lo = 0, hi = items;
  while (lo < hi) {
    const uint32_t mid = lo + (hi - lo) / 2;
    if (prefix[mid] >= target)
      hi = mid;
    else
      lo = mid + 1;
  }
  first[r] = lo;
}

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

// Item records {block, first visit, end visit, slot or ~0}; off counts units of
// `scale` visits (1, or 4 for the matrix schedule's groups) and items hold `size` units.
__global__ void item3_fill_kernel(const uint32_t* first, const uint32_t* slot, const uint32_t* off, uint32_t size,
                                  uint32_t scale, uint64_t nblocks, uint4* info) {
  const uint64_t b = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (b >= nblocks) return;
  const uint32_t vb = off[b], ve = off[b + 1];
  const uint32_t n = first[b + 1] - first[b];
  for (uint32_t j = 0; j < n; ++j) {
    const uint32_t v0 = vb + j * size;
    const uint32_t v1 = v0 + size < ve ? v0 + size : ve;
    info[first[b] + j] = make_uint4(static_cast<uint32_t>(b), scale * v0, scale * v1, n > 1 ? slot[b] + j : 0xFFFFFFFFu);
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

namespace {

// Block b (row-major z, y, x) has visits.
__global__ void block_visited3_kernel(const uint32_t* chunk_off, uint64_t nblocks, uint8_t* visited) {
  const uint64_t b = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (b < nblocks) visited[b] = chunk_off[b + 1] > chunk_off[b] ? 1 : 0;
}

// Per grid row (z, y): the run of visitless blocks of its block row holding column G_x / 2,
// as cell columns [a, b) (empty when that block has visits).
__global__ void zero_span3_kernel(const uint8_t* visited, uint32_t gz, uint32_t gy, uint32_t gx, uint32_t by_count,
                                  uint32_t bx_count, uint32_t side, uint2* span) {
  const uint64_t r = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (r >= static_cast<uint64_t>(gz) * gy) return;
  const uint32_t z = static_cast<uint32_t>(r / gy), y = static_cast<uint32_t>(r % gy);
  const uint8_t* row = visited + (static_cast<uint64_t>(z / side) * by_count + y / side) * bx_count;
  const uint32_t mid = (gx / 2) / side;
  uint2 s = make_uint2(0, 0);
  if (mid < bx_count && row[mid] == 0) {
    uint32_t lo = mid, hi = mid + 1;
    while (lo > 0 && row[lo - 1] == 0) --lo;
    while (hi < bx_count && row[hi] == 0) ++hi;
    s = make_uint2(lo * side, min(hi * side, gx));
  }
  span[r] = s;
}

// Per z-plane: the run of rows around G_y / 2 whose zero span covers the whole row.
__global__ void zero_rows3_kernel(const uint2* zero_cols, uint32_t gz, uint32_t gy, uint32_t gx, uint2* rows) {
  const uint32_t z = blockIdx.x * blockDim.x + threadIdx.x;
  if (z >= gz) return;
  const uint2* row = zero_cols + static_cast<uint64_t>(z) * gy;
  auto empty = [&](uint32_t y) { return row[y].x == 0 && row[y].y >= gx; };
  const uint32_t mid = gy / 2;
  uint2 s = make_uint2(0, 0);
  if (empty(mid)) {
    uint32_t lo = mid, hi = mid + 1;
    while (lo > 0 && empty(lo - 1)) --lo;
    while (hi < gy && empty(hi)) ++hi;
    s = make_uint2(lo, hi);
  }
  rows[z] = s;
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
  require(desc.item_visits <= 4096, "item visits must be in [1, 4096]");
  auto* p = new Plan3();
  p->d = desc;
  if (p->d.scheme == kSchemeDefault) p->d.scheme = static_cast<uint32_t>(tune_get(kTuneGather3Scheme));
  if (p->d.item_visits == 0)
    p->d.item_visits =
      static_cast<uint32_t>(tune_get(p->d.scheme == kSchemeMatrix ? kTuneMatrix3ItemVisits : kTuneItemVisits));
  if (p->d.run_cost == 0) p->d.run_cost = static_cast<uint32_t>(tune_get(kTuneMatrix3RunCost));
  if (p->d.scheme != kSchemeBlocks && p->d.scheme != kSchemeMatrix) {
    delete p;
    throw Error(1, "3-D gather schedule must be 0 (2x2x2 blocks) or 2 (matrix)");
  }
  if (p->d.run_cost < 1 || p->d.run_cost > 64) {
    delete p;
    throw Error(1, "run cost must be in [1, 64]");
  }
  // A sample reaches at most ceil((W + s - 1) / s) blocks per axis (plus three padding
  // visits per block for the matrix schedule); visit offsets are 32-bit.
  const uint64_t side = p->d.scheme == kSchemeMatrix ? 4 : 2;
  const uint64_t reach = (desc.width + 2 * side - 2) / side;
  const uint64_t pad = p->d.scheme == kSchemeMatrix ? 3 * ((nbins + 63) / 64) : 0;
  if (static_cast<uint64_t>(desc.samples) * reach * reach * reach + pad >= (1ull << 32)) {
    delete p;
    throw Error(1, "too many samples for one 3-D plan (visit count must stay below 2^32)");
  }
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
    p->scheme = p->d.scheme;
    p->side = p->scheme == kSchemeMatrix ? 4u : 2u;
    p->run_cost = p->d.run_cost;
    p->item_visits = p->scheme == kSchemeMatrix ? (p->d.item_visits + 3) / 4 * 4 : p->d.item_visits;
    for (int d = 0; d < 3; ++d) p->blocks[d] = (desc.grid[d] + p->side - 1) / p->side;
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
    if (p->scheme == kSchemeMatrix) {
      p->block_chunk_off = alloc3<uint32_t>(p->nblocks + 1, &total);
      p->zero_cols = alloc3<uint2>(static_cast<uint64_t>(desc.grid[0]) * desc.grid[1], &total);
      p->zero_rows = alloc3<uint2>(desc.grid[0], &total);
      p->block_visited = alloc3<uint8_t>(p->nblocks, &total);
    }
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
                  p->counters,   p->sort_scratch, p->deapod[0],    p->deapod[1],  p->deapod[2],
                  p->visit_pos,  p->block_chunk_off, p->item_cost, p->item_scratch, p->run_first,
                  p->zero_cols,  p->block_visited, p->zero_rows};
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
  sa.side = static_cast<int>(p->side);
  sa.off = p->block_visit_off;
  const bool matrix = p->scheme == kSchemeMatrix;
  hipLaunchKernelGGL(visit3_count_kernel, dim3(ceil_div(p->nblocks + 1, 256)), dim3(256), 0, stream, sa);
  KSJ_LAUNCH_CHECK();
  exclusive_scan_u32(p->block_visit_off, p->block_visit_off, p->nblocks + 1, p->sort_scratch, stream);
  uint32_t visits = 0;
  if (matrix) {  // lists padded to groups of four
    hipLaunchKernelGGL(chunk3_count_kernel, dim3(ceil_div(p->nblocks + 1, 256)), dim3(256), 0, stream,
                       p->block_visit_off, p->nblocks, p->block_chunk_off);
    KSJ_LAUNCH_CHECK();
    exclusive_scan_u32(p->block_chunk_off, p->block_chunk_off, p->nblocks + 1, p->sort_scratch, stream);
    uint32_t chunks = 0;
    KSJ_HIP(hipMemcpyAsync(&chunks, p->block_chunk_off + p->nblocks, sizeof(uint32_t), hipMemcpyDeviceToHost, stream));
    KSJ_HIP(hipStreamSynchronize(stream));
    require(chunks <= 0x3FFFFFFFu, "too many gather visits for one 3-D plan");
    visits = 4 * chunks;
    hipLaunchKernelGGL(block_visited3_kernel, dim3(ceil_div(p->nblocks, 256)), dim3(256), 0, stream,
                       p->block_chunk_off, p->nblocks, p->block_visited);
    KSJ_LAUNCH_CHECK();
    const uint64_t grid_rows = static_cast<uint64_t>(p->d.grid[0]) * p->d.grid[1];
    hipLaunchKernelGGL(zero_span3_kernel, dim3(ceil_div(grid_rows, 256)), dim3(256), 0, stream, p->block_visited,
                       p->d.grid[0], p->d.grid[1], p->d.grid[2], p->blocks[1], p->blocks[2], p->side, p->zero_cols);
    KSJ_LAUNCH_CHECK();
    hipLaunchKernelGGL(zero_rows3_kernel, dim3(ceil_div(p->d.grid[0], 256)), dim3(256), 0, stream, p->zero_cols,
                       p->d.grid[0], p->d.grid[1], p->d.grid[2], p->zero_rows);
    KSJ_LAUNCH_CHECK();
  } else {
    KSJ_HIP(hipMemcpyAsync(&visits, p->block_visit_off + p->nblocks, sizeof(uint32_t), hipMemcpyDeviceToHost, stream));
    KSJ_HIP(hipStreamSynchronize(stream));
  }
  if (visits > p->visit_capacity || p->visit_k == nullptr) {
    for (void** q : {reinterpret_cast<void**>(&p->visit_k), reinterpret_cast<void**>(&p->visit_w),
                     reinterpret_cast<void**>(&p->visit_pos)}) {
      if (*q) KSJ_HIP(hipFree(*q));
      *q = nullptr;
    }
    p->visit_capacity = std::max<uint64_t>(visits, p->visit_capacity);
    KSJ_HIP(hipMalloc(&p->visit_k, std::max<size_t>(1, p->visit_capacity) * sizeof(uint32_t)));
    if (matrix)
      KSJ_HIP(hipMalloc(&p->visit_pos, std::max<size_t>(1, p->visit_capacity) * sizeof(uint32_t)));
    else
      KSJ_HIP(hipMalloc(&p->visit_w, std::max<size_t>(1, p->visit_capacity) * 2 * sizeof(float4)));
  }
  p->visit_count = visits;
  sa.visit_k = p->visit_k;
  sa.visit_w = p->visit_w;
  sa.visit_pos = p->visit_pos;
  sa.chunk_off = p->block_chunk_off;
  if (matrix) {
    hipLaunchKernelGGL(visit3_fill_matrix_kernel, dim3(ceil_div(p->nblocks, 256)), dim3(256), 0, stream, sa);
    KSJ_LAUNCH_CHECK();
    hipLaunchKernelGGL(pad3_visits_kernel, dim3(ceil_div(p->nblocks, 256)), dim3(256), 0, stream,
                       p->block_visit_off, p->block_chunk_off, p->nblocks, p->visit_k, p->visit_pos);
    KSJ_LAUNCH_CHECK();
  } else {
    hipLaunchKernelGGL(visit3_fill_kernel, dim3(ceil_div(p->nblocks, 256)), dim3(256), 0, stream, sa);
    KSJ_LAUNCH_CHECK();
  }

  // Items per block: the matrix schedule counts groups of four.
  const uint32_t* unit_off = matrix ? p->block_chunk_off : p->block_visit_off;
  const uint32_t unit_size = matrix ? p->item_visits / 4 : p->item_visits;
  KSJ_HIP(hipMemsetAsync(p->counters, 0, sizeof(uint32_t), stream));
  hipLaunchKernelGGL(item3_count_kernel, dim3(ceil_div(p->nblocks + 1, 256)), dim3(256), 0, stream, unit_off,
                     p->nblocks, unit_size, p->block_first_item, p->multi_blocks, p->counters);
  KSJ_LAUNCH_CHECK();
  exclusive_scan_u32(p->block_first_item, p->block_first_item, p->nblocks + 1, p->sort_scratch, stream);
  uint32_t items = 0, multi = 0;
  KSJ_HIP(hipMemcpyAsync(&items, p->block_first_item + p->nblocks, sizeof(uint32_t), hipMemcpyDeviceToHost, stream));
  KSJ_HIP(hipMemcpyAsync(&multi, p->counters, sizeof(uint32_t), hipMemcpyDeviceToHost, stream));
  KSJ_HIP(hipStreamSynchronize(stream));
  if (items > p->item_capacity || p->item_info == nullptr) {
    for (void** q : {reinterpret_cast<void**>(&p->item_info), reinterpret_cast<void**>(&p->item_cost),
                     &p->item_scratch}) {
      if (*q) KSJ_HIP(hipFree(*q));
      *q = nullptr;
    }
    p->item_capacity = std::max(items, p->item_capacity);
    KSJ_HIP(hipMalloc(&p->item_info, std::max<size_t>(1, p->item_capacity) * sizeof(uint4)));
    if (matrix) {
      KSJ_HIP(hipMalloc(&p->item_cost, (static_cast<size_t>(p->item_capacity) + 1) * sizeof(uint32_t)));
      KSJ_HIP(hipMalloc(&p->item_scratch, std::max<size_t>(256, scan_scratch_bytes(p->item_capacity + 1ull))));
    }
  }
  p->item_count = items;
  hipLaunchKernelGGL(slot3_count_kernel, dim3(ceil_div(p->nblocks + 1, 256)), dim3(256), 0, stream,
                     p->block_first_item, p->nblocks, p->block_first_slot);
  KSJ_LAUNCH_CHECK();
  exclusive_scan_u32(p->block_first_slot, p->block_first_slot, p->nblocks + 1, p->sort_scratch, stream);
  uint32_t slots = 0;
  KSJ_HIP(hipMemcpyAsync(&slots, p->block_first_slot + p->nblocks, sizeof(uint32_t), hipMemcpyDeviceToHost, stream));
  hipLaunchKernelGGL(item3_fill_kernel, dim3(ceil_div(p->nblocks, 256)), dim3(256), 0, stream, p->block_first_item,
                     p->block_first_slot, unit_off, unit_size, matrix ? 4u : 1u, p->nblocks, p->item_info);
  KSJ_LAUNCH_CHECK();
  if (matrix) {  // wave runs of equal cost over the item order
    hipLaunchKernelGGL(item3_cost_kernel, dim3(ceil_div(items + 1ull, 256)), dim3(256), 0, stream, p->item_info,
                       items, p->item_cost);
    KSJ_LAUNCH_CHECK();
    exclusive_scan_u32(p->item_cost, p->item_cost, items + 1ull, p->item_scratch, stream);
    uint32_t total = 0;
    KSJ_HIP(hipMemcpyAsync(&total, p->item_cost + items, sizeof(uint32_t), hipMemcpyDeviceToHost, stream));
    KSJ_HIP(hipStreamSynchronize(stream));
    const uint32_t runs = ceil_div(total, p->run_cost);
    if (runs > p->run_capacity || p->run_first == nullptr) {
      if (p->run_first) KSJ_HIP(hipFree(p->run_first));
      p->run_first = nullptr;
      p->run_capacity = std::max(runs, p->run_capacity);
      KSJ_HIP(hipMalloc(&p->run_first, (static_cast<size_t>(p->run_capacity) + 1) * sizeof(uint32_t)));
    }
    p->run_count = runs;
    hipLaunchKernelGGL(run3_split_kernel, dim3(ceil_div(runs + 1ull, 256)), dim3(256), 0, stream, p->item_cost, items,
                       runs, p->run_cost, p->run_first);
    KSJ_LAUNCH_CHECK();
  }
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
