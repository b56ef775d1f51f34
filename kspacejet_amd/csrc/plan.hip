// Trajectory plan kernels: FP64 coordinates, binning, taps, gather schedule, deapodization.
#include "kb.hpp"
#include "plan.hpp"
#include "sort.hpp"

#include <cmath>
#include <string>

namespace ksj::gpu {
namespace {

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

// ----------------------------------------------------------------------------- gather schedule
// Window of h x w block (BY, BX): bin rows h BY - W/2 + r and columns w BX - W/2 + c
// for r in [0, W + h - 1), c in [0, W + w - 1); cell (h BY + dy, w BX + dx) uses row
// tap dy - r + W - 1 and column tap dx - c + W - 1 of a sample in window bin (r, c)
// (taps outside [0, W) weigh 0).
struct ScheduleArgs {
  const uint32_t* start;
  const float* taps;  // [S][2W]
  uint32_t grid_rows, grid_cols, block_cols, block_group;
  uint64_t nblocks;
  int width, bh, bw;  // block h x w: 2 x 2, 4 x 4 or 4 x 8
  uint32_t* row_off;  // counts, then offsets
  const uint32_t* chunk_off;  // matrix schedule: first group of four of each block (null otherwise)
  uint32_t* visit_k;
  float* visit_w;     // [visits][h*w], or null for packed visits
  int packed;         // visit_k = k | r << 24 | c << 28 (window bin (r, c)); no weights
};

__device__ inline int window_rows(const ScheduleArgs& a) { return a.width + a.bh - 1; }
__device__ inline int window_cols(const ScheduleArgs& a) { return a.width + a.bw - 1; }

// False for the padding blocks past the grid (no visits).
__device__ inline bool block_in_grid(const ScheduleArgs& a, uint64_t blk) {
  uint32_t by, bx;
  block_coords(blk, a.block_cols, a.block_group, &by, &bx);
  return a.bh * by < a.grid_rows && a.bw * bx < a.grid_cols;
}

__device__ inline uint64_t window_bin(const ScheduleArgs& a, uint64_t blk, int r, int c) {
  const int W = a.width;
  const int gr = static_cast<int>(a.grid_rows), gc = static_cast<int>(a.grid_cols);
  uint32_t byu, bxu;
  block_coords(blk, a.block_cols, a.block_group, &byu, &bxu);
  int by = a.bh * static_cast<int>(byu) - W / 2 + r;
  int bx = a.bw * static_cast<int>(bxu) - W / 2 + c;
  by = ((by % gr) + gr) % gr;
  bx = ((bx % gc) + gc) % gc;
  return static_cast<uint64_t>(by) * a.grid_cols + bx;
}

__global__ void visit_count_kernel(ScheduleArgs a) {
  const uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  const int R = window_rows(a);
  const uint64_t rows = a.nblocks * R;
  if (i > rows) return;
  if (i == rows) {
    a.row_off[i] = 0;
    return;
  }
  const uint64_t blk = i / R;
  const int r = static_cast<int>(i % R);
  uint32_t n = 0;
  if (block_in_grid(a, blk)) {
    for (int c = 0; c < window_cols(a); ++c) {
      const uint64_t b = window_bin(a, blk, r, c);
      n += a.start[b + 1] - a.start[b];
    }
  }
  a.row_off[i] = n;
}

// Tap t of a sample's row (or column) taps, 0 outside [0, W).
__device__ inline float tap_or_zero(const float* t, int i, int W) { return (i >= 0 && i < W) ? t[i] : 0.f; }

__global__ void visit_fill_kernel(ScheduleArgs a) {
  const uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  const int W = a.width, h = a.bh, wd = a.bw, R = window_rows(a);
  if (i >= a.nblocks * R) return;
  const uint64_t blk = i / R;
  const int r = static_cast<int>(i % R);
  if (!block_in_grid(a, blk)) return;
  // Matrix schedule: blocks start on groups of four (lists padded by pad_visits_kernel).
  uint32_t pos = a.chunk_off ? 4 * a.chunk_off[blk] + (a.row_off[i] - a.row_off[blk * R]) : a.row_off[i];
  for (int c = 0; c < window_cols(a); ++c) {
    const uint64_t b = window_bin(a, blk, r, c);
    for (uint32_t k = a.start[b]; k < a.start[b + 1]; ++k) {
      if (a.packed) {
        a.visit_k[pos++] = k | static_cast<uint32_t>(r) << 24 | static_cast<uint32_t>(c) << 28;
        continue;
      }
      const float* t = a.taps + static_cast<uint64_t>(k) * (2 * W);
      float* w = a.visit_w + static_cast<uint64_t>(pos) * h * wd;
      for (int dy = 0; dy < h; ++dy) {
        const float wy = tap_or_zero(t, dy - r + W - 1, W);
        for (int dx = 0; dx < wd; ++dx) w[dy * wd + dx] = wy * tap_or_zero(t + W, dx - c + W - 1, W);
      }
      a.visit_k[pos] = k;
      ++pos;
    }
  }
}

// Matrix schedule: groups of four visits per block (chunks[nblocks] = 0 for the scan).
__global__ void chunk_count_kernel(const uint32_t* row_off, uint64_t nblocks, int rows, uint32_t* chunks) {
  const uint64_t b = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (b > nblocks) return;
  chunks[b] = b < nblocks ? (row_off[(b + 1) * rows] - row_off[b * rows] + 3) / 4 : 0u;
}

// Matrix schedule: the padding visits of each block repeat its first sample with zero
// weights (so they add exact zeros and read a row the block reads anyway); packed
// visits carry window bin (15, 15), which no cell of a block reaches.
__global__ void pad_visits_kernel(const uint32_t* row_off, const uint32_t* chunk_off, uint64_t nblocks, int rows,
                                  int cells, uint32_t* visit_k, float* visit_w) {
  const uint64_t b = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (b >= nblocks) return;
  const uint32_t base = 4 * chunk_off[b], end = 4 * chunk_off[b + 1];
  const uint32_t n = row_off[(b + 1) * rows] - row_off[b * rows];
  for (uint32_t p = base + n; p < end; ++p) {
    if (visit_w == nullptr) {
      visit_k[p] = (visit_k[base] & 0x00FFFFFFu) | 0xFF000000u;
      continue;
    }
    visit_k[p] = visit_k[base];
    for (int c = 0; c < cells; ++c) visit_w[static_cast<uint64_t>(p) * cells + c] = 0.f;
  }
}

// first[r] = first item whose cost prefix reaches r * run_cost; first[runs] = items.
__global__ void run_split_kernel(const uint32_t* prefix, uint32_t items, uint32_t runs, uint32_t run_cost,
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

// ----------------------------------------------------------------------------- tile schedule
// Tile t's window: bin rows h ty - W/2 + r and columns w tx - W/2 + c (wrapped),
// r in [0, h + W - 1), c in [0, w + W - 1).  Work items cut its window list into runs of item_samples.
struct TileSchedArgs {
  const uint32_t* start;
  uint32_t grid_rows, grid_cols, tile_cols, ntiles, item_samples;
  int width, tile_w, tile_h;
  uint32_t* tile_samples;
  uint32_t* first;    // items per tile, then (scanned) first item
  uint32_t* multi;    // tiles with more than one item
  uint32_t* counter;
};

__device__ inline uint32_t wrap_index(int v, int n) { return static_cast<uint32_t>(((v % n) + n) % n); }

__global__ void tile_count_kernel(TileSchedArgs a) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t > a.ntiles) return;
  if (t == a.ntiles) {
    a.first[t] = 0;
    return;
  }
  const int W = a.width, RY = a.tile_h + W - 1, RX = a.tile_w + W - 1;
  const int y0 = static_cast<int>(t / a.tile_cols) * a.tile_h - W / 2;
  const int x0 = static_cast<int>(t % a.tile_cols) * a.tile_w - W / 2;
  uint32_t n = 0;
  for (int r = 0; r < RY; ++r) {
    const uint64_t row = wrap_index(y0 + r, static_cast<int>(a.grid_rows));
    for (int c = 0; c < RX; ++c) {
      const uint64_t b = row * a.grid_cols + wrap_index(x0 + c, static_cast<int>(a.grid_cols));
      n += a.start[b + 1] - a.start[b];
    }
  }
  a.tile_samples[t] = n;
  const uint32_t items = n == 0 ? 1u : (n + a.item_samples - 1) / a.item_samples;
  a.first[t] = items;
  if (items > 1) a.multi[atomicAdd(a.counter, 1u)] = t;
}

__global__ void tile_item_fill_kernel(const uint32_t* first, const uint32_t* slot, const uint32_t* tile_samples,
                                      uint32_t ntiles, uint32_t item_samples, uint4* info) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ntiles) return;
  const uint32_t n = tile_samples[t];
  const uint32_t items = first[t + 1] - first[t];
  for (uint32_t j = 0; j < items; ++j) {
    const uint32_t b = j * item_samples;
    const uint32_t e = b + item_samples < n ? b + item_samples : n;
    info[first[t] + j] = make_uint4(t, b, e, items > 1 ? slot[t] + j : 0xFFFFFFFFu);
  }
}

// Block b's work is off[(b+1) stride] - off[b stride] units (visits, with stride =
// window rows per block; or groups of four, stride 1), cut into items of `size` units.
__global__ void item_count_kernel(const uint32_t* off, uint64_t nblocks, int stride, uint32_t size,
                                  uint32_t* first, uint32_t* multi, uint32_t* counter) {
  const uint64_t b = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (b > nblocks) return;
  if (b == nblocks) {
    first[b] = 0;
    return;
  }
  const uint32_t v = off[(b + 1) * stride] - off[b * stride];
  const uint32_t n = v == 0 ? 1u : (v + size - 1) / size;
  first[b] = n;
  if (n > 1) multi[atomicAdd(counter, 1u)] = static_cast<uint32_t>(b);
}

__global__ void slot_count_kernel(const uint32_t* first, uint64_t nblocks, uint32_t* slots) {
  const uint64_t b = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (b > nblocks) return;
  const uint32_t n = b < nblocks ? first[b + 1] - first[b] : 0u;
  slots[b] = n > 1 ? n : 0u;
}

// Cost of an item: its visits / unit plus a fixed cost for loading its header and
// writing its cells.  cost[items] = 0 so the exclusive scan's last entry is the total.
constexpr uint32_t kItemFixedCost = 16;

__global__ void item_cost_kernel(const uint4* info, uint32_t items, uint32_t unit, uint32_t fixed, uint32_t* cost) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < items) cost[i] = (info[i].z - info[i].y) / unit + fixed;
  if (i == items) cost[i] = 0;
}

// first[x] = smallest item whose cost prefix reaches x/8 of the total (first[0] = 0,
// first[8] = items): eight contiguous ranges of about equal cost.
__global__ void xcd_split_kernel(const uint32_t* prefix, uint32_t items, uint32_t* first) {
  const uint32_t x = threadIdx.x;
  if (x > kPlanXcds) return;
  if (x == 0 || x == kPlanXcds) {
    first[x] = x == 0 ? 0u : items;
    return;
  }
  const uint64_t target = static_cast<uint64_t>(prefix[items]) * x / kPlanXcds;
  uint32_t lo = 0, hi = items;
  while (lo < hi) {
    const uint32_t mid = lo + (hi - lo) / 2;
    if (prefix[mid] >= target)
      hi = mid;
    else
      lo = mid + 1;
  }
  first[x] = lo;
}

// Item records {block, first visit, end visit, slot or ~0}; units of `scale` visits.
__global__ void item_fill_kernel(const uint32_t* first, const uint32_t* slot, const uint32_t* off, int stride,
                                 uint32_t size, uint32_t scale, uint64_t nblocks, uint4* info) {
  const uint64_t b = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (b >= nblocks) return;
  const uint32_t vb = off[b * stride], ve = off[(b + 1) * stride];
  const uint32_t n = first[b + 1] - first[b];
  for (uint32_t j = 0; j < n; ++j) {
    const uint32_t v0 = vb + j * size;
    const uint32_t v1 = v0 + size < ve ? v0 + size : ve;
    info[first[b] + j] = make_uint4(static_cast<uint32_t>(b), scale * v0, scale * v1, n > 1 ? slot[b] + j : 0xFFFFFFFFu);
  }
}

// Block (by, bx) has visits: flag in row-major block order.
__global__ void block_visited_kernel(const uint32_t* chunk_off, uint64_t nblocks, uint32_t block_cols, uint32_t group,
                                     uint8_t* visited) {
  const uint64_t b = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (b >= nblocks) return;
  uint32_t by, bx;
  block_coords(b, block_cols, group, &by, &bx);
  visited[static_cast<uint64_t>(by) * block_cols + bx] = chunk_off[b + 1] > chunk_off[b] ? 1 : 0;
}

// Per grid row: the run of visitless blocks of its block row that holds column G_c / 2,
// as cell columns [x, y) (empty when that block has visits).
__global__ void zero_span_kernel(const uint8_t* visited, uint32_t grid_rows, uint32_t grid_cols, uint32_t block_cols,
                                 uint32_t bh, uint32_t bw, uint2* span) {
  const uint32_t y = blockIdx.x * blockDim.x + threadIdx.x;
  if (y >= grid_rows) return;
  const uint8_t* row = visited + static_cast<uint64_t>(y / bh) * block_cols;
  const uint32_t nb = (grid_cols + bw - 1) / bw;  // blocks holding grid columns
  const uint32_t mid = (grid_cols / 2) / bw;
  uint2 s = make_uint2(0, 0);
  if (mid < nb && row[mid] == 0) {
    uint32_t lo = mid, hi = mid + 1;
    while (lo > 0 && row[lo - 1] == 0) --lo;
    while (hi < nb && row[hi] == 0) ++hi;
    s = make_uint2(lo * bw, min(hi * bw, grid_cols));
  }
  span[y] = s;
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

// Route trajectory normalisation (noncartesian_rss_hdf5.cpp:776-834), float arithmetic
// as there: raw [S][kx, ky] -> [S][ky, kx] rad/px.  A failure records
// ((2 s + axis) << 8 | code) with atomicMin, so the host reports the first failing
// coordinate in the reference's order (kx before ky, sample by sample).
struct NormArgs {
  const float* raw;
  float* out;
  uint64_t samples;
  int units;
  float limit_col, limit_row;   // Nyquist bound per axis in input units
  float scale_col, scale_row;   // multiplier to rad/px (float, computed on the host)
  unsigned long long* first_error;
};

__device__ inline int normalize_one(float c, int units, float limit, float scale, float* out) {
  if (!isfinite(c)) return 1;
  if (units < 1 || units > 3) return 3;
  if (fabsf(c) > limit) return 2;
  *out = units == 2 ? c : c * scale;
  return 0;
}

__global__ void normalize_kernel(NormArgs a) {
  const uint64_t s = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (s >= a.samples) return;
  float col = 0.f, row = 0.f;
  int rc = normalize_one(a.raw[2 * s], a.units, a.limit_col, a.scale_col, &col);
  uint64_t axis = 0;
  if (rc == 0) {
    rc = normalize_one(a.raw[2 * s + 1], a.units, a.limit_row, a.scale_row, &row);
    axis = 1;
  }
  if (rc != 0) {
    atomicMin(a.first_error, static_cast<unsigned long long>(((2 * s + axis) << 8) | static_cast<uint64_t>(rc)));
    return;
  }
  a.out[2 * s] = row;
  a.out[2 * s + 1] = col;
}

__global__ void trajectory_flags_init_kernel(uint32_t* flags, uint32_t value) { *flags = value; }

// Bit 0: some value's bits differ from `resident`'s; bit 1: some value is not finite or
// lies outside [-bound, bound] (the comparison is false for NaN).  One OR per wave.
__global__ void trajectory_check_kernel(const float* traj, const float* resident, uint64_t values, float bound,
                                        uint32_t* flags) {
  uint32_t mask = 0;
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
  for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < values; i += stride) {
    const float v = traj[i];
    if (resident && __float_as_uint(v) != __float_as_uint(resident[i])) mask |= 1u;
    if (!(fabsf(v) <= bound)) mask |= 2u;
  }
  const uint32_t differs = __any(mask & 1u) ? 1u : 0u;
  const uint32_t invalid = __any(mask & 2u) ? 2u : 0u;
  if ((differs | invalid) != 0 && (threadIdx.x % warpSize) == 0) atomicOr(flags, differs | invalid);
}

__global__ void ramp_dcf_kernel(const float* traj, float* dcf, uint64_t samples) {
  const uint64_t s = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (s >= samples) return;
  dcf[s] = hypotf(traj[2 * s], traj[2 * s + 1]);
}


template <typename T> T* alloc(size_t count, size_t* total) {
  T* p = nullptr;
  const size_t bytes = count * sizeof(T);
  KSJ_HIP(hipMalloc(&p, bytes == 0 ? 1 : bytes));
  *total += bytes;
  return p;
}

// Grows the item arrays (info, cost prefix and its scan scratch) to hold `items`.
void reserve_items(Plan2* p, uint32_t items) {
  if (items <= p->item_capacity && p->item_cost != nullptr) return;
  for (void** q : {reinterpret_cast<void**>(&p->item_info), reinterpret_cast<void**>(&p->item_cost),
                   &p->item_scratch}) {
    if (*q) KSJ_HIP(hipFree(*q));
    *q = nullptr;
  }
  p->item_capacity = std::max(items, p->item_capacity);
  KSJ_HIP(hipMalloc(&p->item_info, std::max<size_t>(1, p->item_capacity) * sizeof(uint4)));
  KSJ_HIP(hipMalloc(&p->item_cost, (static_cast<size_t>(p->item_capacity) + 1) * sizeof(uint32_t)));
  KSJ_HIP(hipMalloc(&p->item_scratch, std::max<size_t>(256, scan_scratch_bytes(p->item_capacity + 1ull))));
}

// Cost prefix of the items and the eight equal-cost XCD ranges.  The matrix
// schedule counts groups of four visits plus one per item.
void xcd_ranges(Plan2* p, hipStream_t stream) {
  const uint32_t items = p->item_count;
  const bool matrix = p->scheme == kSchemeMatrix;
  hipLaunchKernelGGL(item_cost_kernel, dim3(ceil_div(items + 1ull, 256)), dim3(256), 0, stream, p->item_info, items,
                     matrix ? 4u : 1u, matrix ? 1u : kItemFixedCost, p->item_cost);
  KSJ_LAUNCH_CHECK();
  exclusive_scan_u32(p->item_cost, p->item_cost, items + 1ull, p->item_scratch, stream);
  hipLaunchKernelGGL(xcd_split_kernel, dim3(1), dim3(64), 0, stream, p->item_cost, items, p->xcd_first);
  KSJ_LAUNCH_CHECK();
}

// Tile schedule: window-list length per tile -> items per tile -> item records.
void tile_schedule(Plan2* p, hipStream_t stream) {
  TileSchedArgs ta{};
  ta.start = p->cell_start;
  ta.grid_rows = p->d.grid_rows;
  ta.grid_cols = p->d.grid_cols;
  ta.tile_cols = p->tile_cols;
  ta.ntiles = p->ntiles;
  ta.item_samples = p->item_samples;
  ta.width = static_cast<int>(p->d.width);
  ta.tile_w = static_cast<int>(p->tile_width);
  ta.tile_h = static_cast<int>(p->tile_height);
  ta.tile_samples = p->tile_samples;
  ta.first = p->block_first_item;
  ta.multi = p->multi_blocks;
  ta.counter = p->counters;
  KSJ_HIP(hipMemsetAsync(p->counters, 0, sizeof(uint32_t), stream));
  hipLaunchKernelGGL(tile_count_kernel, dim3(ceil_div(p->ntiles + 1, 256)), dim3(256), 0, stream, ta);
  KSJ_LAUNCH_CHECK();
  exclusive_scan_u32(p->block_first_item, p->block_first_item, p->ntiles + 1ull, p->sort_scratch, stream);
  uint32_t items = 0, multi = 0;
  KSJ_HIP(hipMemcpyAsync(&items, p->block_first_item + p->ntiles, sizeof(uint32_t), hipMemcpyDeviceToHost, stream));
  KSJ_HIP(hipMemcpyAsync(&multi, p->counters, sizeof(uint32_t), hipMemcpyDeviceToHost, stream));
  KSJ_HIP(hipStreamSynchronize(stream));
  reserve_items(p, items);
  p->item_count = items;
  hipLaunchKernelGGL(slot_count_kernel, dim3(ceil_div(p->ntiles + 1, 256)), dim3(256), 0, stream,
                     p->block_first_item, static_cast<uint64_t>(p->ntiles), p->block_first_slot);
  KSJ_LAUNCH_CHECK();
  exclusive_scan_u32(p->block_first_slot, p->block_first_slot, p->ntiles + 1ull, p->sort_scratch, stream);
  uint32_t slots = 0;
  KSJ_HIP(hipMemcpyAsync(&slots, p->block_first_slot + p->ntiles, sizeof(uint32_t), hipMemcpyDeviceToHost, stream));
  hipLaunchKernelGGL(tile_item_fill_kernel, dim3(ceil_div(p->ntiles, 256)), dim3(256), 0, stream,
                     p->block_first_item, p->block_first_slot, p->tile_samples, p->ntiles, p->item_samples,
                     p->item_info);
  KSJ_LAUNCH_CHECK();
  xcd_ranges(p, stream);
  KSJ_HIP(hipStreamSynchronize(stream));
  p->slot_count = slots;
  p->multi_count = multi;
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
  require(desc.item_visits <= 4096, "item visits must be in [1, 4096]");
  auto* p = new Plan2();
  p->d = desc;
  if (p->d.scheme == kSchemeDefault) p->d.scheme = static_cast<uint32_t>(tune_get(kTuneGatherScheme));
  if (p->d.item_visits == 0)
    p->d.item_visits =
      static_cast<uint32_t>(tune_get(p->d.scheme == kSchemeMatrix ? kTuneMatrixItemVisits : kTuneItemVisits));
  if (p->d.item_samples == 0) p->d.item_samples = static_cast<uint32_t>(tune_get(kTuneTileItemSamples));
  if (p->d.tile_width == 0) p->d.tile_width = static_cast<uint32_t>(tune_get(kTuneTileWidth));
  if (p->d.block_group == 0) p->d.block_group = static_cast<uint32_t>(tune_get(kTuneBlockGroup));
  if (p->d.run_cost == 0) p->d.run_cost = static_cast<uint32_t>(tune_get(kTuneMatrixRunCost));
  if (p->d.block_width == 0) p->d.block_width = static_cast<uint32_t>(tune_get(kTuneMatrixBlockWidth));
  require(p->d.block_width == 4 || p->d.block_width == 8, "matrix block width must be 4 or 8");
  require(p->d.run_cost >= 1 && p->d.run_cost <= 64, "run cost must be in [1, 64]");
  require(p->d.block_group == 1 || p->d.block_group == 2 || p->d.block_group == 4, "block group must be 1, 2 or 4");
  require(p->d.tile_width == 4 || p->d.tile_width == 8, "tile width must be 4 or 8");
  require(p->d.scheme == kSchemeBlocks || p->d.scheme == kSchemeTiles || p->d.scheme == kSchemeMatrix,
          "unknown gather schedule");
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
    p->block_group = p->d.block_group;
    const bool matrix = p->d.scheme == kSchemeMatrix;
    p->block_h = matrix ? 4u : 2u;
    p->block_w = matrix ? p->d.block_width : 2u;
    p->block_rows = ceil_div(ceil_div(desc.grid_rows, p->block_h), p->block_group) * p->block_group;
    p->block_cols = ceil_div(ceil_div(desc.grid_cols, p->block_w), p->block_group) * p->block_group;
    p->nblocks = static_cast<uint64_t>(p->block_rows) * p->block_cols;
    // The matrix gather takes visits four at a time: items start on multiples of 4.
    p->item_visits = p->d.scheme == kSchemeMatrix ? (p->d.item_visits + 3) / 4 * 4 : p->d.item_visits;
    p->scheme = p->d.scheme;
    // Packed visits need the sorted position in 24 bits and window bins below 15
    // (W + 8 - 1 <= 15 for every legal width and block).
    p->packed_visits = matrix && tune_get(kTuneMatrixWeights) == 1 && S < (1ull << 24);
    p->item_samples = p->d.item_samples;
    p->run_cost = p->d.run_cost;
    p->tile_width = p->d.tile_width;
    p->tile_height = kTileCells / p->tile_width;
    p->tile_rows = ceil_div(desc.grid_rows, p->tile_height);
    p->tile_cols = ceil_div(desc.grid_cols, p->tile_width);
    p->ntiles = p->tile_rows * p->tile_cols;
    size_t total = 0;
    p->weight = alloc<float2>(S, &total);
    p->rank = alloc<uint32_t>(S, &total);
    p->src = alloc<uint32_t>(S, &total);
    p->bins = alloc<uint32_t>(S, &total);
    p->cell_start = alloc<uint32_t>(p->nbins + 1, &total);
    p->taps = alloc<float>(S * 2 * desc.width, &total);
    const uint64_t window = desc.width + p->block_h - 1;  // window rows per block
    if (p->scheme != kSchemeTiles)
      p->visit_row_off = alloc<uint32_t>(p->nblocks * window + 1, &total);
    else
      p->tile_samples = alloc<uint32_t>(p->ntiles, &total);
    p->block_first_item = alloc<uint32_t>(p->nblocks + 1, &total);
    p->block_first_slot = alloc<uint32_t>(p->nblocks + 1, &total);
    if (p->scheme == kSchemeMatrix) {
      p->block_chunk_off = alloc<uint32_t>(p->nblocks + 1, &total);
      p->zero_cols = alloc<uint2>(desc.grid_rows, &total);
      p->block_visited = alloc<uint8_t>(p->nblocks, &total);
    }
    p->multi_blocks = alloc<uint32_t>(p->nblocks, &total);
    p->xcd_first = alloc<uint32_t>(kPlanXcds + 1, &total);
    p->frac = alloc<double>(2 * S, &total);
    p->sorted_keys = alloc<uint32_t>(S, &total);
    p->counters = alloc<uint32_t>(4, &total);
    const uint64_t scan_n = std::max<uint64_t>(S, p->nblocks * window + 1);
    p->sort_scratch = alloc<char>(std::max(radix_sort_scratch_bytes(scan_n), scan_scratch_bytes(scan_n)), &total);
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
  void* ptrs[] = {p->weight,        p->rank,        p->src,          p->bins,         p->cell_start,
                  p->taps,          p->visit_row_off, p->visit_k,    p->visit_w,      p->block_first_item,
                  p->block_first_slot, p->item_info,    p->multi_blocks, p->frac,         p->sorted_keys,
                  p->counters,      p->sort_scratch, p->deapod_rows, p->deapod_cols, p->item_cost,
                  p->item_scratch,  p->xcd_first,   p->tile_samples, p->block_chunk_off, p->run_first,
                  p->zero_cols,     p->block_visited};
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
  uint32_t host_counters[4] = {0, 0, 0, 0};
  KSJ_HIP(hipMemcpyAsync(host_counters, p->counters, sizeof(host_counters), hipMemcpyDeviceToHost, stream));
  KSJ_HIP(hipStreamSynchronize(stream));
  if (host_counters[1] & kErrBadCoordinate)
    throw Error(1, "trajectory must be finite and within [-pi, pi] radians per pixel");

  if (p->scheme == kSchemeTiles) {
    tile_schedule(p, stream);
    p->has_trajectory = true;
    return;
  }
  // Gather schedule: visits per (block, window row) -> offsets -> visit records.
  ScheduleArgs sa{};
  sa.start = p->cell_start;
  sa.taps = p->taps;
  sa.grid_rows = p->d.grid_rows;
  sa.grid_cols = p->d.grid_cols;
  sa.block_cols = p->block_cols;
  sa.block_group = p->block_group;
  sa.nblocks = p->nblocks;
  sa.width = static_cast<int>(p->d.width);
  sa.bh = static_cast<int>(p->block_h);
  sa.bw = static_cast<int>(p->block_w);
  sa.row_off = p->visit_row_off;
  const int window = sa.width + sa.bh - 1;
  const uint64_t nrows = p->nblocks * window;
  const uint64_t weight_vec4 = static_cast<uint64_t>(sa.bh) * sa.bw / 4;  // float4s of weights per visit
  hipLaunchKernelGGL(visit_count_kernel, dim3(ceil_div(nrows + 1, 256)), dim3(256), 0, stream, sa);
  KSJ_LAUNCH_CHECK();
  exclusive_scan_u32(p->visit_row_off, p->visit_row_off, nrows + 1, p->sort_scratch, stream);
  const bool matrix = p->scheme == kSchemeMatrix;
  uint32_t visits = 0;
  if (matrix) {  // lists padded to groups of four: visits = 4 * groups
    hipLaunchKernelGGL(chunk_count_kernel, dim3(ceil_div(p->nblocks + 1, 256)), dim3(256), 0, stream,
                       p->visit_row_off, p->nblocks, window, p->block_chunk_off);
    KSJ_LAUNCH_CHECK();
    exclusive_scan_u32(p->block_chunk_off, p->block_chunk_off, p->nblocks + 1, p->sort_scratch, stream);
    uint32_t chunks = 0;
    KSJ_HIP(hipMemcpyAsync(&chunks, p->block_chunk_off + p->nblocks, sizeof(uint32_t), hipMemcpyDeviceToHost, stream));
    KSJ_HIP(hipStreamSynchronize(stream));
    require(chunks <= 0x3FFFFFFFu, "too many gather visits for one plan");
    visits = 4 * chunks;
    hipLaunchKernelGGL(block_visited_kernel, dim3(ceil_div(p->nblocks, 256)), dim3(256), 0, stream,
                       p->block_chunk_off, p->nblocks, p->block_cols, p->block_group, p->block_visited);
    KSJ_LAUNCH_CHECK();
    hipLaunchKernelGGL(zero_span_kernel, dim3(ceil_div(p->d.grid_rows, 256)), dim3(256), 0, stream,
                       p->block_visited, p->d.grid_rows, p->d.grid_cols, p->block_cols, p->block_h, p->block_w,
                       p->zero_cols);
    KSJ_LAUNCH_CHECK();
  } else {
    KSJ_HIP(hipMemcpyAsync(&visits, p->visit_row_off + nrows, sizeof(uint32_t), hipMemcpyDeviceToHost, stream));
    KSJ_HIP(hipStreamSynchronize(stream));
  }
  if (visits > p->visit_capacity) {  // first trajectory on this plan, or a denser one
    if (p->visit_k) KSJ_HIP(hipFree(p->visit_k));
    if (p->visit_w) KSJ_HIP(hipFree(p->visit_w));
    p->visit_k = nullptr;
    p->visit_w = nullptr;
    p->visit_capacity = visits;
    KSJ_HIP(hipMalloc(&p->visit_k, std::max<size_t>(1, visits) * sizeof(uint32_t)));
    if (!p->packed_visits)
      KSJ_HIP(hipMalloc(&p->visit_w, std::max<size_t>(1, visits) * weight_vec4 * sizeof(float4)));
  }
  p->visit_count = visits;
  sa.visit_k = p->visit_k;
  sa.visit_w = reinterpret_cast<float*>(p->visit_w);
  sa.packed = p->packed_visits ? 1 : 0;
  sa.chunk_off = matrix ? p->block_chunk_off : nullptr;
  hipLaunchKernelGGL(visit_fill_kernel, dim3(ceil_div(nrows, 256)), dim3(256), 0, stream, sa);
  KSJ_LAUNCH_CHECK();
  if (matrix) {
    hipLaunchKernelGGL(pad_visits_kernel, dim3(ceil_div(p->nblocks, 256)), dim3(256), 0, stream, p->visit_row_off,
                       p->block_chunk_off, p->nblocks, window, sa.bh * sa.bw, p->visit_k, sa.visit_w);
    KSJ_LAUNCH_CHECK();
  }
  // Work items: ceil(visits / item_visits) per block (at least one, so every cell is
  // written); the matrix schedule counts in groups of four.
  const uint32_t* unit_off = matrix ? p->block_chunk_off : p->visit_row_off;
  const int unit_stride = matrix ? 1 : window;
  const uint32_t unit_size = matrix ? p->item_visits / 4 : p->item_visits;
  KSJ_HIP(hipMemsetAsync(p->counters, 0, sizeof(uint32_t), stream));
  hipLaunchKernelGGL(item_count_kernel, dim3(ceil_div(p->nblocks + 1, 256)), dim3(256), 0, stream, unit_off,
                     p->nblocks, unit_stride, unit_size, p->block_first_item, p->multi_blocks, p->counters);
  KSJ_LAUNCH_CHECK();
  exclusive_scan_u32(p->block_first_item, p->block_first_item, p->nblocks + 1, p->sort_scratch, stream);
  uint32_t items = 0, multi = 0;
  KSJ_HIP(hipMemcpyAsync(&items, p->block_first_item + p->nblocks, sizeof(uint32_t), hipMemcpyDeviceToHost, stream));
  KSJ_HIP(hipMemcpyAsync(&multi, p->counters, sizeof(uint32_t), hipMemcpyDeviceToHost, stream));
  KSJ_HIP(hipStreamSynchronize(stream));
  reserve_items(p, items);
  p->item_count = items;
  // Partial-sum slots: consecutive for the items of each multi-item block.
  hipLaunchKernelGGL(slot_count_kernel, dim3(ceil_div(p->nblocks + 1, 256)), dim3(256), 0, stream,
                     p->block_first_item, p->nblocks, p->block_first_slot);
  KSJ_LAUNCH_CHECK();
  exclusive_scan_u32(p->block_first_slot, p->block_first_slot, p->nblocks + 1, p->sort_scratch, stream);
  uint32_t slots = 0;
  KSJ_HIP(hipMemcpyAsync(&slots, p->block_first_slot + p->nblocks, sizeof(uint32_t), hipMemcpyDeviceToHost, stream));
  hipLaunchKernelGGL(item_fill_kernel, dim3(ceil_div(p->nblocks, 256)), dim3(256), 0, stream, p->block_first_item,
                     p->block_first_slot, unit_off, unit_stride, unit_size, matrix ? 4u : 1u, p->nblocks,
                     p->item_info);
  KSJ_LAUNCH_CHECK();
  xcd_ranges(p, stream);
  if (matrix) {  // wave runs over the cost prefix
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
    hipLaunchKernelGGL(run_split_kernel, dim3(ceil_div(runs + 1ull, 256)), dim3(256), 0, stream, p->item_cost, items,
                       runs, p->run_cost, p->run_first);
    KSJ_LAUNCH_CHECK();
  }
  KSJ_HIP(hipStreamSynchronize(stream));
  p->slot_count = slots;
  // The multi-item list is in append order; each of its blocks is summed on its
  // own, so the order changes which thread sums a block, never the sum.
  p->multi_count = multi;
  KSJ_HIP(hipStreamSynchronize(stream));
  p->has_trajectory = true;
}

void deapod_table(float* d_table, uint32_t n, uint32_t grid, int width, double beta, hipStream_t stream) {
  hipLaunchKernelGGL(deapod_kernel, dim3(ceil_div(n, 256)), dim3(256), 0, stream, d_table, n, grid, width, beta,
                     host_i0(beta));
  KSJ_LAUNCH_CHECK();
}

void cell_starts(const uint32_t* d_sorted_keys, uint32_t samples, uint64_t nbins, uint32_t* d_start,
                 hipStream_t stream) {
  hipLaunchKernelGGL(cell_start_kernel, dim3(ceil_div(static_cast<uint64_t>(samples) + 1, 256)), dim3(256), 0, stream,
                     d_sorted_keys, samples, nbins, d_start);
  KSJ_LAUNCH_CHECK();
}

void golden_angle2(float* d_traj, uint32_t spokes, uint32_t readout, hipStream_t stream) {
  const uint64_t n = static_cast<uint64_t>(spokes) * readout;
  if (n == 0) return;
  hipLaunchKernelGGL(golden_angle_kernel, dim3(ceil_div(n, 256)), dim3(256), 0, stream, d_traj, spokes, readout);
  KSJ_LAUNCH_CHECK();
}

void trajectory_check(const float* d_traj, const float* d_resident, uint64_t values, float bound,
                      uint32_t* d_flags, hipStream_t stream) {
  hipLaunchKernelGGL(trajectory_flags_init_kernel, dim3(1), dim3(1), 0, stream, d_flags, d_resident ? 0u : 1u);
  KSJ_LAUNCH_CHECK();
  if (values == 0) return;
  const unsigned blocks = static_cast<unsigned>(std::min<uint64_t>(ceil_div(values, 256), 1024));
  hipLaunchKernelGGL(trajectory_check_kernel, dim3(blocks), dim3(256), 0, stream, d_traj, d_resident, values, bound,
                     d_flags);
  KSJ_LAUNCH_CHECK();
}

void normalize_trajectory2(const float* d_raw, uint64_t samples, int units, uint32_t encoded_rows,
                           uint32_t encoded_cols, float* d_out, hipStream_t stream) {
  require(units >= 1 && units <= 3, "trajectory units are unspecified");
  require(units != 3 || (encoded_rows >= 2 && encoded_cols >= 2),
          "encoded_matrix_index requires encoded dimensions >= 2");
  if (samples == 0) return;
  const float pi_f = 3.14159265358979323846f;
  NormArgs a{};
  a.raw = d_raw;
  a.out = d_out;
  a.samples = samples;
  a.units = units;
  if (units == 1) {
    a.limit_col = a.limit_row = 0.5f;
    a.scale_col = a.scale_row = 2.0f * pi_f;
  } else if (units == 2) {
    a.limit_col = a.limit_row = pi_f;
    a.scale_col = a.scale_row = 1.0f;
  } else {
    a.limit_col = static_cast<float>(encoded_cols) / 2.0f;
    a.limit_row = static_cast<float>(encoded_rows) / 2.0f;
    a.scale_col = 2.0f * pi_f / static_cast<float>(encoded_cols);
    a.scale_row = 2.0f * pi_f / static_cast<float>(encoded_rows);
  }
  unsigned long long* d_err = nullptr;
  KSJ_HIP(hipMalloc(reinterpret_cast<void**>(&d_err), sizeof(unsigned long long)));
  a.first_error = d_err;
  unsigned long long host_err = ~0ull;
  try {
    KSJ_HIP(hipMemcpyAsync(d_err, &host_err, sizeof(host_err), hipMemcpyHostToDevice, stream));
    hipLaunchKernelGGL(normalize_kernel, dim3(ceil_div(samples, 256)), dim3(256), 0, stream, a);
    KSJ_LAUNCH_CHECK();
    KSJ_HIP(hipMemcpyAsync(&host_err, d_err, sizeof(host_err), hipMemcpyDeviceToHost, stream));
    KSJ_HIP(hipStreamSynchronize(stream));
  } catch (...) {
    (void)hipFree(d_err);
    throw;
  }
  KSJ_HIP(hipFree(d_err));
  if (host_err != ~0ull) {
    switch (host_err & 0xFF) {
      case 1: throw Error(1, "trajectory coordinate must be finite");
      case 2: throw Error(1, "trajectory coordinate is outside the inclusive Nyquist interval of its units");
      default: throw Error(1, "trajectory units are unspecified");
    }
  }
}

void ramp_dcf2(const float* d_traj, float* d_dcf, uint64_t samples, hipStream_t stream) {
  if (samples == 0) return;
  hipLaunchKernelGGL(ramp_dcf_kernel, dim3(ceil_div(samples, 256)), dim3(256), 0, stream, d_traj, d_dcf, samples);
  KSJ_LAUNCH_CHECK();
}

}  // namespace ksj::gpu
