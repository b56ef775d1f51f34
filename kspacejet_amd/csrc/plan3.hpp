// Trajectory plan for 3-D gridding (radial "kooshball" and any other 3-D
// trajectory): the 2-D plan's tables with a third axis, computed once on the
// device and kept resident.
//
//   coordinates  FP64 grid coordinate per axis, u = w * G / (2 pi) wrapped into
//                [0, G), its cell and fraction; bin = (cz * G_y + cy) * G_x + cx.
//   sort         stable LSD radix sort of samples by bin, cell_start[b].
//   weights      per-sample complex factor dcf * exp(+i w . s), s = N/2 - origin
//                per axis, FP64; dcf = 1, |w|^2 or given.
//   taps         per sorted sample W taps per axis (z, y, x), FP64 rounded to float.
//   schedule     per 2x2x2 cell block, the (sample, eight weights) visits of its
//                (W+1)^3 window bins; or (matrix schedule) per 4x4x4 cell block the
//                (sample, window position) visits of its (W+3)^3 window bins, padded
//                to groups of four, with the weights formed from the taps in the
//                gather; cut into work items of at most item_visits.
//   deapod       1 / phi_hat per output index of each axis, FP64.
// Volumes are [z][y][x]; trajectories [S][3] in (z, y, x) order, radians per pixel.
#pragma once

#include "plan.hpp"

namespace ksj::gpu {

enum : uint32_t { kDensityRampSquared = 3 };  // |w|^2, the 3-D radial density compensation

struct Plan3Desc {
  uint32_t image[3] = {0, 0, 0};  // N_z, N_y, N_x
  uint32_t grid[3] = {0, 0, 0};   // G_z, G_y, G_x
  uint32_t width = 4;
  double beta = 0.0;              // <= 0: Beatty value for the oversampling
  double origin[3] = {0.0, 0.0, 0.0};
  uint32_t density = kDensityRampSquared;  // kDensityNone, kDensityRampSquared or kDensityGiven
  uint32_t samples = 0;
  uint32_t item_visits = 64;
  uint32_t scheme = kSchemeDefault;  // kSchemeBlocks (2x2x2) or kSchemeMatrix (4x4x4); default: tuning key
  uint32_t run_cost = 0;             // matrix: cost per wave run, 1..64 (0: tuning key)
};

// Window position of a matrix-schedule visit: rz | ry << 4 | rx << 8; a padding visit
// has rz = 15, outside every window, so all its weights are zero.
constexpr uint32_t kPadVisitPos = 15;

struct Plan3 {
  Plan3Desc d{};
  int device = 0;
  uint64_t nbins = 0;
  int key_bits = 0;
  float2* weight = nullptr;        // [S] original order
  uint32_t* rank = nullptr;        // [S] original -> sorted position
  uint32_t* src = nullptr;         // [S] sorted -> original
  uint32_t* bins = nullptr;        // [S] bins (original order)
  uint32_t* cell_start = nullptr;  // [nbins + 1]
  float* taps = nullptr;           // [S][3W] sorted: z, y, x taps
  // Schedule over s x s x s blocks (block (bz, by, bx) covers cells s b .. s b + s - 1
  // per axis; s = 2, or 4 for the matrix schedule).
  uint32_t scheme = kSchemeBlocks, side = 2;
  uint32_t item_visits = 64;             // a multiple of four for the matrix schedule
  uint32_t blocks[3] = {0, 0, 0};
  uint64_t nblocks = 0;
  uint32_t* block_visit_off = nullptr;  // [nblocks + 1]
  uint32_t* visit_k = nullptr;          // [visit_capacity]
  float4* visit_w = nullptr;            // 2x2x2: [2 * visit_capacity], cells (cz, cy, cx) in cz*4 + cy*2 + cx order
  uint32_t* visit_pos = nullptr;        // matrix: [visit_capacity] window position (kPadVisitPos: padding)
  uint64_t visit_capacity = 0, visit_count = 0;
  uint32_t* block_first_item = nullptr;  // [nblocks + 1]
  uint32_t* block_first_slot = nullptr;  // [nblocks + 1]
  uint4* item_info = nullptr;            // {block, first visit, end visit, slot or ~0}
  uint32_t item_capacity = 0, item_count = 0, slot_count = 0;
  uint32_t* multi_blocks = nullptr;      // [nblocks]
  uint32_t multi_count = 0;
  // Matrix schedule: groups of four per block, item cost prefix (groups + 1 per
  // item) and the wave runs cut from it (at most 64 items each).
  uint32_t* block_chunk_off = nullptr;   // [nblocks + 1]
  uint32_t* item_cost = nullptr;         // [item_capacity + 1]
  void* item_scratch = nullptr;
  uint32_t run_cost = 8;
  uint32_t* run_first = nullptr;         // [run_capacity + 1]
  uint32_t run_capacity = 0, run_count = 0;
  float* deapod[3] = {nullptr, nullptr, nullptr};
  // Matrix schedule: per grid row (z, y), the x columns [a, b) no sample reaches -- the run
  // of visitless blocks around column G_x / 2 (every column for rows outside the sampled
  // ball); the gather does not write them and the x FFT reads them as zero (grid_prune).
  uint2* zero_cols = nullptr;        // [G_z * G_y] or null
  // Per z-plane: the run of rows y around G_y / 2 whose zero span is the whole row (rows
  // outside the sampled ball); the x pass does not store them and the y pass reads them
  // as zero.
  uint2* zero_rows = nullptr;        // [G_z] or null
  uint8_t* block_visited = nullptr;  // [nblocks] scratch of zero_cols
  // build scratch
  double* frac = nullptr;          // [S][3]
  uint32_t* sorted_keys = nullptr; // [S]
  uint32_t* counters = nullptr;    // [4]
  void* sort_scratch = nullptr;
  size_t device_bytes = 0;
  bool has_trajectory = false;
};

Plan3* plan3_create(const Plan3Desc& desc);
void plan3_destroy(Plan3* plan);
// Recomputes every table from a float32 [S][3] trajectory on the device (dcf read
// only for kDensityGiven); synchronises `stream`; throws on a non-finite coordinate
// or one outside [-pi, pi].
void plan3_set_trajectory(Plan3* plan, const float* d_traj, const float* d_dcf, hipStream_t stream);

// Golden-means kooshball (FP64 on the device, rounded to float): spoke p along
// (c, sqrt(1-c^2) sin a, sqrt(1-c^2) cos a), c = frac(p g1), a = 2 pi frac(p g2),
// g1 = 0.4656, g2 = 0.6823; k_j = (j - M/2)/M cycles per pixel; [S][3] (z, y, x).
void golden_means3(float* d_traj, uint32_t spokes, uint32_t readout, hipStream_t stream);

}  // namespace ksj::gpu
