// Trajectory plan for 2-D gridding: everything derived from a trajectory alone,
// computed once on the device and kept resident so later scans on the same
// trajectory reuse it.
//
//   coordinates  FP64 grid coordinate per sample, u = w * G / (2 pi) wrapped into
//                [0, G) (the rule of eigen_radial_gridding.cpp:262-277), its cell
//                and fraction; bin = cell_row * G_cols + cell_col.
//   sort         stable LSD radix sort of samples by bin (ties keep sample order),
//                cell_start[b] = first sorted position of bin b.
//   weights      per-sample complex factor dcf * exp(+i(w_r s_r + w_c s_c)) with
//                s = (centred ? N/2 : 0) - origin, evaluated in FP64.
//   taps         per sorted sample W row taps and W column taps of the kernel
//                (Kaiser-Bessel or linear), evaluated in FP64, stored as float.
//   schedule     either (blocks, matrix) per 2x2 or 4x4 cell block the list of
//                (sample, cell weights) visits, or (tiles) per 64-cell tile the
//                length of its window sample list; all cut into bounded work items
//                so dense k-space centres are split over many workgroups.
//   deapod       1 / phi_hat per output row / column (Kaiser-Bessel only), FP64.
#pragma once

#include "common.hpp"

namespace ksj::gpu {

enum : uint32_t { kKernelLinear = 0, kKernelKaiserBessel = 1 };
// XCD count of the MI355X: the gather can give each XCD a contiguous, cost-balanced item range.
constexpr uint32_t kPlanXcds = 8;
enum : uint32_t { kDensityNone = 0, kDensityRamp = 1, kDensityGiven = 2 };
// Gather schedules: 2x2 blocks with per-visit weights read from HBM, 64-cell
// tiles (64/w rows x w columns) whose window samples are staged in LDS, or 4x4
// blocks whose visits carry all sixteen cell weights and are summed on the matrix
// cores.  kSchemeDefault takes tuning key gather_scheme.
enum : uint32_t { kSchemeBlocks = 0, kSchemeTiles = 1, kSchemeMatrix = 2, kSchemeDefault = 0xFFFFFFFFu };
constexpr uint32_t kTileCells = 64;


// Block numbering: blocks come in groups of g x g (row-major inside a group,
// groups row-major), so consecutive work items -- which one workgroup takes
// together -- cover a square of neighbouring blocks whose windows share sample
// rows in the CU's L1.  block_rows / block_cols are padded to whole groups; the
// blocks past the grid get no visits and write nothing.  g = 1 is plain row-major.
__host__ __device__ inline void block_coords(uint64_t b, uint32_t block_cols, uint32_t g, uint32_t* by,
                                             uint32_t* bx) {
  const uint64_t q = b / (g * g);
  const uint32_t s = static_cast<uint32_t>(b % (g * g));
  const uint32_t groups_per_row = block_cols / g;
  *by = g * static_cast<uint32_t>(q / groups_per_row) + s / g;
  *bx = g * static_cast<uint32_t>(q % groups_per_row) + s % g;
}

struct Plan2Desc {
  uint32_t image_rows = 0, image_cols = 0;  // N
  uint32_t grid_rows = 0, grid_cols = 0;    // G
  uint32_t width = 4;                       // taps per axis (2 for linear)
  uint32_t kernel = kKernelKaiserBessel;
  double beta = 0.0;                        // <= 0: Beatty value for the oversampling
  double row_origin = 0.0, col_origin = 0.0;
  uint32_t density = kDensityRamp;
  uint32_t centred = 1;                     // 1: image pixel r reads bin (r - N/2) mod G; 0: bin r
  uint32_t samples = 0;
  uint32_t item_visits = 64;                // blocks, matrix: visits per gather work item (0: tuning key)
  uint32_t scheme = kSchemeDefault;
  uint32_t item_samples = 0;                // tiles: window samples per work item (0: tuning key)
  uint32_t tile_width = 0;                  // tiles: 4 or 8 columns (0: tuning key)
  uint32_t block_group = 0;                 // blocks: group side 1, 2 or 4 (0: tuning key)
  uint32_t run_cost = 0;                    // matrix: cost per wave run, 1..64 (0: tuning key)
  uint32_t block_width = 0;                 // matrix: block columns, 4 or 8 (0: tuning key)
};

struct Plan2 {
  Plan2Desc d{};
  int device = 0;
  uint64_t nbins = 0;
  int key_bits = 0;
  // resident tables
  float2* weight = nullptr;        // [S] original order
  uint32_t* rank = nullptr;        // [S] original -> sorted position
  uint32_t* src = nullptr;         // [S] sorted position -> original
  uint32_t* bins = nullptr;        // [S] sample bins (original order), then sorted keys
  uint32_t* cell_start = nullptr;  // [nbins + 1]
  float* taps = nullptr;           // [S][2W] sorted: W row taps then W column taps
  // Gather schedule.  Cells are taken in h x w blocks (2 x 2; or 4 x 4 or 4 x 8 for
  // the matrix schedule); a block's visits are the samples of its (W+h-1) x (W+w-1)
  // contributing bins (rows then columns of that bin window, sorted order inside a
  // bin), each with the h*w weights phi(row tap) * phi(column tap) of the block's
  // cells in row-major order (0 outside the footprint).  Blocks are cut into work
  // items of at most item_visits visits (a multiple of 4 for the matrix schedule);
  // a block with several items is summed item by item in item order.
  uint32_t block_h = 2, block_w = 2;     // cells per block row / column
  uint32_t block_rows = 0, block_cols = 0, block_group = 1;  // rows / columns padded to whole groups
  uint64_t nblocks = 0;
  uint32_t item_visits = 64;
  uint32_t* visit_row_off = nullptr;  // [nblocks * (W+h-1) + 1] first visit of each (block, window row)
  uint32_t* visit_k = nullptr;        // [visit_capacity] sorted sample position (packed_visits: position in
                                      // bits 0-23, window row in bits 24-27, window column in bits 28-31)
  float4* visit_w = nullptr;          // [visit_capacity * h*w/4] cell weights, row-major per visit (null when
                                      // packed_visits: the gather forms them from the taps)
  bool packed_visits = false;         // matrix schedule with weights formed from the taps
  uint64_t visit_capacity = 0;
  uint64_t visit_count = 0;           // host copy
  uint32_t* block_first_item = nullptr;  // [nblocks + 1]
  uint32_t* block_first_slot = nullptr;  // [nblocks + 1] first partial slot (multi-item blocks only)
  uint4* item_info = nullptr;            // [item_capacity] {block, first visit, end visit, slot or ~0}
  uint32_t item_capacity = 0;
  uint32_t item_count = 0;               // host copy
  uint32_t slot_count = 0;               // partial slots (items of multi-item blocks), host copy
  uint32_t* item_cost = nullptr;         // [item_capacity + 1] prefix sums of item costs (visits + fixed)
  void* item_scratch = nullptr;          // scan scratch for item_cost
  uint32_t* xcd_first = nullptr;         // [kPlanXcds + 1] first item of each XCD's contiguous range
  uint32_t* multi_blocks = nullptr;      // [nblocks] blocks with more than one item
  uint32_t multi_count = 0;              // host copy
  // Matrix schedule (scheme == kSchemeMatrix).  Every block's visit list is padded
  // to a multiple of four (padding visits repeat the block's first sample with zero
  // weights), so items start on groups of four and the items of consecutive blocks
  // are contiguous.  Items cost (visits / 4 + 1); wave run r takes the items whose
  // cost prefix lies in [r run_cost, (r+1) run_cost), at most 64 of them.
  uint32_t* block_chunk_off = nullptr;   // [nblocks + 1] first group of four of each block
  uint32_t run_cost = 32;
  uint32_t* run_first = nullptr;         // [run_capacity + 1] first item of each run
  uint32_t run_capacity = 0;
  uint32_t run_count = 0;                // host copy
  // Tile schedule (scheme == kSchemeTiles).  Tile (ty, tx) covers cells
  // [h ty, h ty + h) x [w tx, w tx + w) with w = tile_width, h = 64 / w; its window
  // is the (h+W-1) x (w+W-1) bins whose samples reach those cells, rows then
  // columns, sorted order inside a bin.  Work item = {tile, first, end
  // window-list index, slot or ~0}, at most item_samples long; block_first_item /
  // block_first_slot / multi_blocks are indexed by tile.
  uint32_t scheme = kSchemeBlocks;
  uint32_t tile_width = 8, tile_height = 8;
  uint32_t tile_rows = 0, tile_cols = 0, ntiles = 0, item_samples = 0;
  uint32_t* tile_samples = nullptr;      // [ntiles] window-list length
  float* deapod_rows = nullptr;    // [N_r] or null
  float* deapod_cols = nullptr;    // [N_c] or null
  // Matrix schedule: per grid row, the columns [x, y) that no sample reaches -- the run of
  // blocks without visits around column G_c / 2 (the corners of k-space outside the
  // sampled disk, in the grid's uncentred layout).  The gather does not write those
  // cells and the row FFT reads them as zero (tuning key grid_prune).
  uint2* zero_cols = nullptr;      // [G_r] or null
  uint8_t* block_visited = nullptr;  // [block_rows * block_cols] scratch of zero_cols
  // scratch used only while building
  double* frac = nullptr;          // [S][2]
  uint32_t* sorted_keys = nullptr; // [S]
  uint32_t* counters = nullptr;    // [4]: multi-item block count, error flags
  void* sort_scratch = nullptr;
  size_t device_bytes = 0;
  bool has_trajectory = false;
};

Plan2* plan2_create(const Plan2Desc& desc);
void plan2_destroy(Plan2* plan);
// Recomputes every table from a float32 [S][2] radian-per-pixel trajectory on the
// device.  dcf is read only when density == kDensityGiven.  Synchronises `stream`
// (the schedule is sized on the host) and throws on a trajectory that cannot be
// mapped (non-finite or outside [-pi, pi]).
void plan2_set_trajectory(Plan2* plan, const float* d_traj, const float* d_dcf, hipStream_t stream);

// 1 / phi_hat(j) for output index r (j = r - n/2) of an n-point crop of a grid-point
// axis: the continuous FT of the width-W Kaiser-Bessel kernel at 2 pi j / grid, FP64.
void deapod_table(float* d_table, uint32_t n, uint32_t grid, int width, double beta, hipStream_t stream);
// start[b] = first position of bin b in the sorted keys (start[nbins] = samples).
void cell_starts(const uint32_t* d_sorted_keys, uint32_t samples, uint64_t nbins, uint32_t* d_start,
                 hipStream_t stream);

// Golden-angle radial trajectory (FP64 on the device, rounded to float):
// spoke p at angle p * pi / phi, readout k_j = (j - M/2) / M cycles per pixel,
// written as [S][2] (ky, kx) in radians per pixel.
void golden_angle2(float* d_traj, uint32_t spokes, uint32_t readout, hipStream_t stream);

// Route trajectory normalisation (noncartesian_rss_hdf5.cpp:776-834): raw [S][kx, ky] in
// units 1 = cycles/FOV, 2 = rad/px, 3 = encoded matrix index -> [S][ky, kx] rad/px, float
// arithmetic as the reference.  Synchronises `stream`; throws Error(1) naming the first
// failing coordinate's reason (non-finite, outside the Nyquist interval, unspecified units).
void normalize_trajectory2(const float* d_raw, uint64_t samples, int units, uint32_t encoded_rows,
                           uint32_t encoded_cols, float* d_out, hipStream_t stream);

// Device-side trajectory check: *d_flags = bit 0 when some of the `values` floats differ
// bitwise from d_resident (always when d_resident is null), bit 1 when some value is not
// finite or lies outside [-bound, bound].  Asynchronous on `stream`.
void trajectory_check(const float* d_traj, const float* d_resident, uint64_t values, float bound,
                      uint32_t* d_flags, hipStream_t stream);

// Analytic ramp DCF hypot(w_r, w_c) per sample in float (eigen_radial_gridding.cpp:279-295).
void ramp_dcf2(const float* d_traj, float* d_dcf, uint64_t samples, hipStream_t stream);

// ISMRMRD acquisitions, packed one after another each as [C][ns_a], -> the route's k-space
// frame [C][S] (the replay's copy loop, noncartesian_rss_hdf5.cpp:1077-1083).  Counts are
// host values.  Every acquisition with the same count: asynchronous, no workspace; otherwise
// the offsets go to d_workspace (assemble_workspace_bytes) and `stream` is synchronised.
size_t assemble_workspace_bytes(uint32_t acquisitions);
void assemble_acquisitions(const float2* d_packed, const uint32_t* samples_per_acquisition, uint32_t acquisitions,
                           uint32_t channels, float2* d_frame, void* d_workspace, size_t workspace_bytes,
                           hipStream_t stream);

}  // namespace ksj::gpu
