// Gather gridding onto the (oversampled) Cartesian grid, driven by a Plan2.
#pragma once

#include "plan.hpp"

namespace ksj::gpu {

// Padded coil count Cp of the sorted sample rows and of the grid: the next power
// of two >= max(8, coils) (lanes are mapped to coils).
uint32_t padded_coils(uint32_t coils);

constexpr uint32_t kMaxStackedCoils = 64;

// Coil planes of one gather pass: coil c reads samples planes[c][0..S).  Several
// scans on the same trajectory are stacked along the coil axis this way.
struct CoilPlanes {
  const float2* plane[kMaxStackedCoils];
  uint32_t coils = 0;
};

// Applies the plan's per-sample weight and moves the coil planes into sorted,
// coil-fastest rows out[k][Cp] (pad zeroed).
void grid_prep(const Plan2& plan, const CoilPlanes& planes, float2* d_sorted, hipStream_t stream);
// The same from a plan's per-sample weights and sorted ranks (any dimension).
void grid_prep_rows(const float2* d_weight, const uint32_t* d_rank, uint32_t samples, const CoilPlanes& planes,
                    float2* d_sorted, hipStream_t stream);

// Bytes of the per-item partial-sum buffer grid_gather needs (blocks with several items).
size_t grid_partial_bytes(const Plan2& plan, uint32_t coils);

// Gathers every grid cell of all coils into the coil-fastest grid [G_r][G_c][Cp].
// Every cell is written (no prior clear needed) and the result is
// bit-reproducible for a given plan.
void grid_gather(const Plan2& plan, const float2* d_sorted, uint32_t coils, float2* d_grid, float2* d_partial,
                 hipStream_t stream);
// Sums the partial results of blocks cut into several work items into the grid
// (item order, so the sum is deterministic); a no-op when no block was cut.
void grid_reduce(const Plan2& plan, uint32_t coils, float2* d_grid, const float2* d_partial, hipStream_t stream);

// One item of a fused uneven batch: a plan (2x2-block or matrix schedule), its [C][S] k-space
// and the item's own sorted-row, grid and partial-sum buffers.
struct GridFusedItem {
  const Plan2* plan = nullptr;
  const float2* kspace = nullptr;
  uint32_t channels = 0;
  float2* sorted = nullptr;
  float2* grid = nullptr;
  float2* partial = nullptr;
};
// Device bytes of the per-launch argument tables grid_fused writes.
size_t grid_fused_table_bytes();
// prep, gather and reduce of up to 64 items, one launch per stage and padded coil
// count; every item's results equal grid_prep + grid_gather + grid_reduce on it.
void grid_fused(const GridFusedItem* items, uint32_t count, void* d_tables, hipStream_t stream);

// The same three stages as one persistent launch fed by a per-XCD device work queue
// (every item's tasks on one XCD; waits bounded).  Same bits as grid_fused.  Takes
// matrix-schedule plans with packed visits of one block width, with 16 visits per batch;
// grid_chain_supported says whether `items` qualify and the device's XCD ids allow it.
size_t grid_chain_table_bytes();
bool grid_chain_supported(const GridFusedItem* items, uint32_t count, hipStream_t stream);
void grid_chain(const GridFusedItem* items, uint32_t count, void* d_tables, hipStream_t stream);
// Probes the current device's XCD ids once (it synchronises; not while `stream` is captured).
void grid_chain_prepare(hipStream_t stream);
// Chain launches on the current device whose bounded waits ran out (0 in correct runs).
uint64_t grid_chain_faults();

}  // namespace ksj::gpu
