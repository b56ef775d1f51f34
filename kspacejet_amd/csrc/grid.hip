// Gather gridding kernels (no float atomics, fixed summation order).
//
// prep:   k-space [C][S] x per-sample weight -> sorted rows [S][Cp] (coil fastest);
//         staged through LDS so reads are coalesced per coil plane and each
//         sorted row is written as whole cache lines.
// gather: the plan's schedule of 2x2 cell blocks, cut into work items of at most
//         item_visits visits (so the dense k-space centre is spread over many
//         waves); one sub-wave per item with lanes mapped to coils, accumulating
//         the four cells in registers; blocks with several items are summed item
//         by item by a second kernel.  Every cell is written exactly once into the
//         coil-fastest grid [G_r][G_c][Cp].
#include "grid.hpp"

namespace ksj::gpu {
namespace {

// ----------------------------------------------------------------------------- prep
constexpr int kPrepSamples = 64;

struct PrepArgs {
  CoilPlanes planes;     // coil c -> [S] samples
  const float2* weight;  // [S]
  const uint32_t* rank;  // [S]
  float2* sorted;        // [S][Cp]
  uint32_t samples, coils, cp;
};

__global__ __launch_bounds__(256) void prep_kernel(PrepArgs a) {
  extern __shared__ float2 tile[];  // [kPrepSamples][cp + 1]; the pad spreads a coil column over the banks
  const uint32_t ts = a.cp + 1;
  const uint32_t s0 = blockIdx.x * kPrepSamples;
  const uint32_t ns = a.samples - s0 < static_cast<uint32_t>(kPrepSamples) ? a.samples - s0 : kPrepSamples;
  for (uint32_t i = threadIdx.x; i < kPrepSamples * a.cp; i += blockDim.x) {
    const uint32_t c = i / kPrepSamples, sl = i % kPrepSamples;
    float2 v = make_float2(0.f, 0.f);
    if (c < a.coils && sl < ns) v = cmul(a.planes.plane[c][s0 + sl], a.weight[s0 + sl]);
    tile[sl * ts + c] = v;
  }
  __syncthreads();
  // Each row of cp complex values is written by cp/2 consecutive lanes (16 B each).
  const uint32_t pieces = a.cp / 2;
  for (uint32_t i = threadIdx.x; i < ns * pieces; i += blockDim.x) {
    const uint32_t sl = i / pieces, pc = i % pieces;
    const float2 lo = tile[sl * ts + 2 * pc], hi = tile[sl * ts + 2 * pc + 1];
    float4* dst = reinterpret_cast<float4*>(a.sorted + static_cast<uint64_t>(a.rank[s0 + sl]) * a.cp);
    dst[pc] = make_float4(lo.x, lo.y, hi.x, hi.y);
  }
}

// ----------------------------------------------------------------------------- gather
struct GatherArgs {
  const float2* sorted;             // [S][Cp]
  const uint32_t* visit_k;          // [visits]
  const float4* visit_w;            // [visits]
  const uint4* item_info;           // [items] {block, first visit, end visit, slot or ~0}
  const uint32_t* block_first_item; // [nblocks + 1]
  const uint32_t* block_first_slot; // [nblocks + 1]
  const uint32_t* multi_blocks;     // [multi_count]
  float2* partial;                  // [slots][4][Cp]
  float2* grid;                     // [G_r][G_c][Cp]
  uint32_t grid_rows, grid_cols, block_cols, cp, item_count, multi_count;
};

__device__ inline void fma4(float2 (&acc)[4], float4 w, float2 d) {
  acc[0].x += w.x * d.x;
  acc[0].y += w.x * d.y;
  acc[1].x += w.y * d.x;
  acc[1].y += w.y * d.y;
  acc[2].x += w.z * d.x;
  acc[2].y += w.z * d.y;
  acc[3].x += w.w * d.x;
  acc[3].y += w.w * d.y;
}

// One sub-wave of SW lanes (lane = coil) per work item, 64/SW items per wave.  The
// sub-wave walks its item's visits in batches of B: lane j < B loads the sample
// position of visit j (one contiguous load), the positions are shared with
// cross-lane shuffles, then all B sample-row loads (SW consecutive coils each) are
// issued before the first use, so B loads are in flight per lane.  Weights are
// broadcast loads independent of the positions.  Sums run in visit order.
// One batch of B visits; `kj` holds (on lanes c < B) the sample positions of the
// batch, loaded one batch ahead by the caller.
template <int SW, int B, bool TAIL>
__device__ inline void gather_batch(const GatherArgs& a, float2 (&acc)[4], uint32_t kj, uint32_t v, uint32_t n,
                                    int sw, int c) {
  const float2* rows = a.sorted + c;
  uint32_t k[B];
  float4 w[B];
  float2 d[B];
#pragma unroll
  for (int i = 0; i < B; ++i) {
    k[i] = __shfl(kj, sw * SW + i, 64);
    w[i] = (!TAIL || static_cast<uint32_t>(i) < n) ? a.visit_w[v + i] : make_float4(0.f, 0.f, 0.f, 0.f);
  }
#pragma unroll
  for (int i = 0; i < B; ++i) d[i] = (!TAIL || static_cast<uint32_t>(i) < n) ? rows[k[i] * a.cp] : make_float2(0.f, 0.f);
#pragma unroll
  for (int i = 0; i < B; ++i) fma4(acc, w[i], d[i]);
}

template <int B>
__device__ inline uint32_t load_positions(const GatherArgs& a, uint32_t v, uint32_t end, int c) {
  return (c < B && v + c < end) ? a.visit_k[v + c] : 0u;
}

// One sub-wave of SW lanes (lane = coil) per work item, 64/SW items per wave.  The
// sub-wave walks its item's visits in batches of B: lane j < B loads the sample
// position of visit j (one contiguous load), the positions are shared with
// cross-lane shuffles, then all B sample-row loads (SW consecutive coils each) are
// issued before the first use, so B loads are in flight per lane.  Weights are
// broadcast loads independent of the positions.  Sums run in visit order.
template <int SW, int B, bool TAIL>
__device__ inline void gather_batch(const GatherArgs& a, float2 (&acc)[4], uint32_t v, uint32_t n, int sw, int c) {
  const float2* rows = a.sorted + c;
  const uint32_t kj = (c < B && (!TAIL || static_cast<uint32_t>(c) < n)) ? a.visit_k[v + c] : 0u;
  uint32_t k[B];
  float4 w[B];
  float2 d[B];
#pragma unroll
  for (int i = 0; i < B; ++i) {
    k[i] = __shfl(kj, sw * SW + i, 64);
    w[i] = (!TAIL || static_cast<uint32_t>(i) < n) ? a.visit_w[v + i] : make_float4(0.f, 0.f, 0.f, 0.f);
  }
#pragma unroll
  for (int i = 0; i < B; ++i) d[i] = (!TAIL || static_cast<uint32_t>(i) < n) ? rows[k[i] * a.cp] : make_float2(0.f, 0.f);
#pragma unroll
  for (int i = 0; i < B; ++i) fma4(acc, w[i], d[i]);
}

// One sub-wave of SW lanes (lane = coil) per work item, 64/SW items per wave.  The
// sub-wave walks its item's visits in batches of B: lane j < B loads the sample
// position of visit j (one contiguous load), the positions are shared with
// cross-lane shuffles, then all B sample-row loads (SW consecutive coils each) are
// issued before the first use, so B loads are in flight per lane.  Weights are
// broadcast loads independent of the positions.  Sums run in visit order.
template <int SW, int B>
__global__ __launch_bounds__(256) void gather_item_kernel(GatherArgs a) {
  constexpr int NSW = 64 / SW;
  const int lane = threadIdx.x & 63, sw = lane / SW, c = lane % SW;
  const uint32_t item = (blockIdx.x * 4 + (threadIdx.x >> 6)) * NSW + sw;
  const bool live = item < a.item_count;
  const uint4 info = live ? a.item_info[item] : make_uint4(0, 0, 0, 0);
  float2 acc[4];
#pragma unroll
  for (int q = 0; q < 4; ++q) acc[q] = make_float2(0.f, 0.f);
  // Software pipeline: the positions of batch i+1 are in flight while batch i's rows load.
  uint32_t v = info.y;
  uint32_t kj = load_positions<B>(a, v, info.z, c);
  for (; v + B <= info.z; v += B) {
    const uint32_t kn = load_positions<B>(a, v + B, info.z, c);
    gather_batch<SW, B, false>(a, acc, kj, v, B, sw, c);
    kj = kn;
  }
  if (v < info.z) gather_batch<SW, B, true>(a, acc, kj, v, info.z - v, sw, c);
  if (!live) return;
  const uint32_t by = info.x / a.block_cols, bx = info.x % a.block_cols;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    if (info.w == 0xFFFFFFFFu) {
      const uint32_t y = 2 * by + q / 2, x = 2 * bx + q % 2;
      if (y < a.grid_rows && x < a.grid_cols)
        a.grid[(static_cast<uint64_t>(y) * a.grid_cols + x) * a.cp + c] = acc[q];
    } else {
      a.partial[(static_cast<uint64_t>(info.w) * 4 + q) * a.cp + c] = acc[q];
    }
  }
}

// Blocks with several items: thread per (block, cell, coil) sums the item partials in item order.
__global__ __launch_bounds__(256) void multi_reduce_kernel(GatherArgs a) {
  const uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= static_cast<uint64_t>(a.multi_count) * 4 * a.cp) return;
  const uint32_t c = static_cast<uint32_t>(i % a.cp);
  const uint32_t q = static_cast<uint32_t>((i / a.cp) % 4);
  const uint32_t b = a.multi_blocks[i / (4ull * a.cp)];
  const uint32_t s0 = a.block_first_slot[b];
  const uint32_t n = a.block_first_item[b + 1] - a.block_first_item[b];
  const float2* p = a.partial + (static_cast<uint64_t>(s0) * 4 + q) * a.cp + c;
  const uint64_t stride = 4ull * a.cp;
  float2 s = make_float2(0.f, 0.f);
  uint32_t j = 0;
  for (; j + 8 <= n; j += 8) {  // eight loads in flight, summed in item order
    float2 t[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) t[u] = p[(j + u) * stride];
#pragma unroll
    for (int u = 0; u < 8; ++u) s = cadd(s, t[u]);
  }
  for (; j < n; ++j) s = cadd(s, p[j * stride]);
  const uint32_t y = 2 * (b / a.block_cols) + q / 2, x = 2 * (b % a.block_cols) + q % 2;
  if (y < a.grid_rows && x < a.grid_cols) a.grid[(static_cast<uint64_t>(y) * a.grid_cols + x) * a.cp + c] = s;
}

template <int SW>
void launch_gather(GatherArgs a, hipStream_t stream) {
  constexpr int NSW = 64 / SW;
  const uint32_t waves = ceil_div(a.item_count, NSW);
  hipLaunchKernelGGL((gather_item_kernel<SW, 8>), dim3(ceil_div(waves, 4)), dim3(256), 0, stream, a);
  KSJ_LAUNCH_CHECK();
  if (a.multi_count > 0) {
    hipLaunchKernelGGL(multi_reduce_kernel, dim3(ceil_div(static_cast<uint64_t>(a.multi_count) * 4 * a.cp, 256)),
                       dim3(256), 0, stream, a);
    KSJ_LAUNCH_CHECK();
  }
}

}  // namespace

uint32_t padded_coils(uint32_t coils) {
  uint32_t cp = 8;
  while (cp < coils) cp *= 2;
  return cp;
}

size_t grid_partial_bytes(const Plan2& plan, uint32_t coils) {
  return static_cast<size_t>(plan.slot_count) * 4 * padded_coils(coils) * sizeof(float2);
}

void grid_prep(const Plan2& plan, const CoilPlanes& planes, float2* d_sorted, hipStream_t stream) {
  require(plan.has_trajectory, "plan has no trajectory");
  require(planes.coils >= 1 && planes.coils <= kMaxStackedCoils, "gridding takes 1..64 stacked coils");
  const uint32_t coils = planes.coils;
  PrepArgs a{};
  a.planes = planes;
  a.weight = plan.weight;
  a.rank = plan.rank;
  a.sorted = d_sorted;
  a.samples = plan.d.samples;
  a.coils = coils;
  a.cp = padded_coils(coils);
  const size_t shmem = static_cast<size_t>(kPrepSamples) * (a.cp + 1) * sizeof(float2);
  hipLaunchKernelGGL(prep_kernel, dim3(ceil_div(a.samples, kPrepSamples)), dim3(256), shmem, stream, a);
  KSJ_LAUNCH_CHECK();
}

void grid_gather(const Plan2& plan, const float2* d_sorted, uint32_t coils, float2* d_grid, float2* d_partial,
                 hipStream_t stream) {
  require(plan.has_trajectory, "plan has no trajectory");
  require(plan.multi_count == 0 || d_partial != nullptr, "multi-item blocks need a partial-sum buffer");
  GatherArgs a{};
  a.sorted = d_sorted;
  a.visit_k = plan.visit_k;
  a.visit_w = plan.visit_w;
  a.item_info = plan.item_info;
  a.block_first_item = plan.block_first_item;
  a.block_first_slot = plan.block_first_slot;
  a.multi_blocks = plan.multi_blocks;
  a.partial = d_partial;
  a.grid = d_grid;
  a.grid_rows = plan.d.grid_rows;
  a.grid_cols = plan.d.grid_cols;
  a.block_cols = plan.block_cols;
  a.cp = padded_coils(coils);
  a.item_count = plan.item_count;
  a.multi_count = plan.multi_count;
  switch (a.cp) {
    case 8: return launch_gather<8>(a, stream);
    case 16: return launch_gather<16>(a, stream);
    case 32: return launch_gather<32>(a, stream);
    case 64: return launch_gather<64>(a, stream);
    default: throw Error(1, "gridding supports at most 64 coils");
  }
}

}  // namespace ksj::gpu
