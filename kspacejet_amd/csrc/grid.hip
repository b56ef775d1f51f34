// Gather gridding kernels (no float atomics, fixed summation order).
//
// prep:   k-space [C][S] x per-sample weight -> sorted rows [S][Cp] (coil fastest);
//         staged through LDS so reads are coalesced per coil plane and each
//         sorted row is written as whole cache lines.
// gather (blocks): the plan's schedule of 2x2 cell blocks, cut into work items of at most
//         item_visits visits (so the dense k-space centre is spread over many
//         waves); one sub-wave per item with lanes mapped to coils, accumulating
//         the four cells in registers; blocks with several items are summed item
//         by item by a second kernel.  Every cell is written exactly once into the
//         coil-fastest grid [G_r][G_c][Cp].
// gather (tiles): one workgroup per work item of a 64-cell tile; the tile's
//         window samples are staged in LDS and every cell sums its footprint
//         from there (below).
#include "grid.hpp"

#include "fused.hpp"
#include "matrix.hpp"

#include <algorithm>
#include <atomic>
#include <cstring>
#include <memory>
#include <mutex>
#include <utility>
#include <vector>

namespace ksj::gpu {
namespace {

// ----------------------------------------------------------------------------- prep
constexpr int kPrepSamples = 64;  // one sample per lane of a wave
constexpr int kPrepWaves = 4;

// threadIdx.x behind an empty asm: values derived from it are then recomputed where they
// are used instead of being hoisted out of a persistent kernel's task loop, where the
// hoisted values of every inlined stage body would stay live at once.
__device__ inline uint32_t opaque_tid() {
  uint32_t t = threadIdx.x;
  asm volatile("" : "+v"(t));
  return t;
}

struct PrepArgs {
  CoilPlanes planes;     // coil c -> [S] samples
  const float2* weight;  // [S]
  const uint32_t* rank;  // [S]
  float2* sorted;        // [S][Cp]
  uint32_t samples, coils;
};

// Block = 64 consecutive samples; wave w reads coils w, w+4, ... (one coalesced
// 512-byte read per coil plane, all issued before the first wait), weights them
// and transposes through LDS; then every sorted row of CP complex values is
// written as CP/2 16-byte pieces at the sample's rank.
// Coil planes of one fused item: plane c at base + c * stride (a [C][S] k-space).
struct StridedPlanes {
  const float2* base;
  uint64_t stride;
};
__device__ inline const float2* plane_of(const CoilPlanes& p, uint32_t c) { return p.plane[c]; }
__device__ inline const float2* plane_of(const StridedPlanes& p, uint32_t c) { return p.base + c * p.stride; }

// LDS of one prep block: the padded tile, then the block's ranks.
template <int CP>
constexpr size_t prep_lds_bytes() {
  return sizeof(float2) * kPrepSamples * (CP + 1) + sizeof(uint32_t) * kPrepSamples;
}

template <int CP, typename Planes>
__device__ inline void prep_block_in(const Planes& planes, const float2* weight, const uint32_t* rank, float2* sorted,
                                     uint32_t samples, uint32_t coils, uint32_t block, float2* tile,
                                     uint32_t* rank_s) {
  constexpr int kTs = CP + 1;               // padded row: a coil column spreads over the banks
  constexpr int kPerWave = CP / kPrepWaves;  // coils read by one wave
  const uint32_t tid = opaque_tid();
  const uint32_t lane = tid & 63;
  const uint32_t wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const uint32_t s0 = block * kPrepSamples;
  const uint32_t s = s0 + lane;
  const bool live = s < samples;
  const float2 w = live ? weight[s] : make_float2(0.f, 0.f);
  const uint32_t r = live ? rank[s] : 0u;
  float2 v[kPerWave];
  // Branch-free, so all of a wave's plane loads are in flight together (a load under a
  // branch waits for its plane pointer first): a coil past the end reads plane 0 and a
  // sample past the end the last sample, and both are multiplied by 0 (a live value by
  // 1, exactly itself).  Such values only reach padded coil columns, which no output
  // reads, and rows past the end, which are not stored.
  const uint32_t sc = live ? s : samples - 1;
  const float2* base[kPerWave];
#pragma unroll
  for (int j = 0; j < kPerWave; ++j) {
    const uint32_t c = wave + kPrepWaves * j;  // wave-uniform: the plane pointer is a scalar load
    base[j] = plane_of(planes, c < coils ? c : 0u);
  }
#pragma unroll
  for (int j = 0; j < kPerWave; ++j) {
    const uint32_t c = wave + kPrepWaves * j;
    const float keep = (c < coils && live) ? 1.f : 0.f;
    // Each k-space value is read exactly once: keep it out of the caches.
    typedef float F2v __attribute__((ext_vector_type(2)));
    const F2v t = __builtin_nontemporal_load(reinterpret_cast<const F2v*>(base[j] + sc));
    v[j] = make_float2(t.x * keep, t.y * keep);
  }
  if (wave == 0) rank_s[lane] = r;  // after the sample loads are issued
#pragma unroll
  for (int j = 0; j < kPerWave; ++j) tile[lane * kTs + wave + kPrepWaves * j] = cmul(v[j], w);
  __syncthreads();
  const uint32_t ns = samples - s0 < static_cast<uint32_t>(kPrepSamples) ? samples - s0 : kPrepSamples;
  constexpr int kPieces = CP / 2;  // float4 pieces per sorted row
  constexpr int kIters = kPrepSamples * kPieces / (kPrepSamples * kPrepWaves);
#pragma unroll
  for (int it = 0; it < kIters; ++it) {
    const uint32_t i = tid + it * kPrepSamples * kPrepWaves;
    const uint32_t sl = i / kPieces, pc = i % kPieces;
    if (sl < ns) {
      const float2 lo = tile[sl * kTs + 2 * pc], hi = tile[sl * kTs + 2 * pc + 1];
      float4* dst = reinterpret_cast<float4*>(sorted + static_cast<uint64_t>(rank_s[sl]) * CP);
      dst[pc] = make_float4(lo.x, lo.y, hi.x, hi.y);
    }
  }
}

template <int CP, typename Planes>
__device__ inline void prep_block(const Planes& planes, const float2* weight, const uint32_t* rank, float2* sorted,
                                  uint32_t samples, uint32_t coils, uint32_t block) {
  __shared__ float2 tile[kPrepSamples * (CP + 1)];
  __shared__ uint32_t rank_s[kPrepSamples];
  prep_block_in<CP>(planes, weight, rank, sorted, samples, coils, block, tile, rank_s);
}

template <int CP>
__global__ __launch_bounds__(kPrepSamples * kPrepWaves) void prep_kernel(PrepArgs a) {
  prep_block<CP>(a.planes, a.weight, a.rank, a.sorted, a.samples, a.coils, blockIdx.x);
}

// ----------------------------------------------------------------------------- gather
constexpr uint32_t kXcds = 8;  // MI355X: workgroups are dispatched round-robin over 8 XCDs
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
  const uint32_t* xcd_first;        // [9] cost-balanced XCD item ranges (gather_map 2)
  uint32_t xcd_chunk, wg_per_xcd;   // workgroup -> item range mapping (gather_map)
  uint32_t grid_rows, grid_cols, block_cols, block_group, cp, item_count, multi_count;
  uint32_t bh, bw;                  // block cells per row / column: 2 x 2, or 4 x 4 / 4 x 8 (matrix)
  const uint32_t* run_first;        // matrix schedule: [runs + 1] first item of each wave run
  uint32_t run_count;
  uint32_t nt_store;                // matrix schedule: grid cells written with non-temporal stores
  uint32_t slab;                    // matrix schedule: workgroups per XCD slab (gather_map 0)
  uint32_t persist;                 // matrix schedule: 1 = resident waves loop over their XCD's runs
  const float* taps;                // [S][2W] sorted taps (matrix schedule with packed visits)
  uint32_t width;                   // W
  uint32_t from_taps;               // 1: visit_k packs (sample, window position), weights formed from taps
  const uint2* zero_cols;           // matrix schedule: per grid row, columns no sample reaches (not written), or null
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
// sub-wave walks its item's visits in batches of B and issues all B sample-row
// loads (SW consecutive coils each) before the first use.  The next batch's
// positions and weights are loaded right after this batch's are distributed, so
// their latency overlaps this batch's row loads and the row loads never wait for
// metadata.  Sums run in visit order.
//
// Per-visit metadata is loaded once per 16-lane row and distributed inside the row:
// lane r of every row loads position r and the weights (float4) of visit r % B of
// the batch, and DPP row_newbcast (broadcast of lane K of each 16-lane row) hands
// them to every lane.
// A broadcast float4 load per visit would move 1 KiB through the texture units for
// 16 bytes of data -- twice the bytes of the sample row itself.  Sub-waves of 8 lanes
// (two per row) fall back to cross-lane shuffles.
// Visit index v + r clamped into the item (an empty item reads visit 0 of the
// plan's arrays, which hold at least one entry): the metadata loads are
// unconditional, and values past the item are never used -- their rows are not
// loaded and their sample values are zero.
__device__ inline uint32_t clamp_visit(uint32_t v, uint32_t end) {
  const uint32_t last = end > 0 ? end - 1 : 0;
  return v < last ? v : last;
}

template <int B>
__device__ inline uint32_t load_positions(const GatherArgs& a, uint32_t v, uint32_t end, int r) {
  return a.visit_k[clamp_visit(v + static_cast<uint32_t>(r), end)];
}

template <int K>
__device__ inline uint32_t row_bcast(uint32_t x) {
  return static_cast<uint32_t>(__builtin_amdgcn_mov_dpp(static_cast<int>(x), 0x150 + K, 0xF, 0xF, false));
}

// Weight component F (= 4 * visit + cell) of a batch, held by lane F % 16 in load F / 16.
template <int F, int NL>
__device__ inline float weight_comp(const uint32_t (&wl)[NL]) {
  return __uint_as_float(row_bcast<F % 16>(wl[F / 16]));
}

// CPL coils per lane: a lane reads CPL consecutive coils of a sample row with one
// 8*CPL-byte load (16-byte loads move twice the data per texture-data cycle of
// 8-byte ones), so an item of Cp coils takes SW = Cp / CPL lanes.
template <int CPL>
struct CoilVec;
template <>
struct CoilVec<1> {
  using T = float2;
  __device__ static void split(const T& v, float2 (&o)[1]) { o[0] = v; }
};
template <>
struct CoilVec<2> {
  using T = float4;
  __device__ static void split(const T& v, float2 (&o)[2]) {
    o[0] = make_float2(v.x, v.y);
    o[1] = make_float2(v.z, v.w);
  }
};

// Weights of the batch starting at visit v, as lane r of every 16-lane row holds
// them (visit v + r % B).
template <int B>
__device__ inline float4 load_weights(const GatherArgs& a, uint32_t v, uint32_t end, int r) {
  return a.visit_w[clamp_visit(v + static_cast<uint32_t>(r % B), end)];
}

template <int SW, int B, int CPL, bool TAIL>
__device__ inline void gather_batch(const GatherArgs& a, float2 (&acc)[4][CPL], uint32_t& kj, float4& wl,
                                    uint32_t v, uint32_t end, int sw, int c) {
  using V = typename CoilVec<CPL>::T;
  const V* rows = reinterpret_cast<const V*>(a.sorted) + c;
  const uint32_t rstride = a.cp / CPL;  // row length in V elements
  const uint32_t n = end - v;
  uint32_t k[B];
  float4 w[B];
  if constexpr (SW >= 16) {
    // Lane r of every 16-lane row holds visit r % B's weights (one float4) and, for
    // r < B, its position, both loaded one batch ahead; DPP row broadcasts hand
    // them to every lane of the row, then the next batch's are loaded.
    const int r = c & 15;
    [&]<int... I>(std::integer_sequence<int, I...>) {
      ((k[I] = row_bcast<I>(kj)), ...);
    }(std::make_integer_sequence<int, B>{});
    [&]<int... I>(std::integer_sequence<int, I...>) {
      ((w[I] = make_float4(__uint_as_float(row_bcast<I>(__float_as_uint(wl.x))),
                           __uint_as_float(row_bcast<I>(__float_as_uint(wl.y))),
                           __uint_as_float(row_bcast<I>(__float_as_uint(wl.z))),
                           __uint_as_float(row_bcast<I>(__float_as_uint(wl.w))))),
       ...);
    }(std::make_integer_sequence<int, B>{});
    if (!TAIL) {
      kj = load_positions<B>(a, v + B, end, r);
      wl = load_weights<B>(a, v + B, end, r);
    }
  } else {
#pragma unroll
    for (int i = 0; i < B; ++i) k[i] = __shfl(kj, sw * SW + i, 64);
    if (!TAIL) kj = load_positions<B>(a, v + B, end, c);
#pragma unroll
    for (int i = 0; i < B; ++i)
      w[i] = (!TAIL || static_cast<uint32_t>(i) < n) ? a.visit_w[v + i] : make_float4(0.f, 0.f, 0.f, 0.f);
  }
  V d[B];
#pragma unroll
  for (int i = 0; i < B; ++i) {
    if (!TAIL || static_cast<uint32_t>(i) < n) {
      d[i] = rows[static_cast<uint64_t>(k[i]) * rstride];
    } else {
      d[i] = V{};
    }
  }
#pragma unroll
  for (int i = 0; i < B; ++i) {
    float2 e[CPL];
    CoilVec<CPL>::split(d[i], e);
#pragma unroll
    for (int j = 0; j < CPL; ++j) {
      float2 q[4] = {acc[0][j], acc[1][j], acc[2][j], acc[3][j]};
      fma4(q, w[i], e[j]);
      acc[0][j] = q[0];
      acc[1][j] = q[1];
      acc[2][j] = q[2];
      acc[3][j] = q[3];
    }
  }
}

// One work item: gather its visits into four accumulators per coil and write its
// cells (or its partial-sum slot).  `live` is false for lanes past the item range.
template <int SW, int B, int CPL>
__device__ inline void gather_one_item(const GatherArgs& a, uint32_t item, bool live, int sw, int c) {
  const uint4 info = live ? a.item_info[item] : make_uint4(0, 0, 0, 0);
  float2 acc[4][CPL];
#pragma unroll
  for (int q = 0; q < 4; ++q)
#pragma unroll
    for (int j = 0; j < CPL; ++j) acc[q][j] = make_float2(0.f, 0.f);
  uint32_t v = info.y;
  uint32_t kj = load_positions<B>(a, v, info.z, SW >= 16 ? (c & 15) : c);
  float4 wl = SW >= 16 ? load_weights<B>(a, v, info.z, c & 15) : make_float4(0.f, 0.f, 0.f, 0.f);
  for (; v + B <= info.z; v += B) gather_batch<SW, B, CPL, false>(a, acc, kj, wl, v, info.z, sw, c);
  if (v < info.z) gather_batch<SW, B, CPL, true>(a, acc, kj, wl, v, info.z, sw, c);
  if (!live) return;
  using V = typename CoilVec<CPL>::T;
  uint32_t by, bx;
  block_coords(info.x, a.block_cols, a.block_group, &by, &bx);
  const uint32_t rstride = a.cp / CPL;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    V out;
    if constexpr (CPL == 1) {
      out = acc[q][0];
    } else {
      out = make_float4(acc[q][0].x, acc[q][0].y, acc[q][1].x, acc[q][1].y);
    }
    if (info.w == 0xFFFFFFFFu) {
      const uint32_t y = 2 * by + q / 2, x = 2 * bx + q % 2;
      if (y < a.grid_rows && x < a.grid_cols)
        reinterpret_cast<V*>(a.grid)[(static_cast<uint64_t>(y) * a.grid_cols + x) * rstride + c] = out;
    } else {
      reinterpret_cast<V*>(a.partial)[(static_cast<uint64_t>(info.w) * 4 + q) * rstride + c] = out;
    }
  }
}

// Item -> wave mappings (tuning key gather_map):
//  0: items in launch order, one group of 64/SW items per wave;
//  1: workgroup w (dispatched to XCD w % 8) takes the w/8-th group of an equal share
//     of items, so neighbouring 2x2 blocks -- which read the same sample rows --
//     share one XCD's L2;
//  2: as 1, with the eight ranges cut at equal cost (visits + a fixed cost per item)
//     by the plan, each XCD's workgroups striding through their range.
template <int SW, int B, int CPL>
__global__ __launch_bounds__(256) void gather_item_kernel(GatherArgs a) {
  constexpr int NSW = 64 / SW;
  const int lane = threadIdx.x & 63, sw = lane / SW, c = lane % SW;
  const uint32_t wave = threadIdx.x >> 6;
  if (a.xcd_chunk == 2) {
    const uint32_t x = blockIdx.x % kXcds, local = blockIdx.x / kXcds;
    const uint32_t end = a.xcd_first[x + 1];
    for (uint32_t base = a.xcd_first[x] + (local * 4 + wave) * NSW; base < end; base += a.wg_per_xcd * 4 * NSW) {
      const uint32_t item = base + sw;
      gather_one_item<SW, B, CPL>(a, item, item < end, sw, c);
    }
    return;
  }
  const uint32_t wg = a.xcd_chunk ? (blockIdx.x % kXcds) * a.wg_per_xcd + blockIdx.x / kXcds : blockIdx.x;
  const uint32_t item = (wg * 4 + wave) * NSW + sw;
  gather_one_item<SW, B, CPL>(a, item, item < a.item_count, sw, c);
}

// Blocks with several items: thread per (block, cell, coil) sums the item partials in item order.
__device__ inline void multi_reduce_one(const GatherArgs& a, uint64_t i) {
  const uint32_t cells = a.bh * a.bw;
  if (i >= static_cast<uint64_t>(a.multi_count) * cells * a.cp) return;
  const uint32_t c = static_cast<uint32_t>(i % a.cp);
  const uint32_t q = static_cast<uint32_t>((i / a.cp) % cells);
  const uint32_t b = a.multi_blocks[i / (static_cast<uint64_t>(cells) * a.cp)];
  const uint32_t s0 = a.block_first_slot[b];
  const uint32_t n = a.block_first_item[b + 1] - a.block_first_item[b];
  const float2* p = a.partial + (static_cast<uint64_t>(s0) * cells + q) * a.cp + c;
  const uint64_t stride = static_cast<uint64_t>(cells) * a.cp;
  float2 s = make_float2(0.f, 0.f);
  // Eight loads in flight, summed in item order; the last group reads clamped indices and
  // adds +0 for items past the end (s starts at +0 and never becomes -0, so s + 0 = s).
  for (uint32_t j = 0; j < n; j += 8) {
    float2 t[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const uint32_t i = j + u < n ? j + u : n - 1;
      t[u] = keep_bits(p[i * stride], j + u < n ? ~0u : 0u);
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) s = cadd(s, t[u]);
  }
  uint32_t by, bx;
  block_coords(b, a.block_cols, a.block_group, &by, &bx);
  const uint32_t y = a.bh * by + q / a.bw, x = a.bw * bx + q % a.bw;
  if (y < a.grid_rows && x < a.grid_cols) a.grid[(static_cast<uint64_t>(y) * a.grid_cols + x) * a.cp + c] = s;
}

__global__ __launch_bounds__(256) void multi_reduce_kernel(GatherArgs a) {
  multi_reduce_one(a, static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x);
}

// ----------------------------------------------------------------------------- matrix gather
// The 4x4 (or 4x8) block schedule as small dense products on the matrix cores: for
// four visits at a time, each 16-cell tile's 16 cells x 4 visits weight matrix times
// the 4 visits x 2 Cp floats of their sample rows (real weights act on real and
// imaginary parts alike), accumulated by v_mfma_f32_16x16x4_f32 in FP32.
//
// Operand layout of v_mfma_f32_16x16x4_f32 (wave64, one block): lane l holds
// A[l % 16][l / 16] and B[l / 16][l % 16]; accumulator register r of lane l is
// D[4 (l / 16) + r][l % 16].  So lane l takes visit q = l / 16 of the four and
//  * A: the weight of cell l % 16 of the tile (a visit's 16 T weights are contiguous);
//  * B: NF = Cp / 8 consecutive floats of the visit's row at column j = l % 16, one
//    per product: product n sums float NF j + n of every row;
//  * D: tile cells 4q .. 4q+3 at floats NF j .. NF j + NF - 1, written as one vector
//    per cell.
// Block visit lists are padded to whole groups of four with zero-weight visits.
// Sums run four visits at a time in visit order, so results are deterministic.
// One wave per run of consecutive items (at most 64): the items' groups of four are
// contiguous, so the wave streams through them in batches of U groups with the
// next batch's metadata (positions and weights) loaded right after this batch's
// row loads are issued.  When a group belongs to the next item, the current item's
// 16 cells are written (to the grid, or to its partial-sum slot when its block is
// cut into several items) and its accumulators cleared; items without visits
// (blocks no sample reaches) write zeros.  Lane t holds item t's record; the wave
// reads it with readlane.
//
// WT (weights from taps): a visit is one packed word, sorted sample k (bits 0-23) and the
// sample's bin in the block's window, row r (bits 24-27) and column c (bits 28-31).  The
// lane forms its cell's weight as row tap (dy - r + W - 1) times column tap
// (dx - c + W - 1) of the sample (0 outside [0, W)), the same float product the weight
// table holds, so both forms give the same bits.  Padding visits carry r = c = 15, which
// no cell reaches.  4 bytes of metadata per visit instead of 4 + 4 T x 16.
template <int CP, int U, int P, int T, bool WT, bool TS>
__device__ inline void matrix_run(const GatherArgs& a, uint32_t run) {
  constexpr int NF = CP / 8;
  constexpr int kCells = 16 * T;  // 4 x 4 (T = 1) or 4 x 8 (T = 2) cells, T products of 16 cells each
  constexpr uint32_t kKeyMask = WT ? 0x00FFFFFFu : 0xFFFFFFFFu;
  const int lane = static_cast<int>(opaque_tid() & 63), q = lane >> 4, j = lane & 15;
  // The run's bounds and its first and last visit through the scalar cache (run is
  // wave-uniform; plan tables are read-only here), so the first batch's visit loads do
  // not wait for the per-lane item records.
  const ConstWord run_first = (ConstWord)(a.run_first);
  const uint32_t i0 = run_first[run], i1 = run_first[run + 1];
  if (i0 >= i1) return;
  const uint32_t nitems = i1 - i0;  // <= 64 by construction of the runs
  const ConstWord info = (ConstWord)(a.item_info);
  const uint32_t c0 = info[4 * static_cast<uint64_t>(i0) + 1] / 4, c1 = info[4 * static_cast<uint64_t>(i1 - 1) + 2] / 4;
  const uint4 rec = static_cast<uint32_t>(lane) < nitems ? a.item_info[i0 + lane] : make_uint4(0, 0, 0, 0);
  auto field = [&](uint32_t v, uint32_t t) {
    return static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(v), static_cast<int>(t)));
  };
  const uint32_t last_visit = c1 > c0 ? 4 * c1 - 1 : 0;
  const float* wts = reinterpret_cast<const float*>(a.visit_w);
  const float* rows = reinterpret_cast<const float*>(a.sorted) + piece_offset<NF>(j);
  MfmaAcc acc[T][NF];
#pragma unroll
  for (int h = 0; h < T; ++h)
#pragma unroll
    for (int n = 0; n < NF; ++n) acc[h][n] = MfmaAcc{0.f, 0.f, 0.f, 0.f};

  // Cell 16 h + 4 q + r of the block (product h, lane group q, register r) is block
  // row (16 h + 4 q + r) / bw, column (16 h + 4 q + r) % bw.
  auto flush = [&](uint32_t t) {
    const uint32_t blk = field(rec.x, t), slot = field(rec.w, t);
    uint32_t by, bx;
    block_coords(blk, a.block_cols, a.block_group, &by, &bx);
    if (a.zero_cols != nullptr && slot == 0xFFFFFFFFu && 4 * by < a.grid_rows) {
      // A block inside its rows' zero span (no sample reaches it): the row FFT reads
      // those cells as zero, so they are not written.
      const uint2 zs = a.zero_cols[4 * by];
      if (4 * T * bx >= zs.x && 4 * T * (bx + 1) <= zs.y) {
#pragma unroll
        for (int h = 0; h < T; ++h)
#pragma unroll
          for (int n = 0; n < NF; ++n) acc[h][n] = MfmaAcc{0.f, 0.f, 0.f, 0.f};
        return;
      }
    }
#pragma unroll
    for (int h = 0; h < T; ++h) {
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        float o[NF];
#pragma unroll
        for (int n = 0; n < NF; ++n) o[n] = acc[h][n][r];
        const uint32_t cell = 16 * h + 4 * q + r;
        if (slot == 0xFFFFFFFFu) {
          const uint32_t y = 4 * by + cell / (4 * T), x = 4 * T * bx + cell % (4 * T);
          if (y < a.grid_rows && x < a.grid_cols) {
            float* dst =
              reinterpret_cast<float*>(a.grid) + (static_cast<uint64_t>(y) * a.grid_cols + x) * (2 * CP) + piece_offset<NF>(j);
            if (a.nt_store)
              store_piece_nt<NF>(dst, o);
            else
              store_piece<NF>(dst, o);
          }
        } else {
          store_piece<NF>(
            reinterpret_cast<float*>(a.partial) + (static_cast<uint64_t>(slot) * kCells + cell) * (2 * CP) + piece_offset<NF>(j),
            o);
        }
      }
    }
#pragma unroll
    for (int h = 0; h < T; ++h)
#pragma unroll
      for (int n = 0; n < NF; ++n) acc[h][n] = MfmaAcc{0.f, 0.f, 0.f, 0.f};
  };

  auto load_meta = [&](uint32_t c, uint32_t (&k)[U], float (&w)[U][T]) {  // groups c + u, clamped into the run
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const uint32_t vi = 4 * (c + u) + q;
      const uint32_t vc = vi < last_visit ? vi : last_visit;
      k[u] = a.visit_k[vc];
      if constexpr (!WT) {
#pragma unroll
        for (int h = 0; h < T; ++h) w[u][h] = wts[static_cast<uint64_t>(vc) * kCells + 16 * h + j];
      }
    }
  };
  auto load_rows = [&](const uint32_t (&k)[U], RowPiece<NF> (&b)[U]) {
#pragma unroll
    for (int u = 0; u < U; ++u) b[u] = load_piece<NF>(rows + static_cast<uint64_t>(k[u] & kKeyMask) * (2 * CP));
  };
  // WT: the lane's row and column tap of each visit (a clamped, always in-bounds index;
  // `live` marks the taps inside [0, W)).  Cell 16 h + j is block row (16 h + j) / (4 T),
  // column (16 h + j) % (4 T).
  const int W = static_cast<int>(a.width);
  auto load_taps = [&](const uint32_t (&k)[U], float (&tr)[U][T], float (&tc)[U][T], bool (&live)[U][T]) {
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const float* t = a.taps + static_cast<uint64_t>(k[u] & 0x00FFFFFFu) * (2 * W);
      const int r = static_cast<int>((k[u] >> 24) & 15u), cw = static_cast<int>(k[u] >> 28);
#pragma unroll
      for (int h = 0; h < T; ++h) {
        const int cell = 16 * h + j;
        const int ir = cell / (4 * T) - r + W - 1, ic = cell % (4 * T) - cw + W - 1;
        live[u][h] = static_cast<unsigned>(ir) < static_cast<unsigned>(W) &&
                     static_cast<unsigned>(ic) < static_cast<unsigned>(W);
        tr[u][h] = t[live[u][h] ? ir : 0];
        tc[u][h] = t[W + (live[u][h] ? ic : 0)];
      }
    }
  };

  uint32_t t = 0, t_end = info[4 * static_cast<uint64_t>(i0) + 2] / 4;  // current item and its end group
  auto compute = [&](uint32_t c, const RowPiece<NF> (&b)[U], const float (&wc)[U][T]) {
#pragma unroll
    for (int u = 0; u < U; ++u) {
      if (c + u < c1) {
        while (c + u >= t_end) {  // group c + u starts a later item
          flush(t);
          ++t;
          t_end = field(rec.z, t) / 4;
        }
#pragma unroll
        for (int h = 0; h < T; ++h)
#pragma unroll
          for (int n = 0; n < NF; ++n)
            acc[h][n] = __builtin_amdgcn_mfma_f32_16x16x4f32(wc[u][h], b[u].v[n], acc[h][n], 0, 0, 0);
      }
    }
  };
  auto copy_w = [](float (&dst)[U][T], const float (&src)[U][T]) {
#pragma unroll
    for (int u = 0; u < U; ++u)
#pragma unroll
      for (int h = 0; h < T; ++h) dst[u][h] = src[u][h];
  };

  uint32_t c = c0;
  if constexpr (WT) {
    // One batch of rows and taps in flight, then the positions of batch c + U.  With TS
    // the batch's 4 U packed visits are loaded one per lane (lane l: visit l mod 4 U) and
    // spread to the lanes of each group by bpermute.
    uint32_t k[U];
    float w[U][T];
    uint32_t kw = 0;
    auto load_keys = [&](uint32_t cc) {
      const uint32_t vi = 4 * cc + (static_cast<uint32_t>(lane) & (4 * U - 1));
      kw = a.visit_k[vi < last_visit ? vi : last_visit];
    };
    if (c < c1) {
      if constexpr (TS)
        load_keys(c);
      else
        load_meta(c, k, w);
    }
    for (; c < c1; c += U) {
      RowPiece<NF> b[U];
      float tr[U][T], tc[U][T], wc[U][T];
      bool live[U][T];
      if constexpr (TS) {
#pragma unroll
        for (int u = 0; u < U; ++u)
          k[u] = static_cast<uint32_t>(__builtin_amdgcn_ds_bpermute(4 * (4 * u + q), static_cast<int>(kw)));
      }
      load_rows(k, b);
      if constexpr (TS) {
        // One tap word per lane (lane j of visit q's 16 holds tap min(j, 2W - 1)); each
        // lane then takes its row and column tap from the visit's lanes by bpermute.
        uint32_t word[U];
        int src_r[U][T], src_c[U][T];
#pragma unroll
        for (int u = 0; u < U; ++u) {
          const float* t = a.taps + static_cast<uint64_t>(k[u] & 0x00FFFFFFu) * (2 * W);
          word[u] = __float_as_uint(t[j < 2 * W ? j : 2 * W - 1]);
          const int r = static_cast<int>((k[u] >> 24) & 15u), cw = static_cast<int>(k[u] >> 28);
#pragma unroll
          for (int h = 0; h < T; ++h) {
            const int cell = 16 * h + j;
            const int ir = cell / (4 * T) - r + W - 1, ic = cell % (4 * T) - cw + W - 1;
            live[u][h] = static_cast<unsigned>(ir) < static_cast<unsigned>(W) &&
                         static_cast<unsigned>(ic) < static_cast<unsigned>(W);
            src_r[u][h] = 4 * (16 * q + (live[u][h] ? ir : 0));
            src_c[u][h] = 4 * (16 * q + W + (live[u][h] ? ic : 0));
          }
        }
        load_keys(c + U);
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int u = 0; u < U; ++u)
#pragma unroll
          for (int h = 0; h < T; ++h) {
            const float pr = __uint_as_float(static_cast<uint32_t>(
              __builtin_amdgcn_ds_bpermute(src_r[u][h], static_cast<int>(word[u]))));
            const float pc = __uint_as_float(static_cast<uint32_t>(
              __builtin_amdgcn_ds_bpermute(src_c[u][h], static_cast<int>(word[u]))));
            wc[u][h] = live[u][h] ? pr * pc : 0.f;
          }
      } else {
        load_taps(k, tr, tc, live);
        load_meta(c + U, k, w);
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int u = 0; u < U; ++u)
#pragma unroll
          for (int h = 0; h < T; ++h) wc[u][h] = live[u][h] ? tr[u][h] * tc[u][h] : 0.f;
      }
      compute(c, b, wc);
    }
  } else if constexpr (P == 1) {
    // One batch of rows in flight: rows of batch c, then metadata of batch c + U.
    uint32_t k[U];
    float w[U][T];
    if (c < c1) load_meta(c, k, w);
    for (; c < c1; c += U) {
      RowPiece<NF> b[U];
      float wc[U][T];
      load_rows(k, b);
      copy_w(wc, w);
      load_meta(c + U, k, w);
      // Keep every load of the batch ahead of the products: the scheduler would
      // otherwise sink each row load to its first use and expose its full latency.
      __builtin_amdgcn_sched_barrier(0);
      compute(c, b, wc);
    }
  } else if (c < c1) {
    // Two batches of rows in flight: while batch c is summed, the rows of batch c + U
    // and the metadata of batch c + 2U are loading.  Two register sets, A and B,
    // alternate (written out twice so no in-flight register is ever copied).
    uint32_t kA[U], kB[U];
    float wA[U][T], wB[U][T], wc[U][T];
    RowPiece<NF> bA[U], bB[U];
    load_meta(c, kA, wA);
    load_rows(kA, bA);
    load_meta(c + U, kB, wB);
    while (true) {
      copy_w(wc, wA);
      load_rows(kB, bB);
      load_meta(c + 2 * U, kA, wA);
      __builtin_amdgcn_sched_barrier(0);
      compute(c, bA, wc);
      c += U;
      if (c >= c1) break;
      copy_w(wc, wB);
      load_rows(kA, bA);
      load_meta(c + 2 * U, kB, wB);
      __builtin_amdgcn_sched_barrier(0);
      compute(c, bB, wc);
      c += U;
      if (c >= c1) break;
    }
  }
  for (; t < nitems; ++t) flush(t);  // the last item with visits, then any empty ones
}

// Four waves per workgroup, one run per wave.  gather_map 0: runs in launch order, or
// with matrix_slab s > 1, in slabs of s consecutive workgroups dealt round-robin to the
// XCDs (slab i on XCD i % 8), so each slab's neighbouring blocks share one L2 while
// every XCD still sweeps the whole grid; 1 or 2: workgroup w (dispatched to XCD w % 8)
// takes the w/8-th group of an equal share of runs (runs are already of equal cost).
//
// Persistent form (matrix_persistent > 0): only as many workgroups as the CUs hold at
// once are launched, wg_per_xcd per XCD; XCD x takes the x-th eighth of the runs (a
// contiguous stretch of blocks, whose shared sample rows stay in that XCD's L2) and its
// waves stride through it.  Each run is computed exactly as in the one-wave-per-run
// form, so the results are the same bits.
template <int CP, int U, int P, int T, bool WT, bool TS>
__global__ __launch_bounds__(256) void gather_matrix_kernel(GatherArgs a) {
  const uint32_t wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  if (a.persist) {
    const uint32_t x = blockIdx.x % kXcds, l = blockIdx.x / kXcds;
    const uint32_t per = (a.run_count + kXcds - 1) / kXcds;
    const uint32_t r1 = min(a.run_count, (x + 1) * per);
    for (uint32_t run = x * per + 4 * l + wave; run < r1; run += 4 * a.wg_per_xcd) matrix_run<CP, U, P, T, WT, TS>(a, run);
    return;
  }
  uint32_t wg = blockIdx.x;
  if (a.xcd_chunk) {
    wg = (blockIdx.x % kXcds) * a.wg_per_xcd + blockIdx.x / kXcds;
  } else if (a.slab > 1) {
    const uint32_t x = blockIdx.x % kXcds, l = blockIdx.x / kXcds;
    wg = ((l / a.slab) * kXcds + x) * a.slab + l % a.slab;
  }
  const uint32_t run = wg * 4 + wave;
  if (run < a.run_count) matrix_run<CP, U, P, T, WT, TS>(a, run);
}

// ----------------------------------------------------------------------------- tile gather
// One workgroup per tile work item; a tile is 64 cells, TY = 64 / TX rows of TX
// columns.  The item's window samples are staged in LDS kTileChunk at a time:
// their sorted rows (coalesced float4 copies) and, per sample, the 64 weights
// phi(row tap) * phi(column tap) of the tile's cells (0 outside the sample's
// footprint).  Wave w owns tile rows [TY/4 w, TY/4 (w+1)) (16 cells, one register
// accumulator each) and sweeps the window rows that reach them: every sample of
// such a row updates all TX cells of each owned tile row it reaches, so a window
// row is one contiguous run of the window list with no per-cell bookkeeping.
// Narrow tiles waste fewer multiply-adds on zero weights (a sample reaches W of
// the TX + W - 1 window columns) at the price of a wider window halo.  Lanes are
// coils; with fewer than 64 coils the 64 / Cp lane groups take every (64 / Cp)-th
// sample of a run and are summed at the end in a fixed butterfly order, so
// results are deterministic.
// Window samples staged at once: 64, or 32 for 64 coils (so four workgroups fit a CU's LDS).
template <int CP>
constexpr int tile_chunk() {
  return CP >= 64 ? 32 : 64;
}
constexpr int kWgtStride = 68;  // floats per staged sample: 16-byte rows whose groups fall in different banks

struct TileArgs {
  const float2* sorted;        // [S][Cp]
  const uint32_t* cell_start;  // [nbins + 1]
  const float* taps;           // [S][2W]
  const uint4* item_info;      // {tile, first, end window-list index, slot or ~0}
  const uint32_t* tile_first_item;
  const uint32_t* tile_first_slot;
  const uint32_t* multi_tiles;
  float2* partial;             // [slots][64][Cp]
  float2* grid;                // [G_r][G_c][Cp]
  uint32_t grid_rows, grid_cols, tile_cols, tile_w, cp, multi_count;
};

__device__ inline uint32_t wrap_bin(int v, int n) { return static_cast<uint32_t>(((v % n) + n) % n); }

template <int W>
__device__ inline float pick(const float (&v)[W], int i) {
  float r = 0.f;
#pragma unroll
  for (int u = 0; u < W; ++u) r = i == u ? v[u] : r;
  return r;
}

template <int CP, int W, int TX>
__global__ __launch_bounds__(256) void tile_gather_kernel(TileArgs a) {
  constexpr int TY = static_cast<int>(kTileCells) / TX;
  constexpr int RW = TY / 4;                            // tile rows per wave
  constexpr int RY = TY + W - 1, RX = TX + W - 1;       // window rows, columns
  constexpr int NB = RY * RX;
  constexpr int G = 64 / CP;
  constexpr int kTileChunk = tile_chunk<CP>();
  constexpr int PIECES = CP / 2;                        // float4 per sample row
  constexpr int ROW_LOADS = kTileChunk * PIECES / 256;  // per thread and chunk
  constexpr int U = RW > 2 ? 2 : 4;                     // run samples per lane group loaded before use
  static_assert(NB < 256, "window bins are one per thread");
  static_assert(ROW_LOADS >= 1, "at least one row piece per thread");
  static_assert(TX % 4 == 0 && RW * TX == 16, "16 cells per wave, weights read as float4");
  __shared__ uint32_t pos[NB + 1];
  __shared__ uint32_t gst[NB];
  __shared__ uint32_t kidx[kTileChunk];
  __shared__ uint32_t kbin[kTileChunk];
  __shared__ uint32_t wave_sum[4];
  __shared__ __align__(16) float wgt[kTileChunk * kWgtStride];
  __shared__ float4 rows[kTileChunk * PIECES];

  const uint32_t t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const uint4 info = a.item_info[blockIdx.x];
  const int gr = static_cast<int>(a.grid_rows), gc = static_cast<int>(a.grid_cols);
  const int y0 = static_cast<int>(info.x / a.tile_cols) * TY, x0 = static_cast<int>(info.x % a.tile_cols) * TX;

  // Window bins: first sorted sample and count; exclusive scan of the counts.
  uint32_t cnt = 0;
  if (t < static_cast<uint32_t>(NB)) {
    const int r = static_cast<int>(t) / RX, c = static_cast<int>(t) % RX;
    const uint64_t b = static_cast<uint64_t>(wrap_bin(y0 - W / 2 + r, gr)) * a.grid_cols + wrap_bin(x0 - W / 2 + c, gc);
    const uint32_t s0 = a.cell_start[b];
    cnt = a.cell_start[b + 1] - s0;
    gst[t] = s0;
  }
  uint32_t inc = cnt;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const uint32_t v = __shfl_up(inc, d, 64);
    if (lane >= static_cast<uint32_t>(d)) inc += v;
  }
  if (lane == 63) wave_sum[wave] = inc;
  __syncthreads();
  uint32_t off = 0;
  for (uint32_t w = 0; w < wave; ++w) off += wave_sum[w];
  const uint32_t my_pos = off + inc - cnt;  // this thread's bin's first window-list index
  if (t <= static_cast<uint32_t>(NB)) pos[t] = my_pos;  // pos[NB] = window total
  __syncthreads();

  const uint32_t end = info.z < pos[NB] ? info.z : pos[NB];
  const uint32_t coil = lane % CP, g = lane / CP;
  const int dyw = RW * static_cast<int>(wave);
  float2 acc[16];
#pragma unroll
  for (int q = 0; q < 16; ++q) acc[q] = make_float2(0.f, 0.f);

  for (uint32_t c0 = info.y; c0 < end; c0 += kTileChunk) {
    const uint32_t n = end - c0 < static_cast<uint32_t>(kTileChunk) ? end - c0 : kTileChunk;
    // Each window bin writes the sorted positions of its samples in this chunk.
    if (t < static_cast<uint32_t>(NB) && cnt > 0) {
      const uint32_t lo = my_pos > c0 ? my_pos : c0, hi = my_pos + cnt < c0 + n ? my_pos + cnt : c0 + n;
      for (uint32_t j = lo; j < hi; ++j) {
        kidx[j - c0] = gst[t] + (j - my_pos);
        kbin[j - c0] = t;
      }
    }
    __syncthreads();
    // Rows: all float4 loads in flight before the LDS writes.
    float4 v[ROW_LOADS];
#pragma unroll
    for (int i = 0; i < ROW_LOADS; ++i) {
      const uint32_t f = t + 256u * i, s = f / PIECES;
      const uint32_t k = kidx[s < n ? s : 0];  // slots past the chunk copy its sample 0: never summed
      v[i] = reinterpret_cast<const float4*>(a.sorted)[static_cast<uint64_t>(k) * PIECES + f % PIECES];
    }
    if (lane < n) {  // weights of sample `lane` (< kTileChunk) for this wave's tile rows
      const uint32_t b = kbin[lane];
      const int r = static_cast<int>(b) / RX, c = static_cast<int>(b) % RX;
      const float* tp = a.taps + static_cast<uint64_t>(kidx[lane]) * (2 * W);
      float ty[W], tx[W];
#pragma unroll
      for (int u = 0; u < W; ++u) {
        ty[u] = tp[u];
        tx[u] = tp[W + u];
      }
      float wx[TX];
#pragma unroll
      for (int dx = 0; dx < TX; ++dx) wx[dx] = pick<W>(tx, dx - c + W - 1);
#pragma unroll
      for (int e = 0; e < RW; ++e) {
        const float wy = pick<W>(ty, dyw + e - r + W - 1);
        float4* dst = reinterpret_cast<float4*>(wgt + lane * kWgtStride + (dyw + e) * TX);
#pragma unroll
        for (int x4 = 0; x4 < TX / 4; ++x4)
          dst[x4] = make_float4(wy * wx[4 * x4], wy * wx[4 * x4 + 1], wy * wx[4 * x4 + 2], wy * wx[4 * x4 + 3]);
      }
    }
#pragma unroll
    for (int i = 0; i < ROW_LOADS; ++i) rows[t + 256u * i] = v[i];
    __syncthreads();
    const float2* lrows = reinterpret_cast<const float2*>(rows);
    // Window row dyw + i reaches owned tile row dyw + e when 0 <= i - e < W.
#pragma unroll
    for (int i = 0; i < RW + W - 1; ++i) {
      const int r = dyw + i;
      uint32_t jb = __builtin_amdgcn_readfirstlane(pos[r * RX]), je = __builtin_amdgcn_readfirstlane(pos[r * RX + RX]);
      jb = jb > c0 ? jb - c0 : 0u;
      je = je > c0 ? je - c0 : 0u;
      je = je < n ? je : n;
      for (uint32_t j = jb; j < je; j += U * G) {
        float2 d[U];
        float4 w[U][RW][TX / 4];
#pragma unroll
        for (int u = 0; u < U; ++u) {
          const uint32_t l = j + u * G + g;
          const uint32_t lc = l < je ? l : je - 1;
          d[u] = lrows[lc * CP + coil];
          const float4* wr = reinterpret_cast<const float4*>(wgt + lc * kWgtStride + dyw * TX);
#pragma unroll
          for (int e = 0; e < RW; ++e)
            if (i - e >= 0 && i - e < W)
#pragma unroll
              for (int x4 = 0; x4 < TX / 4; ++x4) w[u][e][x4] = wr[e * (TX / 4) + x4];
          d[u] = l < je ? d[u] : make_float2(0.f, 0.f);  // a slot past the run adds zero
        }
#pragma unroll
        for (int u = 0; u < U; ++u)
#pragma unroll
          for (int e = 0; e < RW; ++e)
            if (i - e >= 0 && i - e < W)
#pragma unroll
              for (int x4 = 0; x4 < TX / 4; ++x4) {
                const float wv[4] = {w[u][e][x4].x, w[u][e][x4].y, w[u][e][x4].z, w[u][e][x4].w};
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                  float2& acc_c = acc[e * TX + 4 * x4 + k];
                  acc_c.x += wv[k] * d[u].x;
                  acc_c.y += wv[k] * d[u].y;
                }
              }
      }
    }
    __syncthreads();
  }
  if constexpr (G > 1) {
#pragma unroll
    for (int m = CP; m < 64; m <<= 1)
#pragma unroll
      for (int q = 0; q < 16; ++q) {
        acc[q].x += __shfl_xor(acc[q].x, m, 64);
        acc[q].y += __shfl_xor(acc[q].y, m, 64);
      }
  }
  if (g != 0) return;
#pragma unroll
  for (int q = 0; q < 16; ++q) {
    const int dy = dyw + q / TX, dx = q % TX;
    if (info.w == 0xFFFFFFFFu) {
      const int y = y0 + dy, x = x0 + dx;
      if (y < gr && x < gc) a.grid[(static_cast<uint64_t>(y) * a.grid_cols + x) * CP + coil] = acc[q];
    } else {
      a.partial[(static_cast<uint64_t>(info.w) * kTileCells + dy * TX + dx) * CP + coil] = acc[q];
    }
  }
}

// Tiles with several items: thread per (tile, cell, coil) sums the item partials in item order.
__global__ __launch_bounds__(256) void tile_reduce_kernel(TileArgs a) {
  const uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= static_cast<uint64_t>(a.multi_count) * kTileCells * a.cp) return;
  const uint32_t c = static_cast<uint32_t>(i % a.cp);
  const uint32_t cell = static_cast<uint32_t>((i / a.cp) % kTileCells);
  const uint32_t tile = a.multi_tiles[i / (static_cast<uint64_t>(kTileCells) * a.cp)];
  const uint32_t s0 = a.tile_first_slot[tile];
  const uint32_t n = a.tile_first_item[tile + 1] - a.tile_first_item[tile];
  const float2* p = a.partial + (static_cast<uint64_t>(s0) * kTileCells + cell) * a.cp + c;
  const uint64_t stride = static_cast<uint64_t>(kTileCells) * a.cp;
  float2 s = make_float2(0.f, 0.f);
  uint32_t j = 0;
  for (; j + 8 <= n; j += 8) {
    float2 t[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) t[u] = p[(j + u) * stride];
#pragma unroll
    for (int u = 0; u < 8; ++u) s = cadd(s, t[u]);
  }
  for (; j < n; ++j) s = cadd(s, p[j * stride]);
  const uint32_t th = kTileCells / a.tile_w;
  const uint32_t y = (tile / a.tile_cols) * th + cell / a.tile_w;
  const uint32_t x = (tile % a.tile_cols) * a.tile_w + cell % a.tile_w;
  if (y < a.grid_rows && x < a.grid_cols) a.grid[(static_cast<uint64_t>(y) * a.grid_cols + x) * a.cp + c] = s;
}

TileArgs tile_args(const Plan2& plan, const float2* d_sorted, uint32_t coils, float2* d_grid, float2* d_partial) {
  TileArgs a{};
  a.sorted = d_sorted;
  a.cell_start = plan.cell_start;
  a.taps = plan.taps;
  a.item_info = plan.item_info;
  a.tile_first_item = plan.block_first_item;
  a.tile_first_slot = plan.block_first_slot;
  a.multi_tiles = plan.multi_blocks;
  a.partial = d_partial;
  a.grid = d_grid;
  a.grid_rows = plan.d.grid_rows;
  a.grid_cols = plan.d.grid_cols;
  a.tile_cols = plan.tile_cols;
  a.tile_w = plan.tile_width;
  a.cp = padded_coils(coils);
  a.multi_count = plan.multi_count;
  return a;
}

template <int CP, int TX>
void launch_tile_gather(const TileArgs& a, uint32_t width, uint32_t items, hipStream_t stream) {
  const dim3 grid(items), block(256);
  switch (width) {
    case 2: hipLaunchKernelGGL((tile_gather_kernel<CP, 2, TX>), grid, block, 0, stream, a); break;
    case 4: hipLaunchKernelGGL((tile_gather_kernel<CP, 4, TX>), grid, block, 0, stream, a); break;
    case 6: hipLaunchKernelGGL((tile_gather_kernel<CP, 6, TX>), grid, block, 0, stream, a); break;
    case 8: hipLaunchKernelGGL((tile_gather_kernel<CP, 8, TX>), grid, block, 0, stream, a); break;
    default: throw Error(1, "kernel width must be 2, 4, 6 or 8");
  }
  KSJ_LAUNCH_CHECK();
}

template <int CP>
void launch_tile_gather(const TileArgs& a, uint32_t width, uint32_t items, hipStream_t stream) {
  if (a.tile_w == 4) return launch_tile_gather<CP, 4>(a, width, items, stream);
  return launch_tile_gather<CP, 8>(a, width, items, stream);
}

// ----------------------------------------------------------------------------- fused items
// One launch per stage and padded coil count over the items of an uneven batch.
struct PrepItemArgs {
  StridedPlanes planes;
  const float2* weight;
  const uint32_t* rank;
  float2* sorted;
  uint32_t samples, coils;
};

template <int CP>
__global__ __launch_bounds__(kPrepSamples * kPrepWaves) void prep_fused_kernel(const FusedTable<PrepItemArgs>* t) {
  const uint32_t item = fused_item(t, blockIdx.x);
  const PrepItemArgs a = fused_args(t, item);
  prep_block<CP>(a.planes, a.weight, a.rank, a.sorted, a.samples, a.coils,
                 static_cast<uint32_t>(blockIdx.x - fused_first(t, item)));
}

template <int SW, int B, int CPL>
__global__ __launch_bounds__(256) void gather_fused_kernel(const FusedTable<GatherArgs>* t) {
  constexpr int NSW = 64 / SW;
  const int lane = threadIdx.x & 63, sw = lane / SW, c = lane % SW;
  const uint32_t it = fused_item(t, blockIdx.x);
  const GatherArgs a = fused_args(t, it);
  const uint32_t wg = static_cast<uint32_t>(blockIdx.x - fused_first(t, it));
  const uint32_t item = (wg * 4 + (threadIdx.x >> 6)) * NSW + sw;
  gather_one_item<SW, B, CPL>(a, item, item < a.item_count, sw, c);
}

template <int CP, int U, int P, int T, bool WT, bool TS>
__global__ __launch_bounds__(256) void gather_matrix_fused_kernel(const FusedTable<GatherArgs>* t) {
  const uint32_t it = fused_item(t, blockIdx.x);
  const GatherArgs a = fused_args(t, it);
  const uint32_t wg = static_cast<uint32_t>(blockIdx.x - fused_first(t, it));
  const uint32_t run = wg * 4 + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  if (run < a.run_count) matrix_run<CP, U, P, T, WT, TS>(a, run);
}

__global__ __launch_bounds__(256) void reduce_fused_kernel(const FusedTable<GatherArgs>* t) {
  const uint32_t item = fused_item(t, blockIdx.x);
  const GatherArgs a = fused_args(t, item);
  multi_reduce_one(a, (blockIdx.x - fused_first(t, item)) * 256ull + threadIdx.x);
}

template <int SW, int CPL>
void launch_gather_fused(const FusedTableBuilder<GatherArgs>& b, void* d_table, hipStream_t stream) {
  b.upload(d_table, stream);
  const auto* t = static_cast<const FusedTable<GatherArgs>*>(d_table);
  const dim3 grid(static_cast<uint32_t>(b.blocks())), block(256);
  switch (tune_get(kTuneGatherBatch)) {
    case 4: hipLaunchKernelGGL((gather_fused_kernel<SW, 4, CPL>), grid, block, 0, stream, t); break;
    case 16: hipLaunchKernelGGL((gather_fused_kernel<SW, 16, CPL>), grid, block, 0, stream, t); break;
    default: hipLaunchKernelGGL((gather_fused_kernel<SW, 8, CPL>), grid, block, 0, stream, t); break;
  }
  KSJ_LAUNCH_CHECK();
}

template <int CP>
void launch_gather_matrix(GatherArgs a, hipStream_t stream) {
  uint32_t wgs = ceil_div(a.run_count, 4);
  a.xcd_chunk = static_cast<uint32_t>(tune_get(kTuneGatherMap));
  a.slab = static_cast<uint32_t>(tune_get(kTuneMatrixSlab));
  a.persist = 0;
  const uint32_t per_cu = static_cast<uint32_t>(tune_get(kTuneMatrixPersistent));
  if (per_cu > 0 && wgs > per_cu * device_cu_count()) {
    a.persist = 1;
    a.wg_per_xcd = per_cu * ceil_div(device_cu_count(), kXcds);
    wgs = a.wg_per_xcd * kXcds;
  } else if (a.xcd_chunk) {
    a.wg_per_xcd = ceil_div(wgs, kXcds);
    wgs = a.wg_per_xcd * kXcds;  // the padding workgroups find no run and exit
  } else if (a.slab > 1) {
    wgs = ceil_div(wgs, kXcds * a.slab) * kXcds * a.slab;
  }
  a.nt_store = static_cast<uint32_t>(tune_get(kTuneMatrixNtStore));
  with_matrix_shape(a.bw, a.from_taps != 0, [&]<int U, int P, int T, bool WT, bool TS>() {
    hipLaunchKernelGGL((gather_matrix_kernel<CP, U, P, T, WT, TS>), dim3(wgs), dim3(256), 0, stream, a);
  });
  KSJ_LAUNCH_CHECK();
}

// Calls f.template operator()<U, P, T, WT, TS>() for the tuned batch (U groups of four
// visits), pipeline depth, the plans' block width (T products of 16 cells) and visit
// form (WT: packed positions, weights from taps; its loop keeps one batch in flight; TS:
// each lane loads one tap word and takes its two taps from the visit's lanes).
template <typename F>
void with_matrix_shape(uint32_t block_w, bool from_taps, F&& f) {
  const int u = static_cast<int>(tune_get(kTuneMatrixBatch)) / 4;
  const bool two = tune_get(kTuneMatrixPipeline) == 2;
  const bool wide = block_w == 8;
  const bool share = tune_get(kTuneMatrixTapShare) == 1;
  auto by_p = [&]<int U>() {
    if (from_taps && share) {
      if (wide) f.template operator()<U, 1, 2, true, true>(); else f.template operator()<U, 1, 1, true, true>();
    } else if (from_taps) {
      if (wide) f.template operator()<U, 1, 2, true, false>(); else f.template operator()<U, 1, 1, true, false>();
    } else if (two) {
      if (wide) f.template operator()<U, 2, 2, false, false>(); else f.template operator()<U, 2, 1, false, false>();
    } else {
      if (wide) f.template operator()<U, 1, 2, false, false>(); else f.template operator()<U, 1, 1, false, false>();
    }
  };
  if (u == 1) by_p.template operator()<1>();
  else if (u == 4) by_p.template operator()<4>();
  else by_p.template operator()<2>();
}

template <int CP>
void launch_gather_matrix_fused(const FusedTableBuilder<GatherArgs>& b, uint32_t block_w, bool from_taps,
                                void* d_table, hipStream_t stream) {
  b.upload(d_table, stream);
  const auto* t = static_cast<const FusedTable<GatherArgs>*>(d_table);
  const dim3 grid(static_cast<uint32_t>(b.blocks())), block(256);
  with_matrix_shape(block_w, from_taps, [&]<int U, int P, int T, bool WT, bool TS>() {
    hipLaunchKernelGGL((gather_matrix_fused_kernel<CP, U, P, T, WT, TS>), grid, block, 0, stream, t);
  });
  KSJ_LAUNCH_CHECK();
}

template <int SW, int CPL>
void launch_gather(GatherArgs a, hipStream_t stream) {
  constexpr int NSW = 64 / SW;
  const uint32_t waves = ceil_div(a.item_count, NSW);
  uint32_t wgs = ceil_div(waves, 4);
  a.xcd_chunk = static_cast<uint32_t>(tune_get(kTuneGatherMap));
  if (a.xcd_chunk) {
    a.wg_per_xcd = ceil_div(wgs, kXcds);
    wgs = a.wg_per_xcd * kXcds;  // the padding workgroups find no item and exit
  }
  switch (tune_get(kTuneGatherBatch)) {
    case 4: hipLaunchKernelGGL((gather_item_kernel<SW, 4, CPL>), dim3(wgs), dim3(256), 0, stream, a); break;
    case 16: hipLaunchKernelGGL((gather_item_kernel<SW, 16, CPL>), dim3(wgs), dim3(256), 0, stream, a); break;
    default: hipLaunchKernelGGL((gather_item_kernel<SW, 8, CPL>), dim3(wgs), dim3(256), 0, stream, a); break;
  }
  KSJ_LAUNCH_CHECK();
}

// ----------------------------------------------------------------------------- chain (work queue)
// The gridding stages of an uneven batch -- prep, gather and reduce of every item, all
// padded coil counts -- as one persistent launch fed by a device work queue.
//  * Items are assigned to XCDs, largest first onto the least loaded one, and each XCD
//    has its own queue of tasks: prep (kChainPrepBlocks blocks of 64 samples), gather
//    (4 kChainRunsPerWave runs, kChainRunsPerWave per wave) and reduce
//    (kChainReduceIters x 256 partial sums).  Queues interleave items as prep k,
//    gather k - 1, reduce k - 2, so most dependencies are met before a task is taken.
//  * A workgroup reads its XCD from the hardware (XCC_ID) and takes the next task of
//    that XCD's queue with one atomic.  A gather task waits until every prep task of its
//    item has finished, a reduce task until every gather task has.
//  * All tasks of an item run on one XCD, so its sorted rows and partial sums pass
//    through that XCD's L2 only: a finished task waits for its own stores (s_waitcnt)
//    and then counts itself done with an atomic; no L2 write-back or invalidate.
//  * Tasks are taken in queue order and wait only on earlier tasks of the same queue,
//    which running workgroups hold, so every wait ends.  Each wait is still bounded:
//    past `spin_limit` sleeps the workgroup marks the launch faulty (every workgroup then
//    stops taking tasks) and raises the fault counter, so a broken queue ends the kernel
//    with wrong results instead of hanging it.
// Every task computes exactly what the per-stage kernels compute (same blocks, runs and
// reduce order), so results are the same bits as the fused and per-item forms.
constexpr uint32_t kChainPrepBlocks = 8;
constexpr uint32_t kChainRunsPerWave = 2;
constexpr uint32_t kChainReduceIters = 4;
constexpr uint32_t kChainMaxSegs = 3 * kMaxFusedItems;
constexpr uint32_t kChainSpinLimit = 1u << 22;  // sleeps of 8 x 64 cycles: about a second
enum : uint32_t { kTaskPrep = 0, kTaskGather = 1, kTaskReduce = 2 };

struct ChainItem {
  PrepItemArgs prep;
  GatherArgs gather;
  uint32_t cp, prep_tasks, gather_tasks, reduce_tasks;
};

struct ChainTable {
  uint32_t queues;                         // XCD queues in use (1..8)
  uint32_t pad;
  uint32_t seg_begin[kXcds + 1];           // queue x owns segments [seg_begin[x], seg_begin[x + 1])
  uint32_t task_count[kXcds];              // tasks of each queue
  uint32_t seg_first[kChainMaxSegs];       // first task of each segment (within its queue)
  uint32_t seg_info[kChainMaxSegs];        // type << 16 | item
  ChainItem items[kMaxFusedItems];         // last: only the used entries are uploaded
};

struct ChainState {
  uint32_t head[kXcds * 32];  // next task of each queue, one 128-byte line apiece
  uint32_t prep_done[kMaxFusedItems];
  uint32_t gather_done[kMaxFusedItems];
  uint32_t fault;
};

// Launches whose waits ran out (read by ksj_gpu_chain_faults).
__device__ unsigned int g_chain_faults;

__device__ inline uint32_t hw_xcc_id() {
  return static_cast<uint32_t>(__builtin_amdgcn_s_getreg((3 << 11) | (0 << 6) | 20));  // HW_REG_XCC_ID, bits 0-3
}

__device__ inline uint32_t cword(const uint32_t* p) { return *(ConstWord)(p); }

template <typename A>
__device__ inline A const_load(const A* p) {
  static_assert(sizeof(A) % 4 == 0, "table entries are read as 32-bit words");
  const ConstWord src = (ConstWord)(p);
  uint32_t w[sizeof(A) / 4];
#pragma unroll
  for (uint32_t i = 0; i < sizeof(A) / 4; ++i) w[i] = src[i];
  A a;
  __builtin_memcpy(&a, w, sizeof(A));
  return a;
}

__global__ void chain_reset_kernel(ChainState* st) {
  uint32_t* w = reinterpret_cast<uint32_t*>(st);
  for (uint32_t i = threadIdx.x; i < sizeof(ChainState) / 4; i += blockDim.x) w[i] = 0;
}

__global__ void xcc_probe_kernel(uint32_t* out) {
  if (threadIdx.x == 0) out[blockIdx.x] = hw_xcc_id();
}

template <int CPMAX, int U, int T, bool TS>
__global__ __launch_bounds__(256) void grid_chain_kernel(const ChainTable* tab, ChainState* st, uint32_t spin_limit) {
  extern __shared__ float2 chain_lds[];
  __shared__ uint32_t s_task, s_ok;
  const uint32_t nq = cword(&tab->queues);
  const uint32_t q = hw_xcc_id() % nq;
  const uint32_t sb = cword(&tab->seg_begin[q]), se = cword(&tab->seg_begin[q + 1]);
  const uint32_t ntasks = cword(&tab->task_count[q]);
  const uint32_t wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  for (;;) {
    if (threadIdx.x == 0) {
      uint32_t t = ~0u;
      if (__hip_atomic_load(&st->fault, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0)
        t = __hip_atomic_fetch_add(&st->head[32 * q], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      s_task = t;
    }
    __syncthreads();
    const uint32_t task = __builtin_amdgcn_readfirstlane(s_task);
    if (task >= ntasks) break;
    uint32_t lo = sb, hi = se;  // the segment holding the task: seg_first[lo] <= task < seg_first[lo + 1]
    while (hi - lo > 1) {
      const uint32_t mid = (lo + hi) >> 1;
      if (cword(&tab->seg_first[mid]) <= task)
        lo = mid;
      else
        hi = mid;
    }
    const uint32_t info = cword(&tab->seg_info[lo]);
    const uint32_t type = info >> 16, item = info & 0xFFFFu;
    const uint32_t idx = task - cword(&tab->seg_first[lo]);
    const ChainItem* it = &tab->items[item];
    if (type != kTaskPrep) {
      if (threadIdx.x == 0) {
        const uint32_t* ctr = type == kTaskGather ? &st->prep_done[item] : &st->gather_done[item];
        const uint32_t need = type == kTaskGather ? cword(&it->prep_tasks) : cword(&it->gather_tasks);
        uint32_t n = 0;
        bool ok = true, timed_out = false;
        while (__hip_atomic_load(ctr, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < need) {
          if (__hip_atomic_load(&st->fault, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) {
            ok = false;  // another workgroup's wait ran out: stop
            break;
          }
          if (++n > spin_limit) {
            ok = false;
            timed_out = true;
            break;
          }
          __builtin_amdgcn_s_sleep(8);
        }
        if (timed_out && __hip_atomic_exchange(&st->fault, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0)
          __hip_atomic_fetch_add(&g_chain_faults, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        s_ok = ok ? 1u : 0u;
      }
      __syncthreads();
      if (__builtin_amdgcn_readfirstlane(s_ok) == 0) break;
    }
    const uint32_t cp = cword(&it->cp);
    if (type == kTaskPrep) {
      const PrepItemArgs a = const_load(&it->prep);
      const uint32_t nb = (a.samples + kPrepSamples - 1) / kPrepSamples;
      const uint32_t b0 = idx * kChainPrepBlocks, b1 = min(b0 + kChainPrepBlocks, nb);
      auto* rank_s = reinterpret_cast<uint32_t*>(chain_lds + kPrepSamples * (cp + 1));
      for (uint32_t b = b0; b < b1; ++b) {
        if (cp == 8)
          prep_block_in<8>(a.planes, a.weight, a.rank, a.sorted, a.samples, a.coils, b, chain_lds, rank_s);
        else if (cp == 16)
          prep_block_in<16>(a.planes, a.weight, a.rank, a.sorted, a.samples, a.coils, b, chain_lds, rank_s);
        else if constexpr (CPMAX >= 32) {
          if (cp == 32)
            prep_block_in<32>(a.planes, a.weight, a.rank, a.sorted, a.samples, a.coils, b, chain_lds, rank_s);
          else if constexpr (CPMAX >= 64)
            prep_block_in<64>(a.planes, a.weight, a.rank, a.sorted, a.samples, a.coils, b, chain_lds, rank_s);
        }
        __syncthreads();  // the tile is read out before the next block writes it
      }
    } else if (type == kTaskGather) {
      const GatherArgs a = const_load(&it->gather);
#pragma unroll 1
      for (uint32_t r = 0; r < kChainRunsPerWave; ++r) {
        const uint32_t run = (idx * kChainRunsPerWave + r) * 4 + wave;
        if (run >= a.run_count) break;
        if (cp == 8)
          matrix_run<8, U, 1, T, true, TS>(a, run);
        else if (cp == 16)
          matrix_run<16, U, 1, T, true, TS>(a, run);
        else if constexpr (CPMAX >= 32) {
          if (cp == 32)
            matrix_run<32, U, 1, T, true, TS>(a, run);
          else if constexpr (CPMAX >= 64)
            matrix_run<64, U, 1, T, true, TS>(a, run);
        }
      }
    } else {
      const GatherArgs a = const_load(&it->gather);
#pragma unroll 1
      for (uint32_t i = 0; i < kChainReduceIters; ++i)
        multi_reduce_one(a, (static_cast<uint64_t>(idx) * kChainReduceIters + i) * 256u + threadIdx.x);
    }
    // Done: this workgroup's stores have reached the XCD's L2, then one count.
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    if (threadIdx.x == 0 && type != kTaskReduce)
      __hip_atomic_fetch_add(type == kTaskPrep ? &st->prep_done[item] : &st->gather_done[item], 1u, __ATOMIC_RELAXED,
                             __HIP_MEMORY_SCOPE_AGENT);
  }
}

}  // namespace

uint32_t padded_coils(uint32_t coils) {
  uint32_t cp = 8;
  while (cp < coils) cp *= 2;
  return cp;
}

size_t grid_partial_bytes(const Plan2& plan, uint32_t coils) {
  const size_t cells = plan.scheme == kSchemeTiles ? kTileCells : static_cast<size_t>(plan.block_h) * plan.block_w;
  return static_cast<size_t>(plan.slot_count) * cells * padded_coils(coils) * sizeof(float2);
}

void grid_prep(const Plan2& plan, const CoilPlanes& planes, float2* d_sorted, hipStream_t stream) {
  require(plan.has_trajectory, "plan has no trajectory");
  grid_prep_rows(plan.weight, plan.rank, plan.d.samples, planes, d_sorted, stream);
}

void grid_prep_rows(const float2* d_weight, const uint32_t* d_rank, uint32_t samples, const CoilPlanes& planes,
                    float2* d_sorted, hipStream_t stream) {
  require(planes.coils >= 1 && planes.coils <= kMaxStackedCoils, "gridding takes 1..64 stacked coils");
  PrepArgs a{};
  a.planes = planes;
  a.weight = d_weight;
  a.rank = d_rank;
  a.sorted = d_sorted;
  a.samples = samples;
  a.coils = planes.coils;
  const dim3 grid(ceil_div(a.samples, kPrepSamples)), block(kPrepSamples * kPrepWaves);
  switch (padded_coils(planes.coils)) {
    case 8: hipLaunchKernelGGL(prep_kernel<8>, grid, block, 0, stream, a); break;
    case 16: hipLaunchKernelGGL(prep_kernel<16>, grid, block, 0, stream, a); break;
    case 32: hipLaunchKernelGGL(prep_kernel<32>, grid, block, 0, stream, a); break;
    default: hipLaunchKernelGGL(prep_kernel<64>, grid, block, 0, stream, a); break;
  }
  KSJ_LAUNCH_CHECK();
}

namespace {
GatherArgs gather_args(const Plan2& plan, const float2* d_sorted, uint32_t coils, float2* d_grid,
                       float2* d_partial) {
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
  a.block_group = plan.block_group;
  a.cp = padded_coils(coils);
  a.item_count = plan.item_count;
  a.multi_count = plan.multi_count;
  a.xcd_first = plan.xcd_first;
  a.bh = plan.block_h;
  a.bw = plan.block_w;
  a.run_first = plan.run_first;
  a.run_count = plan.run_count;
  a.taps = plan.taps;
  a.width = plan.d.width;
  a.from_taps = plan.packed_visits ? 1u : 0u;
  a.zero_cols = tune_get(kTuneGridPrune) == 1 ? plan.zero_cols : nullptr;
  return a;
}
}  // namespace

void grid_reduce(const Plan2& plan, uint32_t coils, float2* d_grid, const float2* d_partial, hipStream_t stream) {
  if (plan.multi_count == 0) return;
  if (plan.scheme == kSchemeTiles) {
    const TileArgs t = tile_args(plan, nullptr, coils, d_grid, const_cast<float2*>(d_partial));
    hipLaunchKernelGGL(tile_reduce_kernel,
                       dim3(ceil_div(static_cast<uint64_t>(t.multi_count) * kTileCells * t.cp, 256)), dim3(256), 0,
                       stream, t);
    KSJ_LAUNCH_CHECK();
    return;
  }
  const GatherArgs a = gather_args(plan, nullptr, coils, d_grid, const_cast<float2*>(d_partial));
  hipLaunchKernelGGL(multi_reduce_kernel,
                     dim3(ceil_div(static_cast<uint64_t>(a.multi_count) * a.bh * a.bw * a.cp, 256)), dim3(256),
                     0, stream, a);
  KSJ_LAUNCH_CHECK();
}

void grid_gather(const Plan2& plan, const float2* d_sorted, uint32_t coils, float2* d_grid, float2* d_partial,
                 hipStream_t stream) {
  require(plan.has_trajectory, "plan has no trajectory");
  require(plan.multi_count == 0 || d_partial != nullptr, "multi-item blocks need a partial-sum buffer");
  if (plan.scheme == kSchemeTiles) {
    const TileArgs t = tile_args(plan, d_sorted, coils, d_grid, d_partial);
    switch (t.cp) {
      case 8: return launch_tile_gather<8>(t, plan.d.width, plan.item_count, stream);
      case 16: return launch_tile_gather<16>(t, plan.d.width, plan.item_count, stream);
      case 32: return launch_tile_gather<32>(t, plan.d.width, plan.item_count, stream);
      case 64: return launch_tile_gather<64>(t, plan.d.width, plan.item_count, stream);
      default: throw Error(1, "gridding supports at most 64 coils");
    }
  }
  const GatherArgs a = gather_args(plan, d_sorted, coils, d_grid, d_partial);
  if (plan.scheme == kSchemeMatrix) {
    switch (a.cp) {
      case 8: return launch_gather_matrix<8>(a, stream);
      case 16: return launch_gather_matrix<16>(a, stream);
      case 32: return launch_gather_matrix<32>(a, stream);
      case 64: return launch_gather_matrix<64>(a, stream);
      default: throw Error(1, "gridding supports at most 64 coils");
    }
  }
  // Two coils per lane from 32 coils up (sub-waves of 16 or 32 lanes keep the DPP
  // row broadcasts); one coil per lane below (tuning key gather_coils_per_lane).
  const bool two = tune_get(kTuneGatherCoilsPerLane) == 2;
  switch (a.cp) {
    case 8: return launch_gather<8, 1>(a, stream);
    case 16: return launch_gather<16, 1>(a, stream);
    case 32: return two ? launch_gather<16, 2>(a, stream) : launch_gather<32, 1>(a, stream);
    case 64: return two ? launch_gather<32, 2>(a, stream) : launch_gather<64, 1>(a, stream);
    default: throw Error(1, "gridding supports at most 64 coils");
  }
}

// Tables: prep per padded coil count, gather per padded coil count and schedule
// (2x2 blocks, matrix), one reduce.
size_t grid_fused_table_bytes() {
  return 4 * fused_table_bytes<PrepItemArgs>() + 9 * fused_table_bytes<GatherArgs>();
}

void grid_fused(const GridFusedItem* items, uint32_t count, void* d_tables, hipStream_t stream) {
  require(count >= 1 && count <= kMaxFusedItems, "fused gridding takes 1..64 items");
  const bool two = tune_get(kTuneGatherCoilsPerLane) == 2;
  FusedTableBuilder<PrepItemArgs> prep[4];
  FusedTableBuilder<GatherArgs> gather[4], matrix[4], reduce;
  uint32_t matrix_w = 0;  // block width of the batch's matrix plans (one width per fused call)
  int matrix_taps = -1;   // visit form of the batch's matrix plans (one form per fused call)
  const uint32_t nt_store = static_cast<uint32_t>(tune_get(kTuneMatrixNtStore));
  for (uint32_t i = 0; i < count; ++i) {
    const GridFusedItem& it = items[i];
    require(it.plan != nullptr && it.plan->has_trajectory, "fused item without a ready plan");
    require(it.plan->scheme == kSchemeBlocks || it.plan->scheme == kSchemeMatrix,
            "fused gridding runs the 2x2-block or matrix schedule");
    require(it.channels >= 1 && it.channels <= kMaxStackedCoils, "fused items take 1..64 channels");
    require(it.plan->multi_count == 0 || it.partial != nullptr, "multi-item blocks need a partial-sum buffer");
    const uint32_t cp = padded_coils(it.channels);
    const int cls = cp == 8 ? 0 : cp == 16 ? 1 : cp == 32 ? 2 : 3;
    PrepItemArgs pa{};
    pa.planes = StridedPlanes{it.kspace, it.plan->d.samples};
    pa.weight = it.plan->weight;
    pa.rank = it.plan->rank;
    pa.sorted = it.sorted;
    pa.samples = it.plan->d.samples;
    pa.coils = it.channels;
    prep[cls].add(pa, ceil_div(pa.samples, kPrepSamples));
    GatherArgs ga = gather_args(*it.plan, it.sorted, it.channels, it.grid, it.partial);
    ga.xcd_chunk = 0;
    if (it.plan->scheme == kSchemeMatrix) {
      require(matrix_w == 0 || matrix_w == it.plan->block_w, "fused matrix items share one block width");
      require(matrix_taps < 0 || matrix_taps == static_cast<int>(ga.from_taps), "fused matrix items share one visit form");
      matrix_w = it.plan->block_w;
      matrix_taps = static_cast<int>(ga.from_taps);
      ga.nt_store = nt_store;
      matrix[cls].add(ga, std::max<uint64_t>(1, ceil_div(ga.run_count, 4)));
    } else {
      const uint32_t sw = (two && cp >= 32) ? cp / 2 : cp;
      gather[cls].add(ga, std::max<uint64_t>(1, ceil_div(ceil_div(ga.item_count, 64 / sw), 4)));
    }
    if (ga.multi_count > 0)
      reduce.add(ga, ceil_div(static_cast<uint64_t>(ga.multi_count) * ga.bh * ga.bw * cp, 256));
  }
  char* tab = static_cast<char*>(d_tables);
  const dim3 prep_block(kPrepSamples * kPrepWaves);
  for (int cls = 0; cls < 4; ++cls) {
    void* d = tab + cls * fused_table_bytes<PrepItemArgs>();
    if (prep[cls].t.count == 0) continue;
    prep[cls].upload(d, stream);
    const auto* t = static_cast<const FusedTable<PrepItemArgs>*>(d);
    const dim3 grid(static_cast<uint32_t>(prep[cls].blocks()));
    switch (cls) {
      case 0: hipLaunchKernelGGL(prep_fused_kernel<8>, grid, prep_block, 0, stream, t); break;
      case 1: hipLaunchKernelGGL(prep_fused_kernel<16>, grid, prep_block, 0, stream, t); break;
      case 2: hipLaunchKernelGGL(prep_fused_kernel<32>, grid, prep_block, 0, stream, t); break;
      default: hipLaunchKernelGGL(prep_fused_kernel<64>, grid, prep_block, 0, stream, t); break;
    }
    KSJ_LAUNCH_CHECK();
  }
  tab += 4 * fused_table_bytes<PrepItemArgs>();
  for (int cls = 0; cls < 4; ++cls) {
    void* d = tab + cls * fused_table_bytes<GatherArgs>();
    if (gather[cls].t.count == 0) continue;
    switch (cls) {
      case 0: launch_gather_fused<8, 1>(gather[cls], d, stream); break;
      case 1: launch_gather_fused<16, 1>(gather[cls], d, stream); break;
      case 2: two ? launch_gather_fused<16, 2>(gather[cls], d, stream) : launch_gather_fused<32, 1>(gather[cls], d, stream); break;
      default: two ? launch_gather_fused<32, 2>(gather[cls], d, stream) : launch_gather_fused<64, 1>(gather[cls], d, stream); break;
    }
  }
  tab += 4 * fused_table_bytes<GatherArgs>();
  for (int cls = 0; cls < 4; ++cls) {
    void* d = tab + cls * fused_table_bytes<GatherArgs>();
    if (matrix[cls].t.count == 0) continue;
    switch (cls) {
      case 0: launch_gather_matrix_fused<8>(matrix[cls], matrix_w, matrix_taps == 1, d, stream); break;
      case 1: launch_gather_matrix_fused<16>(matrix[cls], matrix_w, matrix_taps == 1, d, stream); break;
      case 2: launch_gather_matrix_fused<32>(matrix[cls], matrix_w, matrix_taps == 1, d, stream); break;
      default: launch_gather_matrix_fused<64>(matrix[cls], matrix_w, matrix_taps == 1, d, stream); break;
    }
  }
  if (reduce.t.count > 0) {
    void* d = tab + 4 * fused_table_bytes<GatherArgs>();
    reduce.upload(d, stream);
    hipLaunchKernelGGL(reduce_fused_kernel, dim3(static_cast<uint32_t>(reduce.blocks())), dim3(256), 0, stream,
                       static_cast<const FusedTable<GatherArgs>*>(d));
    KSJ_LAUNCH_CHECK();
  }
}

// ----------------------------------------------------------------------------- chain form (host)
namespace {

constexpr int kMaxChainDevices = 64;
std::mutex g_chain_mu;
int g_chain_queues[kMaxChainDevices] = {};  // 0 not probed, -1 unusable, else XCD queues

// XCC ids seen by 2048 one-wave blocks: the chain needs them to be exactly 0 .. n - 1
// for n in {1, 2, 4, 8} (every queue then has workgroups).
int probe_chain_queues() {
  constexpr uint32_t kBlocks = 2048;
  uint32_t* d = nullptr;
  hipStream_t s = nullptr;
  KSJ_HIP(hipMalloc(&d, kBlocks * sizeof(uint32_t)));
  std::vector<uint32_t> h(kBlocks, 0xFFFFFFFFu);
  try {
    KSJ_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    hipLaunchKernelGGL(xcc_probe_kernel, dim3(kBlocks), dim3(64), 0, s, d);
    KSJ_LAUNCH_CHECK();
    KSJ_HIP(hipMemcpyAsync(h.data(), d, kBlocks * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    KSJ_HIP(hipStreamSynchronize(s));
  } catch (...) {
    if (s) (void)hipStreamDestroy(s);
    (void)hipFree(d);
    throw;
  }
  (void)hipStreamDestroy(s);
  (void)hipFree(d);
  uint32_t seen = 0;
  for (uint32_t v : h) {
    if (v >= 16) return -1;
    seen |= 1u << v;
  }
  for (int n : {8, 4, 2, 1})
    if (seen == (1u << n) - 1) return n;
  return -1;
}

// Queues of the chain on the current device (0: the chain is not usable).  The probe
// synchronises, so it is not run while `stream` is being captured.
uint32_t chain_queues(hipStream_t stream) {
  int dev = 0;
  KSJ_HIP(hipGetDevice(&dev));
  require(dev >= 0 && dev < kMaxChainDevices, "device ordinal beyond the chain table");
  std::lock_guard<std::mutex> lock(g_chain_mu);
  if (g_chain_queues[dev] == 0) {
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    if (stream != nullptr) KSJ_HIP(hipStreamIsCapturing(stream, &cs));
    if (cs != hipStreamCaptureStatusNone) return 0;
    g_chain_queues[dev] = probe_chain_queues();
  }
  return g_chain_queues[dev] > 0 ? static_cast<uint32_t>(g_chain_queues[dev]) : 0u;
}

constexpr size_t chain_table_bytes() { return (sizeof(ChainTable) + 255) & ~static_cast<size_t>(255); }
constexpr size_t chain_state_bytes() { return (sizeof(ChainState) + 255) & ~static_cast<size_t>(255); }

template <int CPMAX, int T, bool TS>
void launch_chain(const ChainTable* t, ChainState* st, hipStream_t stream) {
  constexpr size_t lds = prep_lds_bytes<CPMAX>();
  static std::atomic<int> per_cu{0};
  int n = per_cu.load(std::memory_order_relaxed);
  if (n == 0) {
    KSJ_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, grid_chain_kernel<CPMAX, 4, T, TS>, 256, lds));
    n = std::max(n, 1);
    per_cu.store(n, std::memory_order_relaxed);
  }
  const dim3 grid(static_cast<uint32_t>(n) * device_cu_count());
  hipLaunchKernelGGL((grid_chain_kernel<CPMAX, 4, T, TS>), grid, dim3(256), lds, stream, t, st, kChainSpinLimit);
  KSJ_LAUNCH_CHECK();
}

template <int CPMAX>
void launch_chain_cp(uint32_t block_w, bool share, const ChainTable* t, ChainState* st, hipStream_t stream) {
  if (block_w == 8) {
    if (share) launch_chain<CPMAX, 2, true>(t, st, stream); else launch_chain<CPMAX, 2, false>(t, st, stream);
  } else {
    if (share) launch_chain<CPMAX, 1, true>(t, st, stream); else launch_chain<CPMAX, 1, false>(t, st, stream);
  }
}

}  // namespace

void grid_chain_prepare(hipStream_t stream) { (void)chain_queues(stream); }

uint64_t grid_chain_faults() {
  unsigned int v = 0;
  KSJ_HIP(hipMemcpyFromSymbol(&v, HIP_SYMBOL(g_chain_faults), sizeof(v), 0, hipMemcpyDeviceToHost));
  return v;
}

size_t grid_chain_table_bytes() { return chain_table_bytes() + chain_state_bytes(); }

bool grid_chain_supported(const GridFusedItem* items, uint32_t count, hipStream_t stream) {
  if (count < 1 || count > kMaxFusedItems) return false;
  // The compiled chain shapes: packed visits, 16 visits per batch.
  if (tune_get(kTuneMatrixBatch) != 16) return false;
  uint32_t bw = 0;
  for (uint32_t i = 0; i < count; ++i) {
    const Plan2* p = items[i].plan;
    if (p == nullptr || !p->has_trajectory || p->scheme != kSchemeMatrix || !p->packed_visits) return false;
    if (items[i].channels < 1 || items[i].channels > kMaxStackedCoils) return false;
    if (bw != 0 && bw != p->block_w) return false;
    bw = p->block_w;
  }
  return chain_queues(stream) > 0;
}

void grid_chain(const GridFusedItem* items, uint32_t count, void* d_tables, hipStream_t stream) {
  require(grid_chain_supported(items, count, stream), "the gridding chain takes 1..64 packed matrix-schedule items");
  const uint32_t nq = chain_queues(stream);
  auto t = std::make_unique<ChainTable>();
  std::memset(t.get(), 0, sizeof(ChainTable));
  t->queues = nq;
  const uint32_t nt_store = static_cast<uint32_t>(tune_get(kTuneMatrixNtStore));
  std::vector<double> cost(count);
  uint32_t cpmax = 8;
  for (uint32_t i = 0; i < count; ++i) {
    const GridFusedItem& it = items[i];
    require(it.plan->multi_count == 0 || it.partial != nullptr, "multi-item blocks need a partial-sum buffer");
    const uint32_t cp = padded_coils(it.channels);
    cpmax = std::max(cpmax, cp);
    ChainItem& c = t->items[i];
    c.prep.planes = StridedPlanes{it.kspace, it.plan->d.samples};
    c.prep.weight = it.plan->weight;
    c.prep.rank = it.plan->rank;
    c.prep.sorted = it.sorted;
    c.prep.samples = it.plan->d.samples;
    c.prep.coils = it.channels;
    c.gather = gather_args(*it.plan, it.sorted, it.channels, it.grid, it.partial);
    c.gather.xcd_chunk = 0;
    c.gather.nt_store = nt_store;
    c.cp = cp;
    c.prep_tasks = ceil_div(ceil_div(c.prep.samples, kPrepSamples), kChainPrepBlocks);
    c.gather_tasks = ceil_div(c.gather.run_count, 4 * kChainRunsPerWave);
    c.reduce_tasks = ceil_div(static_cast<uint64_t>(c.gather.multi_count) * c.gather.bh * c.gather.bw * cp,
                              256ull * kChainReduceIters);
    const double g2 = static_cast<double>(it.plan->d.grid_rows) * it.plan->d.grid_cols;
    cost[i] = (24.0 * it.plan->d.samples + 8.0 * g2) * cp;
  }
  // Largest first onto the least loaded XCD.
  std::vector<uint32_t> order(count);
  for (uint32_t i = 0; i < count; ++i) order[i] = i;
  std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) { return cost[a] > cost[b]; });
  std::vector<double> load(nq, 0.0);
  std::vector<std::vector<uint32_t>> on(nq);
  for (uint32_t i : order) {
    const uint32_t x = static_cast<uint32_t>(std::min_element(load.begin(), load.end()) - load.begin());
    load[x] += cost[i];
    on[x].push_back(i);
  }
  uint32_t seg = 0;
  for (uint32_t x = 0; x < nq; ++x) {
    t->seg_begin[x] = seg;
    uint32_t task = 0;
    auto add = [&](uint32_t type, uint32_t item, uint32_t n) {
      if (n == 0) return;
      t->seg_first[seg] = task;
      t->seg_info[seg] = type << 16 | item;
      ++seg;
      task += n;
    };
    const auto& L = on[x];
    const size_t n = L.size();
    for (size_t k = 0; k < n + 2; ++k) {  // prep k, gather k - 1, reduce k - 2
      if (k < n) add(kTaskPrep, L[k], t->items[L[k]].prep_tasks);
      if (k >= 1 && k - 1 < n) add(kTaskGather, L[k - 1], t->items[L[k - 1]].gather_tasks);
      if (k >= 2 && k - 2 < n) add(kTaskReduce, L[k - 2], t->items[L[k - 2]].reduce_tasks);
    }
    t->task_count[x] = task;
  }
  for (uint32_t x = nq; x <= kXcds; ++x) t->seg_begin[x] = seg;
  const size_t used = offsetof(ChainTable, items) + count * sizeof(ChainItem);
  upload_table(d_tables, t.get(), (used + 3) & ~static_cast<size_t>(3), stream);
  auto* tab = static_cast<const ChainTable*>(d_tables);
  auto* st = reinterpret_cast<ChainState*>(static_cast<char*>(d_tables) + chain_table_bytes());
  hipLaunchKernelGGL(chain_reset_kernel, dim3(1), dim3(256), 0, stream, st);
  KSJ_LAUNCH_CHECK();
  const uint32_t bw = items[0].plan->block_w;
  const bool share = tune_get(kTuneMatrixTapShare) == 1;
  if (cpmax <= 16)
    launch_chain_cp<16>(bw, share, tab, st, stream);
  else if (cpmax == 32)
    launch_chain_cp<32>(bw, share, tab, st, stream);
  else
    launch_chain_cp<64>(bw, share, tab, st, stream);
}

}  // namespace ksj::gpu
