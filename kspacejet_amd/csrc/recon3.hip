// 3-D KB adjoint NUFFT on one GPU: coils in passes of up to coils_per_pass, each
// pass prep -> gather (2x2x2 blocks on the VALU, or 4x4x4 blocks on the matrix
// cores; no float atomics) -> multi-item reduce ->
// pruned 3-D inverse FFT (x over the full grid, y over the kept x columns, z over
// the kept (y, x) lines) with deapodization and crop fused into the y and z passes
// -> RSS accumulated across passes in channel order and/or coil-last images.
#include "fft.hpp"
#include "grid.hpp"
#include "matrix.hpp"
#include "plan3.hpp"
#include "prof.hpp"
#include "recon3.hpp"

#include <algorithm>

namespace ksj::gpu {
namespace {

struct Gather3Args {
  const float2* sorted;        // [S][Cp]
  const uint32_t* visit_k;
  const float4* visit_w;       // two per visit
  const uint4* item_info;
  const uint32_t* block_first_item;
  const uint32_t* block_first_slot;
  const uint32_t* multi_blocks;
  float2* partial;             // [slots][side^3][Cp]
  float2* grid;                // [G_z][G_y][G_x][Cp]
  uint32_t g[3], blocks[3];
  uint32_t cp, item_count, multi_count;
  uint32_t side;               // block side: 2, or 4 for the matrix schedule
  // matrix schedule
  const uint32_t* visit_pos;   // window position per visit
  const float* taps;           // [S][3W] sorted: z, y, x taps
  const uint32_t* run_first;   // [runs + 1]
  uint32_t run_count;
  int width;
  // workgroup -> XCD mapping (matrix3_map): 0 launch order; 1 XCD x takes the x-th eighth
  // of the workgroups (contiguous blocks); s > 1 slabs of s workgroups dealt round-robin
  uint32_t map, wg_per_xcd;
  // matrix schedule: per grid row (z, y), the x columns no sample reaches (not written), or null
  const uint2* zero_cols;
};

__device__ inline uint32_t last_visit(uint32_t v, uint32_t end) {
  const uint32_t last = end > 0 ? end - 1 : 0;
  return v < last ? v : last;
}

__device__ inline void block_cells(const Gather3Args& a, uint32_t b, uint32_t* z, uint32_t* y, uint32_t* x) {
  *x = a.side * (b % a.blocks[2]);
  const uint32_t r = b / a.blocks[2];
  *y = a.side * (r % a.blocks[1]);
  *z = a.side * (r / a.blocks[1]);
}

// One sub-wave of CP lanes (lane = coil) per work item, 64 / CP items per wave.
// Visits are taken B at a time: the batch's sample rows are all loaded before the
// first use, and the next batch's positions and weights (clamped into the item,
// values past it multiply zero) are loaded while this batch's rows are in flight.
// Sums run in visit order; the eight cells are written once (or to the item's
// partial slot).
template <int CP, int B>
__global__ __launch_bounds__(256) void gather3_kernel(Gather3Args a) {
  constexpr int NSW = 64 / CP;
  const int lane = threadIdx.x & 63, sw = lane / CP, c = lane % CP;
  const uint32_t item = (blockIdx.x * 4 + (threadIdx.x >> 6)) * NSW + sw;
  const bool live = item < a.item_count;
  const uint4 info = live ? a.item_info[item] : make_uint4(0, 0, 0, 0);
  float2 acc[8];
#pragma unroll
  for (int q = 0; q < 8; ++q) acc[q] = make_float2(0.f, 0.f);
  uint32_t k[B];
  float4 w0[B], w1[B];
#pragma unroll
  for (int i = 0; i < B; ++i) {
    const uint32_t vi = last_visit(info.y + i, info.z);
    k[i] = a.visit_k[vi];
    w0[i] = a.visit_w[2 * static_cast<uint64_t>(vi)];
    w1[i] = a.visit_w[2 * static_cast<uint64_t>(vi) + 1];
  }
  for (uint32_t v = info.y; v < info.z; v += B) {
    const uint32_t n = info.z - v;
    float2 d[B];
#pragma unroll
    for (int i = 0; i < B; ++i)
      d[i] = static_cast<uint32_t>(i) < n ? a.sorted[static_cast<uint64_t>(k[i]) * CP + c] : make_float2(0.f, 0.f);
    uint32_t kn[B];
    float4 w0n[B], w1n[B];
#pragma unroll
    for (int i = 0; i < B; ++i) {
      const uint32_t vi = last_visit(v + B + i, info.z);
      kn[i] = a.visit_k[vi];
      w0n[i] = a.visit_w[2 * static_cast<uint64_t>(vi)];
      w1n[i] = a.visit_w[2 * static_cast<uint64_t>(vi) + 1];
    }
#pragma unroll
    for (int i = 0; i < B; ++i) {
      const float w[8] = {w0[i].x, w0[i].y, w0[i].z, w0[i].w, w1[i].x, w1[i].y, w1[i].z, w1[i].w};
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        acc[q].x += w[q] * d[i].x;
        acc[q].y += w[q] * d[i].y;
      }
    }
#pragma unroll
    for (int i = 0; i < B; ++i) {
      k[i] = kn[i];
      w0[i] = w0n[i];
      w1[i] = w1n[i];
    }
  }
  if (!live) return;
  uint32_t z0, y0, x0;
  block_cells(a, info.x, &z0, &y0, &x0);
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    if (info.w == 0xFFFFFFFFu) {
      const uint32_t z = z0 + q / 4, y = y0 + (q / 2) % 2, x = x0 + q % 2;
      if (z < a.g[0] && y < a.g[1] && x < a.g[2])
        a.grid[((static_cast<uint64_t>(z) * a.g[1] + y) * a.g[2] + x) * CP + c] = acc[q];
    } else {
      a.partial[(static_cast<uint64_t>(info.w) * 8 + q) * CP + c] = acc[q];
    }
  }
}

// Blocks with several items: thread per (block, cell, coil) sums the partials in item
// order.  Cell q of a block is (q / s^2, (q / s) % s, q % s).
__global__ __launch_bounds__(256) void reduce3_kernel(Gather3Args a) {
  const uint32_t cells = a.side * a.side * a.side;
  const uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= static_cast<uint64_t>(a.multi_count) * cells * a.cp) return;
  const uint32_t c = static_cast<uint32_t>(i % a.cp);
  const uint32_t q = static_cast<uint32_t>((i / a.cp) % cells);
  const uint32_t b = a.multi_blocks[i / (static_cast<uint64_t>(cells) * a.cp)];
  const uint32_t s0 = a.block_first_slot[b];
  const uint32_t n = a.block_first_item[b + 1] - a.block_first_item[b];
  const float2* p = a.partial + (static_cast<uint64_t>(s0) * cells + q) * a.cp + c;
  float2 s = make_float2(0.f, 0.f);
  // Eight loads in flight, summed in item order; the last group reads clamped indices and
  // adds +0 for items past the end (s starts at +0 and never becomes -0, so s + 0 = s).
  const uint64_t stride = static_cast<uint64_t>(cells) * a.cp;
  for (uint32_t j = 0; j < n; j += 8) {
    float2 t[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const uint32_t k = j + u < n ? j + u : n - 1;
      t[u] = keep_bits(p[k * stride], j + u < n ? ~0u : 0u);
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) s = cadd(s, t[u]);
  }
  uint32_t z0, y0, x0;
  block_cells(a, b, &z0, &y0, &x0);
  const uint32_t z = z0 + q / (a.side * a.side), y = y0 + (q / a.side) % a.side, x = x0 + q % a.side;
  if (z < a.g[0] && y < a.g[1] && x < a.g[2]) a.grid[((static_cast<uint64_t>(z) * a.g[1] + y) * a.g[2] + x) * a.cp + c] = s;
}

// ----------------------------------------------------------------------------- matrix gather
// 4x4x4 blocks on the matrix cores.  For four visits at a time, each z-plane of the
// block (16 cells) is one v_mfma_f32_16x16x4_f32 per 16 columns of the sample rows:
// A [16 cells x 4 visits] weights, B [4 visits x 16 columns] row floats.  Lane l takes
// visit q = l / 16 of the group and cell (y, x) = (j / 4, j % 4), j = l % 16, of every
// plane; its four weights tz[z - rz + W - 1] ty[y - ry + W - 1] tx[x - rx + W - 1]
// (zero outside [0, W)) come from the sample's taps and the visit's window position,
// so a visit carries 8 bytes instead of 64 weights.  Accumulator register r of lane l
// in plane h is cell (h, q, r).  One wave per run of consecutive items, streamed in
// batches of U groups with the next batch's visits loaded while this batch's rows and
// taps are in flight; an item's 64 cells are written when the stream passes its end.
// Sums run four visits at a time in visit order: deterministic.
template <int CP, int U, bool TS>
__device__ inline void matrix3_run(const Gather3Args& a, uint32_t run) {
  constexpr int NF = CP / 8;
  constexpr int T = 4;  // z-planes of a block
  const int lane = threadIdx.x & 63, q = lane >> 4, j = lane & 15;
  const int cy = j >> 2, cx = j & 3;
  const uint32_t i0 = a.run_first[run], i1 = a.run_first[run + 1];  // run is wave-uniform: scalar loads
  if (i0 >= i1) return;
  const uint32_t nitems = i1 - i0;  // <= 64 by construction of the runs
  const uint4 rec = static_cast<uint32_t>(lane) < nitems ? a.item_info[i0 + lane] : make_uint4(0, 0, 0, 0);
  auto field = [&](uint32_t v, uint32_t t) {
    return static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(v), static_cast<int>(t)));
  };
  const uint32_t c0 = field(rec.y, 0) / 4, c1 = field(rec.z, nitems - 1) / 4;
  const uint32_t last_v = c1 > c0 ? 4 * c1 - 1 : 0;
  const int W = a.width;
  const float* rows = reinterpret_cast<const float*>(a.sorted) + piece_offset<NF>(j);
  MfmaAcc acc[T][NF];
#pragma unroll
  for (int h = 0; h < T; ++h)
#pragma unroll
    for (int n = 0; n < NF; ++n) acc[h][n] = MfmaAcc{0.f, 0.f, 0.f, 0.f};

  auto flush = [&](uint32_t t) {
    const uint32_t blk = field(rec.x, t), slot = field(rec.w, t);
    uint32_t z0, y0, x0;
    block_cells(a, blk, &z0, &y0, &x0);
    if (a.zero_cols != nullptr && slot == 0xFFFFFFFFu && z0 < a.g[0] && y0 < a.g[1]) {
      // A block inside its rows' zero span (no sample reaches it): the x FFT reads those
      // cells as zero, so they are not written.
      const uint2 zs = a.zero_cols[static_cast<uint64_t>(z0) * a.g[1] + y0];
      if (x0 >= zs.x && x0 + 4 <= zs.y) {
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
        if (slot == 0xFFFFFFFFu) {
          const uint32_t z = z0 + h, y = y0 + q, x = x0 + r;
          if (z < a.g[0] && y < a.g[1] && x < a.g[2])
            store_piece_nt<NF>(reinterpret_cast<float*>(a.grid) +
                                 ((static_cast<uint64_t>(z) * a.g[1] + y) * a.g[2] + x) * (2 * CP) + piece_offset<NF>(j),
                               o);
        } else {
          store_piece<NF>(reinterpret_cast<float*>(a.partial) +
                            (static_cast<uint64_t>(slot) * 64 + 16 * h + 4 * q + r) * (2 * CP) + piece_offset<NF>(j),
                          o);
        }
      }
    }
#pragma unroll
    for (int h = 0; h < T; ++h)
#pragma unroll
      for (int n = 0; n < NF; ++n) acc[h][n] = MfmaAcc{0.f, 0.f, 0.f, 0.f};
  };

  uint32_t k[U], pos[U];
  auto load_meta = [&](uint32_t c) {  // groups c + u, clamped into the run
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const uint32_t vi = 4 * (c + u) + q;
      const uint32_t vc = vi < last_v ? vi : last_v;
      k[u] = a.visit_k[vc];
      pos[u] = a.visit_pos[vc];
    }
  };

  uint32_t t = 0, t_end = field(rec.z, 0) / 4;  // current item and its end group
  uint32_t c = c0;
  if (c < c1) load_meta(c);
  for (; c < c1; c += U) {
    RowPiece<NF> b[U];
    float tz[U][T], ty[U], tx[U];
    int rzs[U], iys[U], ixs[U];
    bool vyx[U];
    float tw0[U], tw1[U];  // tap-sharing form: words j and j + 16 of the visit's 3W taps
#pragma unroll
    for (int u = 0; u < U; ++u) {
      b[u] = load_piece<NF>(rows + static_cast<uint64_t>(k[u]) * (2 * CP));
      const float* tp = a.taps + static_cast<uint64_t>(k[u]) * (3 * W);
      const int rz = static_cast<int>(pos[u] & 15), ry = static_cast<int>((pos[u] >> 4) & 15),
                rx = static_cast<int>((pos[u] >> 8) & 15);
      const int iy = cy - ry + W - 1, ix = cx - rx + W - 1;
      vyx[u] = iy >= 0 && iy < W && ix >= 0 && ix < W;
      rzs[u] = rz;
      iys[u] = iy >= 0 && iy < W ? iy : 0;
      ixs[u] = ix >= 0 && ix < W ? ix : 0;
      if constexpr (TS) {
        // The 16 lanes of a visit need 3W distinct taps between them: each loads one or
        // two words and takes its own from the visit's lanes by bpermute after the wait.
        tw0[u] = tp[j < 3 * W ? j : 3 * W - 1];
        tw1[u] = 3 * W > 16 ? tp[j + 16 < 3 * W ? j + 16 : 3 * W - 1] : 0.f;
      } else {
        ty[u] = tp[W + iys[u]];
        tx[u] = tp[2 * W + ixs[u]];
#pragma unroll
        for (int h = 0; h < T; ++h) {
          const int iz = h - rz + W - 1;
          tz[u][h] = tp[iz >= 0 && iz < W ? iz : 0];
        }
      }
    }
    load_meta(c + U);
    // Keep every load of the batch ahead of the products (see the 2-D gather).
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int u = 0; u < U; ++u) {
      if (c + u < c1) {
        while (c + u >= t_end) {  // group c + u starts a later item
          flush(t);
          ++t;
          t_end = field(rec.z, t) / 4;
        }
        if constexpr (TS) {
          // word w of visit q lies in lane 16 q + w (w < 16) of tw0, or 16 q + w - 16 of tw1
          auto tap = [&](int w) {
            const int w0 = w < 16 ? w : w - 16;
            const int addr = 4 * (16 * q + w0);
            const float lo = __int_as_float(__builtin_amdgcn_ds_bpermute(addr, __float_as_int(tw0[u])));
            if (3 * W <= 16) return lo;
            const float hi = __int_as_float(__builtin_amdgcn_ds_bpermute(addr, __float_as_int(tw1[u])));
            return w < 16 ? lo : hi;
          };
          ty[u] = tap(W + iys[u]);
          tx[u] = tap(2 * W + ixs[u]);
#pragma unroll
          for (int h = 0; h < T; ++h) {
            const int iz = h - rzs[u] + W - 1;
            tz[u][h] = tap(iz >= 0 && iz < W ? iz : 0);
          }
        }
        const float wyx = vyx[u] ? ty[u] * tx[u] : 0.f;
#pragma unroll
        for (int h = 0; h < T; ++h) {
          const int iz = h - rzs[u] + W - 1;
          const float wv = (iz >= 0 && iz < W) ? tz[u][h] * wyx : 0.f;
#pragma unroll
          for (int n = 0; n < NF; ++n)
            acc[h][n] = __builtin_amdgcn_mfma_f32_16x16x4f32(wv, b[u].v[n], acc[h][n], 0, 0, 0);
        }
      }
    }
  }
  for (; t < nitems; ++t) flush(t);  // the last item with visits, then any empty ones
}

// Four waves per workgroup, one run per wave.  Hardware places workgroup b on XCD b % 8;
// matrix3_map decides which runs it takes: in launch order (0), the XCD's contiguous
// eighth of the runs (1), or slabs of `map` consecutive workgroups dealt round-robin to
// the XCDs (> 1), so neighbouring blocks, which share sample rows, meet in one L2.
template <int CP, int U, bool TS>
__global__ __launch_bounds__(256) void gather3_matrix_kernel(Gather3Args a) {
  constexpr uint32_t kXcd = 8;
  uint32_t wg = blockIdx.x;
  const uint32_t x = blockIdx.x % kXcd, l = blockIdx.x / kXcd;
  if (a.map == 1)
    wg = x * a.wg_per_xcd + l;
  else if (a.map > 1)
    wg = ((l / a.map) * kXcd + x) * a.map + l % a.map;
  const uint32_t run = wg * 4 + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  if (run < a.run_count) matrix3_run<CP, U, TS>(a, run);
}

// RSS over the coil-fastest z output [P][coils], carried across coil passes in channel
// order: s = (first ? 0 : sumsq) + |x_0|^2 + |x_1|^2 + ...  A block's kRss3Pixels pixels
// hold one contiguous run of values: their |x|^2 are staged in LDS by coalesced loads,
// then each thread sums its pixel's coils.
constexpr uint32_t kRss3Pixels = 256;
__global__ __launch_bounds__(kRss3Pixels) void rss3_kernel(const float2* img, uint32_t coils, uint64_t pixels,
                                                           float* sumsq, float* rss, int first, int last) {
  extern __shared__ float norms[];  // [kRss3Pixels][coils]
  const uint64_t p0 = static_cast<uint64_t>(blockIdx.x) * kRss3Pixels;
  const uint32_t np = pixels - p0 < kRss3Pixels ? static_cast<uint32_t>(pixels - p0) : kRss3Pixels;
  const float2* src = img + p0 * coils;
  for (uint32_t i = threadIdx.x; i < np * coils; i += kRss3Pixels) norms[i] = norm2(src[i]);
  __syncthreads();
  if (threadIdx.x >= np) return;
  const uint64_t p = p0 + threadIdx.x;
  float s = first ? 0.f : sumsq[p];
  for (uint32_t c = 0; c < coils; ++c) s += norms[threadIdx.x * coils + c];
  if (last)
    rss[p] = sqrtf(s);
  else
    sumsq[p] = s;
}
// Coil-fastest z output [P][coils] into coil-last images [P][channels] at channel c0.
__global__ void coil_last3_kernel(const float2* img, uint32_t coils, uint64_t pixels, float2* out, uint32_t c0,
                                  uint32_t channels) {
  const uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= pixels * coils) return;
  const uint64_t p = i / coils;
  const uint32_t c = static_cast<uint32_t>(i - p * coils);
  out[p * channels + c0 + c] = img[i];
}

template <int CP>
void launch_gather3(const Gather3Args& a, hipStream_t stream) {
  if (a.side == 4) {
    Gather3Args m = a;
    uint32_t wgs = std::max<uint32_t>(1, ceil_div(a.run_count, 4));
    m.map = static_cast<uint32_t>(tune_get(kTuneMatrix3Map));
    m.wg_per_xcd = ceil_div(wgs, 8u);
    // padding workgroups find no run and exit
    if (m.map == 1) wgs = m.wg_per_xcd * 8;
    else if (m.map > 1) wgs = ceil_div(wgs, 8 * m.map) * 8 * m.map;
    auto go = [&]<bool TS>() {
      switch (tune_get(kTuneMatrixBatch)) {  // visits per batch: groups of four
        case 4: hipLaunchKernelGGL((gather3_matrix_kernel<CP, 1, TS>), dim3(wgs), dim3(256), 0, stream, m); break;
        case 8: hipLaunchKernelGGL((gather3_matrix_kernel<CP, 2, TS>), dim3(wgs), dim3(256), 0, stream, m); break;
        default: hipLaunchKernelGGL((gather3_matrix_kernel<CP, 4, TS>), dim3(wgs), dim3(256), 0, stream, m); break;
      }
    };
    // 3W taps over 16 lanes and at most two words each: widths up to 10
    if (tune_get(kTuneMatrix3TapShare) == 1 && 3 * a.width <= 32)
      go.template operator()<true>();
    else
      go.template operator()<false>();
    KSJ_LAUNCH_CHECK();
    return;
  }
  constexpr int NSW = 64 / CP;
  const uint32_t wgs = std::max<uint32_t>(1, ceil_div(ceil_div(a.item_count, NSW), 4));
  hipLaunchKernelGGL((gather3_kernel<CP, 4>), dim3(wgs), dim3(256), 0, stream, a);
  KSJ_LAUNCH_CHECK();
}

struct Layout3 {
  size_t sorted, grid, partial, xout, yout, zout, sumsq, total;
};

size_t align256_3(size_t n) { return (n + 255) & ~static_cast<size_t>(255); }

Layout3 layout3(const Plan3& p, uint32_t channels, uint32_t coils_per_pass) {
  const Plan3Desc& d = p.d;
  const uint32_t cpp = std::min(channels, coils_per_pass);
  const size_t cp = padded_coils(cpp);
  const size_t G = static_cast<size_t>(d.grid[0]) * d.grid[1] * d.grid[2];
  const size_t P = static_cast<size_t>(d.image[0]) * d.image[1] * d.image[2];
  Layout3 l{};
  l.sorted = align256_3(static_cast<size_t>(d.samples) * cp * sizeof(float2));
  l.grid = align256_3(G * cp * sizeof(float2));
  const size_t cells = static_cast<size_t>(p.side) * p.side * p.side;
  l.partial = align256_3(std::max<size_t>(1, static_cast<size_t>(p.slot_count) * cells * cp) * sizeof(float2));
  l.xout = align256_3(cpp * static_cast<size_t>(d.grid[0]) * d.grid[1] * d.image[2] * sizeof(float2));
  l.yout = align256_3(cpp * static_cast<size_t>(d.grid[0]) * d.image[1] * d.image[2] * sizeof(float2));
  l.zout = align256_3(cpp * P * sizeof(float2));
  l.sumsq = align256_3(P * sizeof(float));
  l.total = l.sorted + l.grid + l.partial + l.xout + l.yout + l.zout + l.sumsq;
  return l;
}

}  // namespace

size_t adjoint3_workspace_bytes(const Plan3& plan, uint32_t channels, uint32_t coils_per_pass) {
  require(channels >= 1 && coils_per_pass >= 1 && coils_per_pass <= kMaxStackedCoils,
          "3-D adjoint needs channels >= 1 and 1..64 coils per pass");
  require(plan.has_trajectory, "plan has no trajectory");
  return layout3(plan, channels, coils_per_pass).total;
}

void adjoint3_execute(const Plan3& plan, const float2* d_kspace, uint32_t channels, uint32_t coils_per_pass,
                      float2* d_images, float* d_rss, void* d_workspace, size_t workspace_bytes, hipStream_t stream) {
  require(d_kspace != nullptr, "k-space must be non-null");
  require(d_images != nullptr || d_rss != nullptr, "3-D adjoint needs an image or an RSS output");
  const Layout3 l = layout3(plan, channels, coils_per_pass);
  require(adjoint3_workspace_bytes(plan, channels, coils_per_pass) == l.total && d_workspace != nullptr &&
            workspace_bytes >= l.total,
          "3-D adjoint workspace is too small");
  const Plan3Desc& d = plan.d;
  const uint64_t P = static_cast<uint64_t>(d.image[0]) * d.image[1] * d.image[2];
  char* w = static_cast<char*>(d_workspace);
  auto* sorted = reinterpret_cast<float2*>(w);
  auto* grid = reinterpret_cast<float2*>(w + l.sorted);
  auto* partial = reinterpret_cast<float2*>(w + l.sorted + l.grid);
  auto* xout = reinterpret_cast<float2*>(w + l.sorted + l.grid + l.partial);
  auto* yout = reinterpret_cast<float2*>(w + l.sorted + l.grid + l.partial + l.xout);
  auto* zout = reinterpret_cast<float2*>(w + l.sorted + l.grid + l.partial + l.xout + l.yout);
  auto* sumsq = reinterpret_cast<float*>(w + l.sorted + l.grid + l.partial + l.xout + l.yout + l.zout);
  const uint32_t cpp = std::min(channels, coils_per_pass);
  for (uint32_t c0 = 0; c0 < channels; c0 += cpp) {
    const uint32_t pc = std::min(cpp, channels - c0);
    const uint32_t cp = padded_coils(pc);
    CoilPlanes planes;
    planes.coils = pc;
    for (uint32_t c = 0; c < pc; ++c) planes.plane[c] = d_kspace + static_cast<uint64_t>(c0 + c) * d.samples;
    {
      ProfScope ps("grid3_prep", stream);
      grid_prep_rows(plan.weight, plan.rank, d.samples, planes, sorted, stream);
    }
    Gather3Args a{};
    a.sorted = sorted;
    a.visit_k = plan.visit_k;
    a.visit_w = plan.visit_w;
    a.item_info = plan.item_info;
    a.block_first_item = plan.block_first_item;
    a.block_first_slot = plan.block_first_slot;
    a.multi_blocks = plan.multi_blocks;
    a.partial = partial;
    a.grid = grid;
    for (int i = 0; i < 3; ++i) {
      a.g[i] = d.grid[i];
      a.blocks[i] = plan.blocks[i];
    }
    a.cp = cp;
    a.item_count = plan.item_count;
    a.multi_count = plan.multi_count;
    a.side = plan.side;
    a.visit_pos = plan.visit_pos;
    a.taps = plan.taps;
    a.run_first = plan.run_first;
    a.run_count = plan.run_count;
    a.width = static_cast<int>(d.width);
    const uint2* zero_cols = tune_get(kTuneGridPrune) == 1 ? plan.zero_cols : nullptr;
    a.zero_cols = zero_cols;
    {
      ProfScope ps("grid3_gather", stream);
      switch (cp) {
        case 8: launch_gather3<8>(a, stream); break;
        case 16: launch_gather3<16>(a, stream); break;
        case 32: launch_gather3<32>(a, stream); break;
        default: launch_gather3<64>(a, stream); break;
      }
    }
    if (plan.multi_count > 0) {
      ProfScope ps("grid3_reduce", stream);
      const uint64_t cells = static_cast<uint64_t>(plan.side) * plan.side * plan.side;
      hipLaunchKernelGGL(reduce3_kernel, dim3(ceil_div(static_cast<uint64_t>(plan.multi_count) * cells * cp, 256)),
                         dim3(256), 0, stream, a);
      KSJ_LAUNCH_CHECK();
    }
    const uint32_t Gz = d.grid[0], Gy = d.grid[1], Gx = d.grid[2];
    const uint32_t Nz = d.image[0], Ny = d.image[1], Nx = d.image[2];
    // Every pass output is coil-fastest, so each pass loads and stores the same element
    // of consecutive lines (the register form where the length has one).
    // x: lines (z, y, coil) of the coil-fastest grid -> [Gz][Gy][Nx][pc], crop to the centred Nx.
    FftArgs fx;
    fx.in = grid;
    fx.out = xout;
    fx.src.planes = Gz * Gy;
    fx.src.lines = pc;
    fx.src.plane_stride = static_cast<uint64_t>(Gx) * cp;
    fx.src.line_stride = 1;
    fx.src.elem_stride = cp;
    fx.dst.planes = Gz * Gy;
    fx.dst.lines = pc;
    fx.dst.plane_stride = static_cast<uint64_t>(Nx) * pc;
    fx.dst.line_stride = 1;
    fx.dst.elem_stride = pc;
    fx.n = Gx;
    fx.n_out = Nx;
    fx.out_offset = (Gx - Nx / 2) % Gx;
    fx.inverse = 1;
    fx.load_lines_fast = fx.store_lines_fast = 1;
    fx.zero_span = zero_cols;  // per plane (z, y): the columns the gather left unwritten
    const uint2* zero_rows = zero_cols ? plan.zero_rows : nullptr;
    fx.skip_span = zero_rows;  // rows outside the sampled ball: all zero, not stored
    fx.skip_div = Gy;
    {
      ProfScope ps("fft3_x", stream);
      fft_launch(fx, stream);
    }
    // y: lines (z; x, coil) -> [Gz][Ny][Nx][pc]; deapodization of y (per output) and x (per
    // line, shared by a column's coils).
    FftArgs fy;
    fy.in = xout;
    fy.out = yout;
    fy.src.planes = Gz;
    fy.src.lines = Nx * pc;
    fy.src.plane_stride = static_cast<uint64_t>(Gy) * Nx * pc;
    fy.src.line_stride = 1;
    fy.src.elem_stride = static_cast<uint64_t>(Nx) * pc;
    fy.dst.planes = Gz;
    fy.dst.lines = Nx * pc;
    fy.dst.plane_stride = static_cast<uint64_t>(Ny) * Nx * pc;
    fy.dst.line_stride = 1;
    fy.dst.elem_stride = static_cast<uint64_t>(Nx) * pc;
    fy.n = Gy;
    fy.n_out = Ny;
    fy.out_offset = (Gy - Ny / 2) % Gy;
    fy.out_weight = plan.deapod[1];
    fy.line_weight = plan.deapod[2];
    fy.line_weight_div = pc;
    fy.zero_span = zero_rows;  // per plane z: the rows the x pass did not store
    fy.inverse = 1;
    fy.load_lines_fast = fy.store_lines_fast = 1;
    {
      ProfScope ps("fft3_y_deapod", stream);
      fft_launch(fy, stream);
    }
    // z: lines (pixel (y, x); coil) -> deapodization of z; the pass itself writes the coil-last
    // images and carries the RSS sum across coil passes where its epilogue applies, else
    // it stores [Nz][Ny][Nx][pc] for the two combine kernels.
    FftArgs fz;
    fz.in = yout;
    fz.out = zout;
    fz.src.planes = Ny * Nx;
    fz.src.lines = pc;
    fz.src.plane_stride = pc;
    fz.src.line_stride = 1;
    fz.src.elem_stride = static_cast<uint64_t>(Ny) * Nx * pc;
    fz.dst = fz.src;
    fz.n = Gz;
    fz.n_out = Nz;
    fz.out_offset = (Gz - Nz / 2) % Gz;
    fz.out_weight = plan.deapod[0];
    fz.inverse = 1;
    fz.load_lines_fast = fz.store_lines_fast = 1;
    const bool first = c0 == 0, last = c0 + pc >= channels;
    if (fft_columns_supported(fz, pc)) {
      ColumnOutputs co;
      co.channels = pc;
      co.cols = Ny * Nx;
      co.img_stride = channels;
      co.coil_last[0] = d_images ? d_images + c0 : nullptr;
      co.rss[0] = d_rss;
      co.sumsq = d_rss ? sumsq : nullptr;
      co.first = first;
      co.last = last;
      ProfScope ps("fft3_z_deapod", stream);
      fft_launch_columns(fz, co, stream);
      continue;
    }
    {
      ProfScope ps("fft3_z_deapod", stream);
      fft_launch(fz, stream);
    }
    ProfScope ps("coil_combine3", stream);
    if (d_rss) {
      const size_t lds = static_cast<size_t>(kRss3Pixels) * pc * sizeof(float);  // <= 64 KiB at 64 coils
      hipLaunchKernelGGL(rss3_kernel, dim3(ceil_div(P, static_cast<uint64_t>(kRss3Pixels))), dim3(kRss3Pixels), lds,
                         stream, zout, pc, P, sumsq, d_rss, first ? 1 : 0, last ? 1 : 0);
      KSJ_LAUNCH_CHECK();
    }
    if (d_images) {
      hipLaunchKernelGGL(coil_last3_kernel, dim3(ceil_div(P * pc, 256)), dim3(256), 0, stream, zout, pc, P, d_images,
                         c0, channels);
      KSJ_LAUNCH_CHECK();
    }
  }
}

}  // namespace ksj::gpu
