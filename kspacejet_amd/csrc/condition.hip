// K-space conditioning kernels.  All three are streaming passes over channel-major
// k-space, bound by HBM bandwidth except the channel mix at high channel counts
// (8 * in flops per complex input value read).
#include "common.hpp"
#include "condition.hpp"
#include "matrix.hpp"
#include "prof.hpp"

#include <algorithm>

namespace ksj::gpu {
namespace {

typedef float F2 __attribute__((ext_vector_type(2)));

// A non-temporal 8-byte store of one complex value (the output is not read again here).
__device__ inline void store_nt(float2* p, float re, float im) {
  const F2 v = {re, im};
  __builtin_nontemporal_store(v, reinterpret_cast<F2*>(p));
}

constexpr int kMixBlock = 256;
constexpr uint32_t kMixGroup = 16;  // outputs per block along grid y when the grid is small

// One thread per sample: the sample's IN_MAX input values (zero past in_ch) stay in
// registers while the block's outputs are formed from them.  The block's matrix rows sit
// in LDS, zero-padded to IN_MAX columns, so the inner loop has no branches and every lane
// reads the same LDS address (a broadcast).
template <int IN_MAX>
__global__ __launch_bounds__(kMixBlock) void channel_mix_kernel(const float2* __restrict__ w,
                                                                const float2* __restrict__ x,
                                                                float2* __restrict__ y, uint32_t out_ch,
                                                                uint32_t in_ch, uint64_t samples, uint32_t group) {
  __shared__ float4 rows[kMixMaxChannels * IN_MAX / 2];  // [group][IN_MAX] complex, two per float4
  const uint32_t o0 = blockIdx.y * group;
  const uint32_t count = min(out_ch - o0, group);
  auto* lds = reinterpret_cast<float2*>(rows);
  for (uint32_t e = threadIdx.x; e < count * IN_MAX; e += kMixBlock) {
    const uint32_t o = e / IN_MAX, i = e % IN_MAX;
    lds[e] = i < in_ch ? w[static_cast<uint64_t>(o0 + o) * in_ch + i] : make_float2(0.f, 0.f);
  }
  __syncthreads();
  const uint64_t s = static_cast<uint64_t>(blockIdx.x) * kMixBlock + threadIdx.x;
  if (s >= samples) return;
  float2 v[IN_MAX];
#pragma unroll
  for (int i = 0; i < IN_MAX; ++i)
    v[i] = static_cast<uint32_t>(i) < in_ch ? x[static_cast<uint64_t>(i) * samples + s] : make_float2(0.f, 0.f);
  for (uint32_t o = 0; o < count; ++o) {
    const float4* row = rows + o * (IN_MAX / 2);
    float re = 0.f, im = 0.f;
#pragma unroll
    for (int i = 0; i < IN_MAX; i += 2) {
      const float4 c = row[i / 2];  // matrix entries i (x, y) and i + 1 (z, w)
      re = fmaf(c.x, v[i].x, re);
      re = fmaf(-c.y, v[i].y, re);
      im = fmaf(c.x, v[i].y, im);
      im = fmaf(c.y, v[i].x, im);
      re = fmaf(c.z, v[i + 1].x, re);
      re = fmaf(-c.w, v[i + 1].y, re);
      im = fmaf(c.z, v[i + 1].y, im);
      im = fmaf(c.w, v[i + 1].x, im);
    }
    store_nt(&y[static_cast<uint64_t>(o0 + o) * samples + s], re, im);
  }
}

template <int IN_MAX>
void launch_mix(const float2* w, const float2* x, float2* y, uint32_t out_ch, uint32_t in_ch, uint64_t samples,
                hipStream_t stream) {
  const unsigned blocks = ceil_div(samples, kMixBlock);
  // All outputs in one block once the sample blocks alone fill the chip (each input value
  // is then read once); otherwise groups of 16 outputs along grid y.
  const uint32_t group = blocks >= 512 ? out_ch : kMixGroup;
  const dim3 grid(blocks, ceil_div(out_ch, group));
  ProfScope ps("channel_mix", stream);
  hipLaunchKernelGGL(channel_mix_kernel<IN_MAX>, grid, dim3(kMixBlock), 0, stream, w, x, y, out_ch, in_ch, samples,
                     group);
  KSJ_LAUNCH_CHECK();
}

// The mix on the matrix cores (v_mfma_f32_16x16x4_f32): complex Y = W X as the real
// product [Yr; Yi] = [[Wr, -Wi], [Wi, Wr]] [Xr; Xi], one wave per tile of 16 samples.
//   M (16 rows of a row block rb): output o = 8 rb + m / 2, real part for even m, imaginary
//     for odd m;
//   K, per pair p of k-steps: step 2p takes the real parts and step 2p + 1 the imaginary
//     parts of input channels 4p + j (j = 0..3, the k index within a step), so one float2
//     load per lane feeds both steps;
//   N: the 16 samples of the tile.
// Lane l holds A[l % 16][l / 16] and B[l / 16][l % 16]; accumulator register r holds
// D[4 (l / 16) + r][l % 16], i.e. outputs 8 rb + 2 (l / 16) and 8 rb + 2 (l / 16) + 1
// (re, im, re, im) of sample l % 16.  The matrix sits in LDS zero-padded to 8 RB rows and
// 4 PAIRS columns (32 bytes of row padding spread the rows over the banks).
constexpr int kMixMfmaBlock = 512;  // 8 waves

template <int RB, int PAIRS>
__global__ __launch_bounds__(kMixMfmaBlock) void channel_mix_mfma_kernel(const float2* __restrict__ w,
                                                                         const float2* __restrict__ x,
                                                                         float2* __restrict__ y, uint32_t out_ch,
                                                                         uint32_t in_ch, uint64_t samples) {
  constexpr int kRows = 8 * RB, kCols = 4 * PAIRS, kStride = kCols + 4;
  __shared__ float2 ws[kRows * kStride];
  for (int e = threadIdx.x; e < kRows * kCols; e += kMixMfmaBlock) {
    const uint32_t o = e / kCols, i = e % kCols;
    ws[o * kStride + i] = o < out_ch && i < in_ch ? w[static_cast<uint64_t>(o) * in_ch + i] : make_float2(0.f, 0.f);
  }
  __syncthreads();
  const int lane = threadIdx.x & 63;
  const int n = lane & 15, j = lane >> 4;
  const bool even = (n & 1) == 0;
  const float2* wrow = ws + (n >> 1) * kStride + j;  // row o = n / 2 of block 0, column j of pair 0
  const uint64_t tiles = (samples + 15) / 16;
  const uint64_t waves = static_cast<uint64_t>(gridDim.x) * (kMixMfmaBlock / 64);
  // Lane's inputs of tile t: x[4p + j][16 t + n], zero past the samples and the channels.
  // Branch-free, so all of a lane's loads are in flight together: out-of-range positions
  // read an in-range element and are cleared.
  auto load_tile = [&](uint64_t t, float2 (&b)[PAIRS]) {
    const uint64_t s = t * 16 + n;
    const bool in_s = s < samples;
    const uint64_t sc = in_s ? s : samples - 1;
#pragma unroll
    for (int p = 0; p < PAIRS; ++p) {
      const uint32_t c = 4 * p + j;
      const uint32_t keep = in_s && c < in_ch ? ~0u : 0u;
      b[p] = keep_bits(x[static_cast<uint64_t>(c < in_ch ? c : 0u) * samples + sc], keep);
    }
  };
  // With up to 16 input channels the next tile's inputs are loaded while this tile's
  // products run (11.8 -> 11.2 us at 16 x 16); at 32 the extra registers cost as much
  // occupancy as the prefetch gains, so larger matrices load per tile.
  constexpr bool kPrefetch = PAIRS <= 4;
  float2 b[PAIRS];
  uint64_t t = static_cast<uint64_t>(blockIdx.x) * (kMixMfmaBlock / 64) + (threadIdx.x >> 6);
  if constexpr (kPrefetch) load_tile(t, b);
  // Every wave strides over the tiles and leaves when it runs out of them.
  for (; t < tiles; t += waves) {
    const uint64_t s = t * 16 + n;
    const bool live = s < samples;
    float2 next[PAIRS];
    if constexpr (kPrefetch)
      load_tile(t + waves, next);  // past the last tile every lane loads zeros
    else
      load_tile(t, b);
    MfmaAcc acc[RB];
#pragma unroll
    for (int rb = 0; rb < RB; ++rb) acc[rb] = MfmaAcc{0.f, 0.f, 0.f, 0.f};
    // Small matrices stay in registers across tiles; larger ones are read from LDS in
    // every tile (an offset the compiler cannot see through keeps them from being hoisted
    // into registers, which would spill).
    int lds_off = 0;
    if constexpr (RB * PAIRS > 16) asm volatile("" : "+v"(lds_off));
#pragma unroll
    for (int p = 0; p < PAIRS; ++p) {
#pragma unroll
      for (int rb = 0; rb < RB; ++rb) {
        const float2 wv = wrow[lds_off + rb * 8 * kStride + 4 * p];
        const float a_re = even ? wv.x : wv.y;   // coefficient of Re(x): Wr (re row), Wi (im row)
        const float a_im = even ? -wv.y : wv.x;  // coefficient of Im(x): -Wi (re row), Wr (im row)
        acc[rb] = __builtin_amdgcn_mfma_f32_16x16x4f32(a_re, b[p].x, acc[rb], 0, 0, 0);
        acc[rb] = __builtin_amdgcn_mfma_f32_16x16x4f32(a_im, b[p].y, acc[rb], 0, 0, 0);
      }
    }
    if (live) {
#pragma unroll
      for (int rb = 0; rb < RB; ++rb) {
        const uint32_t o = 8 * rb + 2 * j;
        if (o < out_ch) store_nt(&y[static_cast<uint64_t>(o) * samples + s], acc[rb][0], acc[rb][1]);
        if (o + 1 < out_ch) store_nt(&y[static_cast<uint64_t>(o + 1) * samples + s], acc[rb][2], acc[rb][3]);
      }
    }
    if constexpr (kPrefetch) {
#pragma unroll
      for (int p = 0; p < PAIRS; ++p) b[p] = next[p];
    }
  }
}

template <int RB, int PAIRS>
void launch_mix_mfma(const float2* w, const float2* x, float2* y, uint32_t out_ch, uint32_t in_ch, uint64_t samples,
                     hipStream_t stream) {
  const uint64_t tiles = (samples + 15) / 16;
  // About four tiles per wave, at most 2048 blocks: each block fills its LDS copy of the
  // matrix once.  (Two tiles per wave measured no faster at 16 channels and slower at 64.)
  const unsigned blocks = static_cast<unsigned>(std::min<uint64_t>(2048, std::max<uint64_t>(1, (tiles + 31) / 32)));
  ProfScope ps("channel_mix", stream);
  hipLaunchKernelGGL((channel_mix_mfma_kernel<RB, PAIRS>), dim3(blocks), dim3(kMixMfmaBlock), 0, stream, w, x, y,
                     out_ch, in_ch, samples);
  KSJ_LAUNCH_CHECK();
}

template <int RB>
void mix_mfma_cols(const float2* w, const float2* x, float2* y, uint32_t out_ch, uint32_t in_ch, uint64_t samples,
                   hipStream_t stream) {
  if (in_ch <= 16)
    launch_mix_mfma<RB, 4>(w, x, y, out_ch, in_ch, samples, stream);
  else if (in_ch <= 32)
    launch_mix_mfma<RB, 8>(w, x, y, out_ch, in_ch, samples, stream);
  else
    launch_mix_mfma<RB, 16>(w, x, y, out_ch, in_ch, samples, stream);
}

// Grid: x over columns, y over rows (each block strides over the rows), z over channels;
// the column's phase value is loaded once per thread.
__global__ void phase_correct_kernel(const float2* __restrict__ x, const float2* __restrict__ p,
                                     float2* __restrict__ y, uint32_t rows, uint32_t cols) {
  // Each product and each sum rounds to float on its own, as in the reference.
#pragma clang fp contract(off)
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= cols) return;
  const uint32_t c = blockIdx.z;
  const float2 b = p[static_cast<uint64_t>(c) * cols + k];
  for (uint32_t r = blockIdx.y; r < rows; r += gridDim.y) {
    const uint64_t e = (static_cast<uint64_t>(c) * rows + r) * cols + k;
    const float2 a = x[e];
    const float ac = a.x * b.x, bd = a.y * b.y, ad = a.x * b.y, bc = a.y * b.x;
    store_nt(&y[e], ac - bd, ad + bc);
  }
}

// Grid: x over the kept columns, y over lines (each block strides over the lines).
__global__ void readout_crop_kernel(const float2* __restrict__ x, float2* __restrict__ y, uint64_t lines,
                                    uint32_t in_cols, uint32_t out_cols, uint32_t offset) {
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= out_cols) return;
  for (uint64_t l = blockIdx.y; l < lines; l += gridDim.y) {
    const float2 v = x[l * in_cols + offset + k];
    store_nt(&y[l * out_cols + k], v.x, v.y);
  }
}

}  // namespace

void channel_mix(const float2* d_w, const float2* d_x, float2* d_y, uint32_t out_channels, uint32_t in_channels,
                 uint64_t samples, hipStream_t stream) {
  require(out_channels >= 1 && out_channels <= kMixMaxChannels && in_channels >= 1 &&
              in_channels <= kMixMaxChannels,
          "channel counts must be in [1, 64]");
  if (samples == 0) return;
  require(samples <= (uint64_t{1} << 36), "too many samples");  // grid x stays below 2^31
  if (tune_get(kTuneMixScheme) == 1) {
    if (out_channels <= 16)
      mix_mfma_cols<2>(d_w, d_x, d_y, out_channels, in_channels, samples, stream);
    else if (out_channels <= 32)
      mix_mfma_cols<4>(d_w, d_x, d_y, out_channels, in_channels, samples, stream);
    else
      mix_mfma_cols<8>(d_w, d_x, d_y, out_channels, in_channels, samples, stream);
    return;
  }
  if (in_channels <= 4)
    launch_mix<4>(d_w, d_x, d_y, out_channels, in_channels, samples, stream);
  else if (in_channels <= 8)
    launch_mix<8>(d_w, d_x, d_y, out_channels, in_channels, samples, stream);
  else if (in_channels <= 16)
    launch_mix<16>(d_w, d_x, d_y, out_channels, in_channels, samples, stream);
  else if (in_channels <= 32)
    launch_mix<32>(d_w, d_x, d_y, out_channels, in_channels, samples, stream);
  else
    launch_mix<64>(d_w, d_x, d_y, out_channels, in_channels, samples, stream);
}

void phase_correct(const float2* d_x, const float2* d_p, float2* d_y, uint32_t channels, uint32_t rows,
                   uint32_t cols, hipStream_t stream) {
  require(cols >= 1, "cols must be positive");
  require(channels <= 65535, "too many channels");  // grid z
  if (uint64_t{channels} * rows == 0) return;
  const dim3 grid(ceil_div(cols, 256), std::min<uint32_t>(rows, 1024), channels);
  ProfScope ps("phase_correct", stream);
  hipLaunchKernelGGL(phase_correct_kernel, grid, dim3(256), 0, stream, d_x, d_p, d_y, rows, cols);
  KSJ_LAUNCH_CHECK();
}

void readout_crop(const float2* d_x, float2* d_y, uint64_t lines, uint32_t in_cols, uint32_t out_cols,
                  uint32_t offset, hipStream_t stream) {
  require(out_cols >= 1 && uint64_t{offset} + out_cols <= in_cols, "crop range must lie inside the readout");
  if (lines == 0) return;
  const dim3 grid(ceil_div(out_cols, 256), static_cast<unsigned>(std::min<uint64_t>(lines, 32768)));
  ProfScope ps("readout_crop", stream);
  hipLaunchKernelGGL(readout_crop_kernel, grid, dim3(256), 0, stream, d_x, d_y, lines, in_cols, out_cols, offset);
  KSJ_LAUNCH_CHECK();
}

}  // namespace ksj::gpu
