// K-space conditioning kernels.  All three are streaming passes over channel-major
// k-space, bound by HBM bandwidth except the channel mix at high channel counts
// (8 * in flops per complex input value read).
#include "common.hpp"
#include "condition.hpp"

namespace ksj::gpu {
namespace {

constexpr int kMixBlock = 256;
constexpr uint32_t kMixGroup = 16;  // outputs per block along grid y

// One thread per sample: the sample's IN_MAX input values (zero past in_ch) stay in
// registers while the block's group of outputs is formed from them.  The matrix row is
// the same for every lane, so its values arrive through scalar loads.
template <int IN_MAX>
__global__ __launch_bounds__(kMixBlock) void channel_mix_kernel(const float2* __restrict__ w,
                                                                const float2* __restrict__ x,
                                                                float2* __restrict__ y, uint32_t out_ch,
                                                                uint32_t in_ch, uint64_t samples) {
  const uint64_t s = static_cast<uint64_t>(blockIdx.x) * kMixBlock + threadIdx.x;
  if (s >= samples) return;
  const uint32_t o0 = blockIdx.y * kMixGroup;
  const uint32_t o1 = min(out_ch, o0 + kMixGroup);
  float2 v[IN_MAX];
#pragma unroll
  for (int i = 0; i < IN_MAX; ++i)
    v[i] = static_cast<uint32_t>(i) < in_ch ? x[static_cast<uint64_t>(i) * samples + s] : make_float2(0.f, 0.f);
  for (uint32_t o = o0; o < o1; ++o) {
    const float2* row = w + static_cast<uint64_t>(o) * in_ch;
    float re = 0.f, im = 0.f;
#pragma unroll
    for (int i = 0; i < IN_MAX; ++i) {
      if (static_cast<uint32_t>(i) < in_ch) {
        const float2 c = row[i];
        re = fmaf(c.x, v[i].x, re);
        re = fmaf(-c.y, v[i].y, re);
        im = fmaf(c.x, v[i].y, im);
        im = fmaf(c.y, v[i].x, im);
      }
    }
    __builtin_nontemporal_store(re, &y[static_cast<uint64_t>(o) * samples + s].x);
    __builtin_nontemporal_store(im, &y[static_cast<uint64_t>(o) * samples + s].y);
  }
}

template <int IN_MAX>
void launch_mix(const float2* w, const float2* x, float2* y, uint32_t out_ch, uint32_t in_ch, uint64_t samples,
                hipStream_t stream) {
  const dim3 grid(ceil_div(samples, kMixBlock), ceil_div(out_ch, kMixGroup));
  hipLaunchKernelGGL(channel_mix_kernel<IN_MAX>, grid, dim3(kMixBlock), 0, stream, w, x, y, out_ch, in_ch, samples);
  KSJ_LAUNCH_CHECK();
}

__global__ void phase_correct_kernel(const float2* __restrict__ x, const float2* __restrict__ p,
                                     float2* __restrict__ y, uint64_t count, uint32_t cols, uint64_t per_channel) {
  // Each product and each sum rounds to float on its own, as in the reference.
#pragma clang fp contract(off)
  const uint64_t e = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= count) return;
  const uint64_t c = e / per_channel;
  const uint32_t k = static_cast<uint32_t>(e % cols);
  const float2 a = x[e];
  const float2 b = p[c * cols + k];
  const float ac = a.x * b.x, bd = a.y * b.y, ad = a.x * b.y, bc = a.y * b.x;
  y[e] = make_float2(ac - bd, ad + bc);
}

__global__ void readout_crop_kernel(const float2* __restrict__ x, float2* __restrict__ y, uint64_t count,
                                    uint32_t in_cols, uint32_t out_cols, uint32_t offset) {
  const uint64_t e = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= count) return;
  const uint64_t l = e / out_cols;
  const uint32_t k = static_cast<uint32_t>(e % out_cols);
  y[e] = x[l * in_cols + offset + k];
}

}  // namespace

void channel_mix(const float2* d_w, const float2* d_x, float2* d_y, uint32_t out_channels, uint32_t in_channels,
                 uint64_t samples, hipStream_t stream) {
  require(out_channels >= 1 && out_channels <= kMixMaxChannels && in_channels >= 1 &&
              in_channels <= kMixMaxChannels,
          "channel counts must be in [1, 64]");
  if (samples == 0) return;
  require(samples <= (uint64_t{1} << 36), "too many samples");  // grid x stays below 2^31
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
  require(uint64_t{rows} * cols <= (uint64_t{1} << 32) && channels <= 65536, "frame too large");
  const uint64_t per_channel = uint64_t{rows} * cols;
  const uint64_t count = per_channel * channels;
  require(count <= (uint64_t{1} << 36), "frame too large");  // grid x stays below 2^31
  if (count == 0) return;
  hipLaunchKernelGGL(phase_correct_kernel, dim3(ceil_div(count, 256)), dim3(256), 0, stream, d_x, d_p, d_y, count,
                     cols, per_channel);
  KSJ_LAUNCH_CHECK();
}

void readout_crop(const float2* d_x, float2* d_y, uint64_t lines, uint32_t in_cols, uint32_t out_cols,
                  uint32_t offset, hipStream_t stream) {
  require(out_cols >= 1 && uint64_t{offset} + out_cols <= in_cols, "crop range must lie inside the readout");
  require(lines <= (uint64_t{1} << 36) / out_cols, "frame too large");  // grid x stays below 2^31
  const uint64_t count = lines * out_cols;
  if (count == 0) return;
  hipLaunchKernelGGL(readout_crop_kernel, dim3(ceil_div(count, 256)), dim3(256), 0, stream, d_x, d_y, count, in_cols,
                     out_cols, offset);
  KSJ_LAUNCH_CHECK();
}

}  // namespace ksj::gpu
