// LDS Stockham FFT kernels for gfx950 (wave64).
//
// One workgroup transforms `lpb` lines of length N held in LDS.  Each line is
// served by T threads; every pass reads its radix-R butterfly inputs into
// registers, barriers, writes the autosorted outputs back in place and
// barriers again (Stockham autosort: output is in natural order, no bit
// reversal).  Twiddles come from an FP64-generated table staged in LDS.
// Loads and stores are mapped so that consecutive threads touch consecutive
// addresses in global memory for both contiguous rows and strided columns.
#include "fft.hpp"

#include "fused.hpp"

#include <map>
#include <vector>
#include <mutex>
#include <utility>

namespace ksj::gpu {
namespace {

// ----------------------------------------------------------------------------- radix schedules
// T = threads per line.  Every pass satisfies (N / R) % T == 0, so each thread
// owns exactly N / (R T) butterflies in every pass.  Three schedules per length
// (tuning key fft_schedule): 0 the default, 1 the same radices on half the threads
// (twice the butterflies per thread), 2 a radix-4-heavy factorisation.  They differ
// in rounding only (each is an exact factorisation of the DFT).
template <int N, int V> struct Shape;
#define KSJ_SHAPE(NN, VV, TT, ...)                                          \
  template <> struct Shape<NN, VV> {                                        \
    static constexpr int T = TT;                                            \
    static constexpr int R[] = {__VA_ARGS__};                               \
    static constexpr int P = sizeof(R) / sizeof(int);                       \
  };
KSJ_SHAPE(64, 0, 8, 8, 8)
KSJ_SHAPE(64, 1, 4, 8, 8)
KSJ_SHAPE(64, 2, 16, 4, 4, 4)
KSJ_SHAPE(128, 0, 16, 8, 4, 4)
KSJ_SHAPE(128, 1, 8, 8, 4, 4)
KSJ_SHAPE(128, 2, 32, 4, 4, 4, 2)
KSJ_SHAPE(192, 0, 16, 4, 4, 4, 3)
KSJ_SHAPE(192, 1, 8, 4, 4, 4, 3)
KSJ_SHAPE(192, 2, 16, 4, 4, 3, 4)
KSJ_SHAPE(256, 0, 32, 8, 8, 4)
KSJ_SHAPE(256, 1, 16, 8, 8, 4)
KSJ_SHAPE(256, 2, 64, 4, 4, 4, 4)
KSJ_SHAPE(320, 0, 16, 4, 4, 4, 5)
KSJ_SHAPE(320, 1, 8, 4, 4, 4, 5)
KSJ_SHAPE(320, 2, 16, 5, 4, 4, 4)
KSJ_SHAPE(384, 0, 16, 4, 4, 8, 3)
KSJ_SHAPE(384, 1, 8, 4, 4, 8, 3)
KSJ_SHAPE(384, 2, 16, 8, 4, 4, 3)
KSJ_SHAPE(512, 0, 64, 8, 8, 8)
KSJ_SHAPE(512, 1, 32, 8, 8, 8)
KSJ_SHAPE(512, 2, 64, 8, 4, 4, 4)
KSJ_SHAPE(640, 0, 32, 4, 4, 4, 2, 5)
KSJ_SHAPE(640, 1, 16, 4, 4, 4, 2, 5)
KSJ_SHAPE(640, 2, 16, 8, 4, 4, 5)
KSJ_SHAPE(768, 0, 32, 8, 8, 4, 3)
KSJ_SHAPE(768, 1, 16, 8, 8, 4, 3)
KSJ_SHAPE(768, 2, 64, 4, 4, 4, 4, 3)
KSJ_SHAPE(1024, 0, 64, 8, 8, 4, 4)
KSJ_SHAPE(1024, 1, 32, 8, 8, 4, 4)
KSJ_SHAPE(1024, 2, 64, 4, 4, 4, 4, 4)
#undef KSJ_SHAPE

template <int N, int V>
constexpr bool shape_ok() {
  int prod = 1;
  for (int i = 0; i < Shape<N, V>::P; ++i) {
    prod *= Shape<N, V>::R[i];
    if ((N / Shape<N, V>::R[i]) % Shape<N, V>::T != 0) return false;
  }
  return prod == N;
}

// LDS element index of line element i: one pad element per 16 spreads the strided
// Stockham writes (stride R in the first pass, R * Ns blocks in the next) over the
// banks instead of mapping a whole wave onto a few.
__device__ constexpr int pad(int i) { return i + (i >> 4); }
// LDS stride between lines: padded length plus one (odd) so lines start on different banks.
template <int N>
constexpr int line_stride() {
  return N + N / 16 + 1;
}

__device__ inline float2 mul_i(float2 z, bool inv) {  // z * (+i) for inverse, z * (-i) for forward
  return inv ? make_float2(-z.y, z.x) : make_float2(z.y, -z.x);
}

template <int R, bool INV> struct Dft;

template <bool INV> struct Dft<2, INV> {
  __device__ static inline void run(float2* v) {
    const float2 a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
  }
};

template <bool INV> struct Dft<4, INV> {
  __device__ static inline void run(float2* v) {
    const float2 a0 = cadd(v[0], v[2]), a1 = csub(v[0], v[2]);
    const float2 b0 = cadd(v[1], v[3]), b1 = mul_i(csub(v[1], v[3]), INV);
    v[0] = cadd(a0, b0);
    v[2] = csub(a0, b0);
    v[1] = cadd(a1, b1);
    v[3] = csub(a1, b1);
  }
};

template <bool INV> struct Dft<8, INV> {
  __device__ static inline void run(float2* v) {
    constexpr float h = 0.70710678118654752440f;
    float2 e[4] = {v[0], v[2], v[4], v[6]};
    float2 o[4] = {v[1], v[3], v[5], v[7]};
    Dft<4, INV>::run(e);
    Dft<4, INV>::run(o);
    const float s = INV ? 1.0f : -1.0f;
    const float2 w1 = make_float2(h, s * h), w3 = make_float2(-h, s * h);
    const float2 t0 = o[0], t1 = cmul(o[1], w1), t2 = mul_i(o[2], INV), t3 = cmul(o[3], w3);
    v[0] = cadd(e[0], t0);
    v[4] = csub(e[0], t0);
    v[1] = cadd(e[1], t1);
    v[5] = csub(e[1], t1);
    v[2] = cadd(e[2], t2);
    v[6] = csub(e[2], t2);
    v[3] = cadd(e[3], t3);
    v[7] = csub(e[3], t3);
  }
};

template <bool INV> struct Dft<3, INV> {
  __device__ static inline void run(float2* v) {
    constexpr float r3 = 0.86602540378443864676f;
    const float2 t1 = cadd(v[1], v[2]), t2 = csub(v[1], v[2]);
    const float2 m = make_float2(v[0].x - 0.5f * t1.x, v[0].y - 0.5f * t1.y);
    const float2 n = mul_i(cscale(t2, r3), INV);
    v[0] = cadd(v[0], t1);
    v[1] = cadd(m, n);
    v[2] = csub(m, n);
  }
};

template <bool INV> struct Dft<5, INV> {
  __device__ static inline void run(float2* v) {
    constexpr float c1 = 0.30901699437494742410f, c2 = -0.80901699437494742410f;
    constexpr float s1 = 0.95105651629515357212f, s2 = 0.58778525229247312917f;
    const float2 t1 = cadd(v[1], v[4]), t2 = cadd(v[2], v[3]);
    const float2 t3 = csub(v[1], v[4]), t4 = csub(v[2], v[3]);
    const float2 a1 = make_float2(v[0].x + c1 * t1.x + c2 * t2.x, v[0].y + c1 * t1.y + c2 * t2.y);
    const float2 a2 = make_float2(v[0].x + c2 * t1.x + c1 * t2.x, v[0].y + c2 * t1.y + c1 * t2.y);
    const float2 b1 = mul_i(make_float2(s1 * t3.x + s2 * t4.x, s1 * t3.y + s2 * t4.y), INV);
    const float2 b2 = mul_i(make_float2(s2 * t3.x - s1 * t4.x, s2 * t3.y - s1 * t4.y), INV);
    v[0] = cadd(v[0], cadd(t1, t2));
    v[1] = cadd(a1, b1);
    v[4] = csub(a1, b1);
    v[2] = cadd(a2, b2);
    v[3] = csub(a2, b2);
  }
};

template <int N, int V, bool INV, int P, int Ns>
__device__ inline void passes(float2* line, const float2* tw, int t) {
  if constexpr (P < Shape<N, V>::P) {
    constexpr int T = Shape<N, V>::T;
    constexpr int R = Shape<N, V>::R[P];
    constexpr int B = N / R / T;
    static_assert((N / R) % T == 0, "radix schedule must split evenly over the line's threads");
    float2 v[B][R];
#pragma unroll
    for (int b = 0; b < B; ++b) {
      const int j = t + b * T;
      const int k = j % Ns;
#pragma unroll
      for (int r = 0; r < R; ++r) v[b][r] = line[pad(j + r * (N / R))];
      if constexpr (Ns > 1) {
#pragma unroll
        for (int r = 1; r < R; ++r) {
          float2 w = tw[k * r * (N / (Ns * R))];
          if (!INV) w.y = -w.y;
          v[b][r] = cmul(v[b][r], w);
        }
      }
      Dft<R, INV>::run(v[b]);
    }
    __syncthreads();
#pragma unroll
    for (int b = 0; b < B; ++b) {
      const int j = t + b * T;
      const int k = j % Ns;
      const int base = (j / Ns) * Ns * R + k;
#pragma unroll
      for (int r = 0; r < R; ++r) line[pad(base + r * Ns)] = v[b][r];
    }
    __syncthreads();
    passes<N, V, INV, P + 1, Ns * R>(line, tw, t);
  }
}

__device__ inline uint64_t line_base(const FftLines& L, uint64_t line) {
  const uint64_t p = line / L.lines, q = line - p * L.lines;
  return p * L.plane_stride + q * L.line_stride;
}

// Loads `lpb` lines into LDS (line stride ls), rotated by in_offset.  lpb is a power
// of two and the block has lpb * T threads, so every thread moves exactly N / T
// elements: all of its global loads are issued before the first LDS write, so the
// loads of one thread overlap instead of paying one memory latency each.
template <int N, int V>
__device__ inline void load_lines(const FftArgs& a, float2* buf, int ls, int lpb_log2, const uint64_t* src_base) {
  constexpr int E = N / Shape<N, V>::T;
  const int lpb = 1 << lpb_log2;
  const uint32_t es = static_cast<uint32_t>(a.src.elem_stride);
  float2 v[E];
  int dst[E];
#pragma unroll
  for (int i = 0; i < E; ++i) {
    const int idx = threadIdx.x + i * blockDim.x;
    int ly, e;
    if (a.load_lines_fast) {
      ly = idx & (lpb - 1);
      e = idx >> lpb_log2;
    } else {
      ly = idx / N;
      e = idx - ly * N;
    }
    const uint64_t base = src_base[ly];
    const bool live = base != ~0ull;
    uint32_t src = static_cast<uint32_t>(e) + a.in_offset;
    if (src >= static_cast<uint32_t>(N)) src -= N;
    // Unconditional load, so no branch separates the loads and all of them are in
    // flight together: a line past the end transforms a copy of line 0, and its
    // outputs are never stored.
    // Every input element of a pass is read once: non-temporal, out of the caches' way.
    typedef float F2v __attribute__((ext_vector_type(2)));
    const F2v t = __builtin_nontemporal_load(
        reinterpret_cast<const F2v*>(a.in + (live ? base : 0ull) + static_cast<uint64_t>(src * es)));
    v[i] = make_float2(t.x, t.y);
    dst[i] = ly * ls + pad(e);
  }
#pragma unroll
  for (int i = 0; i < E; ++i) buf[dst[i]] = v[i];
}

// Stores the n_out kept outputs of every line, scaled by a.scale x out_weight[e] x
// line_weight[line]; the weights were staged in LDS (ow, lw), so the stores issue
// back to back.
template <int N, int V>
__device__ inline void store_lines(const FftArgs& a, const float2* buf, int ls, int lpb_log2, const float* ow,
                                   const float* lw, const uint64_t* dst_base) {
  constexpr int E = N / Shape<N, V>::T;
  const int lpb = 1 << lpb_log2;
  const int nout = static_cast<int>(a.n_out);
  const int count = lpb * nout;
  const uint32_t es = static_cast<uint32_t>(a.dst.elem_stride);
#pragma unroll
  for (int i = 0; i < E; ++i) {
    const int idx = threadIdx.x + i * blockDim.x;
    if (idx >= count) break;
    int ly, e;
    if (a.store_lines_fast) {
      ly = idx & (lpb - 1);
      e = idx >> lpb_log2;
    } else {
      ly = idx / nout;
      e = idx - ly * nout;
    }
    const uint64_t base = dst_base[ly];
    if (base == ~0ull) continue;
    uint32_t bin = static_cast<uint32_t>(e) + a.out_offset;
    if (bin >= static_cast<uint32_t>(N)) bin -= N;
    a.out[base + static_cast<uint64_t>(static_cast<uint32_t>(e) * es)] =
        cscale(buf[ly * ls + pad(static_cast<int>(bin))], ow[e] * lw[ly]);
  }
}

template <int N, int V, bool INV>
__device__ inline void fft_lds_block(const FftArgs& a, const float2* __restrict__ tw_global, int lpb_log2,
                                     uint64_t block) {
  static_assert(shape_ok<N, V>(), "radix schedule must multiply to N and split evenly over the line's threads");
  constexpr int T = Shape<N, V>::T;
  constexpr int LS = line_stride<N>();
  extern __shared__ float2 smem[];
  __shared__ uint64_t src_base[64], dst_base[64];
  __shared__ float ow[N], lw[64];  // output weights (scale x out_weight) and per-line weights
  const int lpb = 1 << lpb_log2;
  float2* tw = smem;
  float2* buf = smem + N;
  for (int i = threadIdx.x; i < N; i += blockDim.x) {
    tw[i] = tw_global[i];
    ow[i] = (i < static_cast<int>(a.n_out) && a.out_weight) ? a.scale * a.out_weight[i] : a.scale;
  }
  const uint64_t total = static_cast<uint64_t>(a.src.planes) * a.src.lines;
  const uint64_t first = block * lpb;
  for (int ly = threadIdx.x; ly < lpb; ly += blockDim.x) {
    const uint64_t line = first + ly;
    src_base[ly] = line < total ? line_base(a.src, line) : ~0ull;
    dst_base[ly] = line < total ? line_base(a.dst, line) : ~0ull;
    lw[ly] = a.line_weight ? a.line_weight[line % a.src.lines] : 1.0f;
  }
  __syncthreads();
  load_lines<N, V>(a, buf, LS, lpb_log2, src_base);
  __syncthreads();
  const int ly = threadIdx.x / T, t = threadIdx.x % T;
  passes<N, V, INV, 0, 1>(buf + ly * LS, tw, t);
  store_lines<N, V>(a, buf, LS, lpb_log2, ow, lw, dst_base);
}

template <int N, int V, bool INV>
__global__ __launch_bounds__(1024) void fft_lds_kernel(FftArgs a, const float2* __restrict__ tw_global, int lpb_log2) {
  fft_lds_block<N, V, INV>(a, tw_global, lpb_log2, blockIdx.x);
}

// Fused items of one length: the table's blocks run fft_lds_block of their item.
template <int N, int V, bool INV>
__global__ __launch_bounds__(1024) void fft_fused_kernel(const FusedTable<FftArgs>* t,
                                                         const float2* __restrict__ tw_global, int lpb_log2) {
  const uint32_t item = fused_item(t, blockIdx.x);
  const FftArgs a = fused_args(t, item);
  fft_lds_block<N, V, INV>(a, tw_global, lpb_log2, blockIdx.x - fused_first(t, item));
}

// Any length: direct DFT per output bin against the twiddle table (used for
// lengths without a radix schedule, e.g. odd or small reference test sizes).
template <int DUMMY>
__global__ __launch_bounds__(256) void dft_direct_kernel(FftArgs a, const float2* __restrict__ tw_global, int lpb) {
  extern __shared__ float2 smem[];
  const int n = static_cast<int>(a.n);
  const int ls = n + 1;
  float2* tw = smem;
  float2* buf = smem + n;
  float2* res = buf + lpb * ls;
  for (int i = threadIdx.x; i < n; i += blockDim.x) tw[i] = tw_global[i];
  const uint64_t total = static_cast<uint64_t>(a.src.planes) * a.src.lines;
  const uint64_t first = static_cast<uint64_t>(blockIdx.x) * lpb;
  for (int idx = threadIdx.x; idx < lpb * n; idx += blockDim.x) {
    const int ly = idx / n, e = idx % n;
    const uint64_t line = first + ly;
    float2 v = make_float2(0.f, 0.f);
    if (line < total) {
      uint32_t src = static_cast<uint32_t>(e) + a.in_offset;
      if (src >= static_cast<uint32_t>(n)) src -= n;
      v = a.in[line_base(a.src, line) + src * a.src.elem_stride];
    }
    buf[ly * ls + e] = v;
  }
  __syncthreads();
  for (int idx = threadIdx.x; idx < lpb * n; idx += blockDim.x) {
    const int ly = idx / n, k = idx % n;
    float2 acc = make_float2(0.f, 0.f);
    for (int j = 0; j < n; ++j) {
      float2 w = tw[(static_cast<int64_t>(j) * k) % n];
      if (!a.inverse) w.y = -w.y;
      acc = cadd(acc, cmul(buf[ly * ls + j], w));
    }
    res[ly * ls + k] = acc;
  }
  __syncthreads();
  for (int idx = threadIdx.x; idx < lpb * static_cast<int>(a.n_out); idx += blockDim.x) {
    const int ly = idx / a.n_out, e = idx % a.n_out;
    const uint64_t line = first + ly;
    if (line >= total) continue;
    uint32_t bin = static_cast<uint32_t>(e) + a.out_offset;
    if (bin >= static_cast<uint32_t>(n)) bin -= n;
    float w = a.scale;
    if (a.out_weight) w *= a.out_weight[e];
    if (a.line_weight) w *= a.line_weight[line % a.src.lines];
    a.out[line_base(a.dst, line) + static_cast<uint64_t>(e) * a.dst.elem_stride] = cscale(res[ly * ls + bin], w);
  }
}

__global__ void twiddle_kernel(float2* tw, uint32_t n) {
  const uint32_t m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= n) return;
  double s, c;
  sincospi(2.0 * static_cast<double>(m) / static_cast<double>(n), &s, &c);
  tw[m] = make_float2(static_cast<float>(c), static_cast<float>(s));
}

int max_dynamic_lds() {
  static int cached = -1;
  if (cached < 0) {
    int dev = 0, v = 0;
    if (hipGetDevice(&dev) == hipSuccess &&
        hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) == hipSuccess && v > 0)
      cached = v;
    else
      cached = 65536;
  }
  return cached;
}

// Lines per block (log2) of a launch: the tuning key or a per-mapping default, capped
// by 1024 threads and by the LDS.
template <int N, int V>
int lines_per_block_log2(const FftArgs& a, int lpb) {
  constexpr int T = Shape<N, V>::T;
  if (lpb <= 0) lpb = static_cast<int>(tune_get(kTuneFftLinesPerBlock));
  if (lpb <= 0) lpb = (a.load_lines_fast || a.store_lines_fast) ? 16 : (T >= 64 ? 4 : 256 / T);
  while (lpb * T > 1024) lpb /= 2;
  // LDS: twiddles + lpb padded lines, next to the static per-block tables
  // (line bases, output and line weights).
  constexpr int kStatic = 64 * 2 * 8 + N * 4 + 64 * 4;
  const int lds_cap = static_cast<int>(((max_dynamic_lds() - kStatic) / sizeof(float2) - N) / line_stride<N>());
  while (lpb > 1 && lpb > lds_cap) lpb /= 2;
  int lpb_log2 = 0;
  while ((1 << (lpb_log2 + 1)) <= lpb) ++lpb_log2;
  return lpb_log2;
}

template <int N, int V>
size_t fft_shmem(int lpb_log2) {
  return (static_cast<size_t>(N) + (static_cast<size_t>(1) << lpb_log2) * line_stride<N>()) * sizeof(float2);
}

template <int N, int V>
void launch_fast(const FftArgs& a, const float2* tw, hipStream_t stream, int lpb) {
  constexpr int T = Shape<N, V>::T;
  const int lpb_log2 = lines_per_block_log2<N, V>(a, lpb);
  lpb = 1 << lpb_log2;
  const uint64_t total = static_cast<uint64_t>(a.src.planes) * a.src.lines;
  const size_t shmem = fft_shmem<N, V>(lpb_log2);
  const unsigned grid = ceil_div(total, static_cast<uint64_t>(lpb));
  if (a.inverse) {
    if (shmem > 65536)
      KSJ_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(&fft_lds_kernel<N, V, true>),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(shmem)));
    hipLaunchKernelGGL((fft_lds_kernel<N, V, true>), dim3(grid), dim3(lpb * T), shmem, stream, a, tw, lpb_log2);
  } else {
    if (shmem > 65536)
      KSJ_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(&fft_lds_kernel<N, V, false>),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(shmem)));
    hipLaunchKernelGGL((fft_lds_kernel<N, V, false>), dim3(grid), dim3(lpb * T), shmem, stream, a, tw, lpb_log2);
  }
  KSJ_LAUNCH_CHECK();
}

template <int N, int V>
void launch_fused(const FftArgs* const* items, uint32_t count, const float2* tw, void* d_table, hipStream_t stream) {
  constexpr int T = Shape<N, V>::T;
  const int lpb_log2 = lines_per_block_log2<N, V>(*items[0], 0);
  const uint64_t lpb = 1ull << lpb_log2;
  FusedTableBuilder<FftArgs> b;
  for (uint32_t i = 0; i < count; ++i) {
    require(items[i]->inverse == items[0]->inverse && items[i]->load_lines_fast == items[0]->load_lines_fast &&
              items[i]->store_lines_fast == items[0]->store_lines_fast,
            "fused FFT items must share direction and thread mapping");
    b.add(*items[i], ceil_div(static_cast<uint64_t>(items[i]->src.planes) * items[i]->src.lines, lpb));
  }
  b.upload(d_table, stream);
  const auto* t = static_cast<const FusedTable<FftArgs>*>(d_table);
  const size_t shmem = fft_shmem<N, V>(lpb_log2);
  const dim3 grid(static_cast<uint32_t>(b.blocks())), block(static_cast<uint32_t>(lpb * T));
  if (items[0]->inverse) {
    if (shmem > 65536)
      KSJ_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(&fft_fused_kernel<N, V, true>),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(shmem)));
    hipLaunchKernelGGL((fft_fused_kernel<N, V, true>), grid, block, shmem, stream, t, tw, lpb_log2);
  } else {
    if (shmem > 65536)
      KSJ_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(&fft_fused_kernel<N, V, false>),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(shmem)));
    hipLaunchKernelGGL((fft_fused_kernel<N, V, false>), grid, block, shmem, stream, t, tw, lpb_log2);
  }
  KSJ_LAUNCH_CHECK();
}

// The radix schedule of length N: the tuning key, where 3 picks per length -- half the
// threads per line (1) at 512, which overlaps better with the other lanes of a served
// stream (+1.4 % served images/s on the 256x256 radial stream; a serial per-scan call is
// about 2.5 % slower with it), and schedule 0 at every other length, where 1 is slower.
template <int N> int fft_schedule() {
  const int v = static_cast<int>(tune_get(kTuneFftSchedule));
  return v == 3 ? (N == 512 ? 1 : 0) : v;
}

template <int N>
void launch_fused(const FftArgs* const* items, uint32_t count, const float2* tw, void* d_table, hipStream_t stream) {
  switch (fft_schedule<N>()) {
    case 1: return launch_fused<N, 1>(items, count, tw, d_table, stream);
    case 2: return launch_fused<N, 2>(items, count, tw, d_table, stream);
    default: return launch_fused<N, 0>(items, count, tw, d_table, stream);
  }
}

template <int N>
void launch_fast(const FftArgs& a, const float2* tw, hipStream_t stream, int lpb) {
  switch (fft_schedule<N>()) {
    case 1: return launch_fast<N, 1>(a, tw, stream, lpb);
    case 2: return launch_fast<N, 2>(a, tw, stream, lpb);
    default: return launch_fast<N, 0>(a, tw, stream, lpb);
  }
}

struct TwKey {
  int device;
  uint32_t n;
  bool operator<(const TwKey& o) const { return device != o.device ? device < o.device : n < o.n; }
};

}  // namespace

bool fft_fast_length(uint32_t n) {
  switch (n) {
    case 64: case 128: case 192: case 256: case 320: case 384: case 512: case 640: case 768: case 1024:
      return true;
    default:
      return false;
  }
}

const float2* fft_twiddles(uint32_t n, hipStream_t stream) {
  static std::mutex mu;
  static std::map<TwKey, float2*> cache;
  int device = 0;
  KSJ_HIP(hipGetDevice(&device));
  std::lock_guard<std::mutex> lock(mu);
  auto it = cache.find({device, n});
  if (it != cache.end()) return it->second;
  float2* tw = nullptr;
  KSJ_HIP(hipMalloc(&tw, sizeof(float2) * n));
  hipLaunchKernelGGL(twiddle_kernel, dim3(ceil_div(n, 256)), dim3(256), 0, stream, tw, n);
  KSJ_LAUNCH_CHECK();
  KSJ_HIP(hipStreamSynchronize(stream));
  cache.emplace(TwKey{device, n}, tw);
  return tw;
}

void fft_launch(const FftArgs& a, hipStream_t stream, int lpb) {
  require(a.n >= 1 && a.n <= 4096, "FFT length must be in [1, 4096]");
  require(a.n_out >= 1 && a.n_out <= a.n, "FFT pruned output count must be in [1, n]");
  require(a.in_offset < a.n && a.out_offset < a.n, "FFT rotations must be below the length");
  require(a.src.planes == a.dst.planes && a.src.lines == a.dst.lines, "FFT source/destination line counts differ");
  const uint64_t total = static_cast<uint64_t>(a.src.planes) * a.src.lines;
  if (total == 0) return;
  const float2* tw = fft_twiddles(a.n, stream);
  switch (a.n) {
    case 64: return launch_fast<64>(a, tw, stream, lpb);
    case 128: return launch_fast<128>(a, tw, stream, lpb);
    case 192: return launch_fast<192>(a, tw, stream, lpb);
    case 256: return launch_fast<256>(a, tw, stream, lpb);
    case 320: return launch_fast<320>(a, tw, stream, lpb);
    case 384: return launch_fast<384>(a, tw, stream, lpb);
    case 512: return launch_fast<512>(a, tw, stream, lpb);
    case 640: return launch_fast<640>(a, tw, stream, lpb);
    case 768: return launch_fast<768>(a, tw, stream, lpb);
    case 1024: return launch_fast<1024>(a, tw, stream, lpb);
    default: break;
  }
  int l = (a.load_lines_fast || a.store_lines_fast) ? 8 : 1;
  while (l > 1 && (static_cast<size_t>(a.n) + 2ull * l * (a.n + 1)) * sizeof(float2) > 60 * 1024) l /= 2;
  const size_t shmem = (static_cast<size_t>(a.n) + 2ull * l * (a.n + 1)) * sizeof(float2);
  require(shmem <= 160 * 1024, "direct DFT line does not fit in LDS");
  hipLaunchKernelGGL(dft_direct_kernel<0>, dim3(ceil_div(total, l)), dim3(256), shmem, stream, a, tw, l);
  KSJ_LAUNCH_CHECK();
}

size_t fft_fused_table_bytes(uint32_t count) { return static_cast<size_t>(count) * fused_table_bytes<FftArgs>(); }

void fft_launch_fused(const FftArgs* args, uint32_t count, void* d_tables, hipStream_t stream) {
  require(count <= kMaxFusedItems, "fused FFT takes at most 64 items");
  std::vector<uint32_t> done(count, 0);
  char* tab = static_cast<char*>(d_tables);
  for (uint32_t i = 0; i < count; ++i) {
    if (done[i]) continue;
    const FftArgs& a = args[i];
    require(a.n >= 1 && a.n <= 4096 && a.n_out >= 1 && a.n_out <= a.n && a.in_offset < a.n && a.out_offset < a.n &&
              a.src.planes == a.dst.planes && a.src.lines == a.dst.lines,
            "invalid fused FFT item");
    std::vector<const FftArgs*> cls;
    for (uint32_t j = i; j < count; ++j)
      if (!done[j] && args[j].n == a.n) {
        done[j] = 1;
        if (static_cast<uint64_t>(args[j].src.planes) * args[j].src.lines > 0) cls.push_back(&args[j]);
      }
    if (cls.empty()) continue;
    if (!fft_fast_length(a.n)) {  // direct-DFT lengths: one launch per item
      for (const FftArgs* p : cls) fft_launch(*p, stream);
      continue;
    }
    const float2* tw = fft_twiddles(a.n, stream);
    const uint32_t nc = static_cast<uint32_t>(cls.size());
    switch (a.n) {
      case 64: launch_fused<64>(cls.data(), nc, tw, tab, stream); break;
      case 128: launch_fused<128>(cls.data(), nc, tw, tab, stream); break;
      case 192: launch_fused<192>(cls.data(), nc, tw, tab, stream); break;
      case 256: launch_fused<256>(cls.data(), nc, tw, tab, stream); break;
      case 320: launch_fused<320>(cls.data(), nc, tw, tab, stream); break;
      case 384: launch_fused<384>(cls.data(), nc, tw, tab, stream); break;
      case 512: launch_fused<512>(cls.data(), nc, tw, tab, stream); break;
      case 640: launch_fused<640>(cls.data(), nc, tw, tab, stream); break;
      case 768: launch_fused<768>(cls.data(), nc, tw, tab, stream); break;
      default: launch_fused<1024>(cls.data(), nc, tw, tab, stream); break;
    }
    tab += fused_table_bytes<FftArgs>();
  }
}

}  // namespace ksj::gpu
