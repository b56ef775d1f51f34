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
#include <type_traits>
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

// Products of the form a b +- c d are written as one explicit fma and one multiply, so
// the compiler's contraction of the remaining adds has one choice only and every form
// of a pass (staged or register) rounds identically.
__device__ inline float2 tmul(float2 a, float2 b) {
  return make_float2(fmaf(a.x, b.x, -(a.y * b.y)), fmaf(a.x, b.y, a.y * b.x));
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
    const float2 t0 = o[0], t1 = tmul(o[1], w1), t2 = mul_i(o[2], INV), t3 = tmul(o[3], w3);
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
    const float2 m = make_float2(fmaf(-0.5f, t1.x, v[0].x), fmaf(-0.5f, t1.y, v[0].y));
    // m +- i r3 t2 (inverse) or m -+ i r3 t2 (forward), each component one fma: the
    // product r3 t2 is not shared between the two outputs.
    const float s = INV ? r3 : -r3;
    v[0] = cadd(v[0], t1);
    v[1] = make_float2(fmaf(-s, t2.y, m.x), fmaf(s, t2.x, m.y));
    v[2] = make_float2(fmaf(s, t2.y, m.x), fmaf(-s, t2.x, m.y));
  }
};

template <bool INV> struct Dft<5, INV> {
  __device__ static inline void run(float2* v) {
    constexpr float c1 = 0.30901699437494742410f, c2 = -0.80901699437494742410f;
    constexpr float s1 = 0.95105651629515357212f, s2 = 0.58778525229247312917f;
    const float2 t1 = cadd(v[1], v[4]), t2 = cadd(v[2], v[3]);
    const float2 t3 = csub(v[1], v[4]), t4 = csub(v[2], v[3]);
    const float2 a1 = make_float2(fmaf(c2, t2.x, fmaf(c1, t1.x, v[0].x)), fmaf(c2, t2.y, fmaf(c1, t1.y, v[0].y)));
    const float2 a2 = make_float2(fmaf(c1, t2.x, fmaf(c2, t1.x, v[0].x)), fmaf(c1, t2.y, fmaf(c2, t1.y, v[0].y)));
    const float2 b1 = mul_i(make_float2(fmaf(s1, t3.x, s2 * t4.x), fmaf(s1, t3.y, s2 * t4.y)), INV);
    const float2 b2 = mul_i(make_float2(fmaf(s2, t3.x, -(s1 * t4.x)), fmaf(s2, t3.y, -(s1 * t4.y))), INV);
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
          v[b][r] = tmul(v[b][r], w);
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
// Element e of a line is read as zero when e lies in the line's zero span [zs.x, zs.y).
__device__ inline bool in_span(uint32_t e, uint2 zs) { return e - zs.x < zs.y - zs.x; }

// The zero span of a source plane, and whether a line's outputs are skipped.
__device__ inline uint2 plane_zero_span(const FftArgs& a, uint64_t plane) {
  if (!a.zero_span) return make_uint2(0, 0);
  return a.zero_span[a.zero_span_period ? plane % a.zero_span_period : plane];
}
__device__ inline bool plane_skipped(const FftArgs& a, uint64_t plane) {
  if (!a.skip_span) return false;
  return in_span(static_cast<uint32_t>(plane % a.skip_div), a.skip_span[plane / a.skip_div]);
}

template <int N, int V, typename Mid>
__device__ inline void load_lines(const FftArgs& a, float2* buf, int ls, int lpb_log2, const uint64_t* src_base,
                                  const uint2* zspan, Mid mid) {
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
    const uint32_t keep = in_span(src, zspan[ly]) ? 0u : ~0u;  // the zero span reads element 0, then zero
    const F2v t = __builtin_nontemporal_load(
        reinterpret_cast<const F2v*>(a.in + (live ? base : 0ull) + static_cast<uint64_t>((src & keep) * es)));
    v[i] = keep_bits(make_float2(t.x, t.y), keep);
    dst[i] = ly * ls + pad(e);
  }
  mid();  // work that overlaps the loads in flight
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
  __shared__ uint2 zspan[64];
  __shared__ float ow[N], lw[64];  // output weights (scale x out_weight) and per-line weights
  const int lpb = 1 << lpb_log2;
  float2* tw = smem;
  float2* buf = smem + N;
  const uint64_t total = static_cast<uint64_t>(a.src.planes) * a.src.lines;
  const uint64_t first = block * lpb;
  for (int ly = threadIdx.x; ly < lpb; ly += blockDim.x) {
    const uint64_t line = first + ly;
    src_base[ly] = line < total ? line_base(a.src, line) : ~0ull;
    dst_base[ly] = (line < total && !plane_skipped(a, line / a.src.lines)) ? line_base(a.dst, line) : ~0ull;
    zspan[ly] = line < total ? plane_zero_span(a, line / a.src.lines) : make_uint2(0, 0);
    const uint64_t wi = (a.line_weight_by_plane ? line / a.src.lines : line % a.src.lines) / a.line_weight_div;
    lw[ly] = (a.line_weight && line < total) ? a.line_weight[wi] : 1.0f;
  }
  __syncthreads();
  // The line loads go first; the twiddle and weight staging runs while they are in flight.
  load_lines<N, V>(a, buf, LS, lpb_log2, src_base, zspan, [&] {
    for (int i = threadIdx.x; i < N; i += blockDim.x) {
      tw[i] = tw_global[i];
      ow[i] = (i < static_cast<int>(a.n_out) && a.out_weight) ? a.scale * a.out_weight[i] : a.scale;
    }
  });
  __syncthreads();
  const int ly = threadIdx.x / T, t = threadIdx.x % T;
  passes<N, V, INV, 0, 1>(buf + ly * LS, tw, t);
  store_lines<N, V>(a, buf, LS, lpb_log2, ow, lw, dst_base);
}

// ----------------------------------------------------------------------------- register-staged form
// For launches whose loads and stores both take the same element of consecutive lines
// (lines contiguous in memory), thread (ly, t) = (tid % lpb, tid / lpb) holds elements
// t + i T (i < N / T) of line ly.  Those are exactly the inputs of its butterflies in
// the first Stockham pass and the outputs of its butterflies in the last, so the first
// pass runs on the loaded registers and the last pass's results go straight to the
// stores: two LDS round trips fewer than the staged form, and the same arithmetic
// (same butterflies, twiddles and order), so the same bits.  LDS holds the lines
// element-major, element e of line ly at e * (lpb + 1) + ly, so a wave's lanes (lines)
// touch consecutive words.

template <int N, int V, int P> struct StageNs {
  static constexpr int value = StageNs<N, V, P - 1>::value * Shape<N, V>::R[P - 1];
};
template <int N, int V> struct StageNs<N, V, 0> {
  static constexpr int value = 1;
};

// The middle passes (1 .. P-2) between the register first and last passes.
template <int N, int V, bool INV, int P>
__device__ inline void reg_mid_passes(float2* buf, int S, int ly, const float2* tw, int t) {
  if constexpr (P < Shape<N, V>::P - 1) {
    constexpr int T = Shape<N, V>::T;
    constexpr int R = Shape<N, V>::R[P];
    constexpr int B = N / R / T;
    constexpr int Ns = StageNs<N, V, P>::value;
    float2 v[B][R];
#pragma unroll
    for (int b = 0; b < B; ++b) {
      const int j = t + b * T;
      const int k = j % Ns;
#pragma unroll
      for (int r = 0; r < R; ++r) v[b][r] = buf[(j + r * (N / R)) * S + ly];
#pragma unroll
      for (int r = 1; r < R; ++r) {
        float2 w = tw[k * r * (N / (Ns * R))];
        if (!INV) w.y = -w.y;
        v[b][r] = tmul(v[b][r], w);
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
      for (int r = 0; r < R; ++r) buf[(base + r * Ns) * S + ly] = v[b][r];
    }
    __syncthreads();
    reg_mid_passes<N, V, INV, P + 1>(buf, S, ly, tw, t);
  }
}

// A thread's elements of its line in the register form.
template <int N, int V> struct RegLine {
  float2 v[N / Shape<N, V>::T];
};

// x.v[i] = element t + i T of the line on entry and bin t + i T of its transform on exit.
template <int N, int V, bool INV>
__device__ inline void reg_transform(RegLine<N, V>& x, float2* buf, int S, int ly, const float2* tw, int t) {
  constexpr int T = Shape<N, V>::T;
  constexpr int PL = Shape<N, V>::P - 1;
  static_assert(PL >= 1, "the register form needs at least two passes");
  {  // first pass: inputs j + r N / R0 = element t + T (b + r B0), no twiddles
    constexpr int R = Shape<N, V>::R[0];
    constexpr int B = N / R / T;
#pragma unroll
    for (int b = 0; b < B; ++b) {
      float2 v[R];
#pragma unroll
      for (int r = 0; r < R; ++r) v[r] = x.v[b + r * B];
      Dft<R, INV>::run(v);
      const int j = t + b * T;
#pragma unroll
      for (int r = 0; r < R; ++r) buf[(j * R + r) * S + ly] = v[r];
    }
    __syncthreads();
  }
  reg_mid_passes<N, V, INV, 1>(buf, S, ly, tw, t);
  {  // last pass: outputs j + r Ns (j < Ns) = bin t + T (b + r B)
    constexpr int R = Shape<N, V>::R[PL];
    constexpr int B = N / R / T;
    constexpr int Ns = StageNs<N, V, PL>::value;
    static_assert(Ns * R == N, "the last pass must close the transform");
#pragma unroll
    for (int b = 0; b < B; ++b) {
      const int j = t + b * T;
      float2 v[R];
#pragma unroll
      for (int r = 0; r < R; ++r) v[r] = buf[(j + r * (N / R)) * S + ly];
#pragma unroll
      for (int r = 1; r < R; ++r) {
        float2 w = tw[j * r * (N / (Ns * R))];
        if (!INV) w.y = -w.y;
        v[r] = tmul(v[r], w);
      }
      Dft<R, INV>::run(v);
#pragma unroll
      for (int r = 0; r < R; ++r) x.v[b + r * B] = v[r];
    }
  }
}

// Element-major LDS of the register form: element stride lpb + pad (tuning key
// fft_lds_pad, >= 1, so the epilogue's stride lpb + 1 staging of |x|^2 fits too).
template <int N>
__host__ __device__ constexpr int reg_lds_elems(int lpb, int pad) {
  return N * (lpb + pad);
}

// One thread's line of a register-form group: where it is read and stored, the
// span read as zero and its weight.  A line past the end reads line 0 and is never
// stored (its transform keeps every wave's loads unconditional).
struct RegLineView {
  uint64_t src, dst;  // element 0 of the source / destination line; dst ~0: not stored
  uint2 zs;
  float lw;
  uint32_t plane, line;  // the line's plane and index within the plane
  bool live;
};

// l / a.src.lines: a multiply-high by the launch's magic number where it is exact.
__device__ inline uint32_t line_plane(const FftArgs& a, uint32_t l) {
  if (a.lines_magic) return __umulhi(l, a.lines_magic);
  if (a.src.lines == 1) return l;
  return l / a.src.lines;
}

// `one_plane`: every line of the block lies in one plane, so its zero span is read once
// through the scalar cache (no per-lane load ahead of the data loads).
__device__ inline RegLineView reg_line_view(const FftArgs& a, uint32_t line, uint32_t total, bool one_plane) {
  RegLineView r;
  r.live = line < total;
  const uint32_t l = r.live ? line : 0u;
  r.plane = line_plane(a, l);
  r.line = l - r.plane * a.src.lines;
  r.src = r.plane * a.src.plane_stride + r.line * a.src.line_stride;
  r.dst = (r.live && !plane_skipped(a, r.plane)) ? r.plane * a.dst.plane_stride + r.line * a.dst.line_stride : ~0ull;
  if (one_plane && a.zero_span) {
    const uint32_t p = __builtin_amdgcn_readfirstlane(r.plane);
    r.zs = a.zero_span[a.zero_span_period ? p % a.zero_span_period : p];
  } else {
    r.zs = r.live ? plane_zero_span(a, r.plane) : make_uint2(0, 0);
  }
  r.lw = (a.line_weight && r.live)
             ? a.line_weight[(a.line_weight_by_plane ? r.plane : r.line) / a.line_weight_div]
             : 1.0f;
  return r;
}

// Element t + i T of the line into x.v[i] (rotated by in_offset; the zero span reads 0).
template <int N, int V>
__device__ inline void reg_load(const FftArgs& a, const RegLineView& w, RegLine<N, V>& x, int t) {
  constexpr int T = Shape<N, V>::T;
  constexpr int E = N / T;
  const uint32_t es = static_cast<uint32_t>(a.src.elem_stride);
  typedef float F2v __attribute__((ext_vector_type(2)));
  // Every input element of a pass is read once: non-temporal, out of the caches' way.
  if (a.in_offset == 0) {  // no rotation: strided loads, the zero span (if any) skipped
    const float2* p = a.in + w.src + static_cast<uint64_t>(static_cast<uint32_t>(t) * es);
    const uint64_t step = static_cast<uint64_t>(T) * es;
    if (!a.zero_span) {
#pragma unroll
      for (int i = 0; i < E; ++i) {
        const F2v v = __builtin_nontemporal_load(reinterpret_cast<const F2v*>(p + i * step));
        x.v[i] = make_float2(v.x, v.y);
      }
      return;
    }
    // Branch-free, so every load is in flight at once: an element of the zero span reads
    // the thread's first element instead (a cache hit; the cell itself may never have been
    // written) and is replaced by zero.
#pragma unroll
    for (int i = 0; i < E; ++i) {
      const uint32_t keep = in_span(static_cast<uint32_t>(t + i * T), w.zs) ? 0u : ~0u;
      const uint64_t off = (static_cast<uint64_t>(i) * step) & static_cast<uint64_t>(static_cast<int64_t>(static_cast<int32_t>(keep)));
      const F2v v = __builtin_nontemporal_load(reinterpret_cast<const F2v*>(p + off));
      x.v[i] = keep_bits(make_float2(v.x, v.y), keep);
    }
    return;
  }
#pragma unroll
  for (int i = 0; i < E; ++i) {
    uint32_t src = static_cast<uint32_t>(t + i * T) + a.in_offset;
    if (src >= static_cast<uint32_t>(N)) src -= N;
    const uint32_t keep = in_span(src, w.zs) ? 0u : ~0u;  // the zero span reads element 0, then zero
    const F2v v = __builtin_nontemporal_load(reinterpret_cast<const F2v*>(a.in + w.src + static_cast<uint64_t>((src & keep) * es)));
    x.v[i] = keep_bits(make_float2(v.x, v.y), keep);
  }
}

// The register form of line group g, lines [g lpb, (g + 1) lpb) of the launch: thread
// (ly, t) = (tid % lpb, tid / lpb) holds elements t + i T of line g lpb + ly.  Each
// thread finds its own line's addresses (no per-block line tables, no barrier before
// the loads).
template <int N, int V, bool INV, bool EPI>
__device__ inline void fft_reg_group(const FftArgs& a, const ColumnOutputs& co, const float2* __restrict__ tw_global,
                                     int lpb_log2, int pad, uint32_t g) {
  static_assert(shape_ok<N, V>(), "radix schedule must multiply to N and split evenly over the line's threads");
  constexpr int T = Shape<N, V>::T;
  constexpr int E = N / T;
  extern __shared__ float2 smem[];
  __shared__ float ow[N];
  const int lpb = 1 << lpb_log2;
  const int S = lpb + pad;
  float2* tw = smem;
  float2* buf = smem + N;
  const uint32_t total = a.src.planes * a.src.lines;  // below 2^32 (checked at launch)
  const uint32_t nout = a.n_out;
  const int ly = threadIdx.x & (lpb - 1), t = threadIdx.x >> lpb_log2;
  // A block lies in one plane when lines per plane is a multiple of the block's lines
  // and the block holds no line past the end.
  const bool one_plane = a.src.lines % lpb == 0 && (g + 1) * lpb <= total;
  RegLine<N, V> x;
  // The line loads go first: the twiddle and weight staging below waits on its own loads
  // and so overlaps them.
  reg_load<N, V>(a, reg_line_view(a, g * lpb + ly, total, one_plane), x, t);
  for (int i = threadIdx.x; i < N; i += blockDim.x) {
    tw[i] = tw_global[i];
    ow[i] = (i < static_cast<int>(a.n_out) && a.out_weight) ? a.scale * a.out_weight[i] : a.scale;
  }
  __syncthreads();  // twiddles and output weights staged
  reg_transform<N, V, INV>(x, buf, S, ly, tw, t);
  // The view again rather than kept across the transform: fewer live registers.
  const RegLineView cur = reg_line_view(a, g * lpb + ly, total, one_plane);
  {
    if constexpr (!EPI) {
      if (cur.dst != ~0ull) {
        const uint32_t es = static_cast<uint32_t>(a.dst.elem_stride);
#pragma unroll
        for (int i = 0; i < E; ++i) {
          const uint32_t bin = static_cast<uint32_t>(t + i * T);
          const uint32_t e = bin >= a.out_offset ? bin - a.out_offset : bin + N - a.out_offset;
          if (e < nout) a.out[cur.dst + static_cast<uint64_t>(e * es)] = cscale(x.v[i], ow[e] * cur.lw);
        }
      }
    } else {
      // Column epilogue: line = (column c = plane, stacked coil k); group q = k / channels
      // writes coil-last images [N_r][cols][channels] and/or RSS [N_r][cols], the RSS
      // summed in coil order from |x|^2 staged in LDS (every group lies in one block).
      // Channels per group divide the (power of two) lines per block: a power of two.
      const uint32_t c = cur.plane, k = cur.line, ch = co.channels;
      const uint32_t ch_log2 = __builtin_ctz(ch);
      const uint32_t q = k >> ch_log2, kk = k & (ch - 1);
      const uint32_t istride = co.img_stride ? co.img_stride : ch;
      float2* img = cur.live ? co.coil_last[q] : nullptr;
      float* pw = reinterpret_cast<float*>(buf);
      __syncthreads();  // every thread's last-pass reads are done before buf is reused
#pragma unroll
      for (int i = 0; i < E; ++i) {
        const uint32_t bin = static_cast<uint32_t>(t + i * T);
        const uint32_t e = bin >= a.out_offset ? bin - a.out_offset : bin + N - a.out_offset;
        if (e >= nout) continue;
        const float2 v = cscale(x.v[i], ow[e] * cur.lw);
        if (img) img[(static_cast<uint64_t>(e) * co.cols + c) * istride + kk] = v;
        pw[e * (lpb + 1) + ly] = norm2(v);
      }
      __syncthreads();
      const uint32_t gpb_log2 = lpb_log2 - ch_log2;  // groups per block (log2)
      const uint32_t l0g = g * lpb;
      for (uint32_t r = threadIdx.x; r < (nout << gpb_log2); r += blockDim.x) {
        const uint32_t e = r >> gpb_log2, gb = r & ((1u << gpb_log2) - 1);
        const uint32_t l0 = l0g + (gb << ch_log2);
        if (l0 >= total) continue;
        const uint32_t col = line_plane(a, l0);
        float* out = co.rss[(l0 - col * a.src.lines) >> ch_log2];
        const uint64_t o = static_cast<uint64_t>(e) * co.cols + col;
        if (co.sumsq) {  // carried across launches (one group per position)
          float acc = co.first ? 0.f : co.sumsq[o];
          for (uint32_t m = 0; m < ch; ++m) acc += pw[e * (lpb + 1) + gb * ch + m];
          if (!co.last)
            co.sumsq[o] = acc;
          else if (out)
            out[o] = sqrtf(acc);
          continue;
        }
        if (!out) continue;
        float acc = 0.f;
        for (uint32_t m = 0; m < ch; ++m) acc += pw[e * (lpb + 1) + gb * ch + m];
        out[o] = sqrtf(acc);
      }
    }
  }
}

// Group of block b when the groups split evenly over the 8 XCDs (block b runs on XCD
// b % 8): consecutive groups on one XCD, so lines that share cache lines meet in one L2.
__device__ inline uint32_t xcd_group(uint32_t groups) {
  const uint32_t b = blockIdx.x;
  if (groups % 8 != 0) return b;
  return (b % 8) * (groups / 8) + b / 8;
}

template <int N, int V, bool INV, bool EPI>
__global__ __launch_bounds__(1024) void fft_reg_kernel(FftArgs a, ColumnOutputs co, const float2* __restrict__ tw_global,
                                                       int lpb_log2, int pad) {
  fft_reg_group<N, V, INV, EPI>(a, co, tw_global, lpb_log2, pad, xcd_group(gridDim.x));
}

// Fused items of one length in the register form: each block runs its item's lines.
template <int N, int V, bool INV>
__global__ __launch_bounds__(1024) void fft_fused_reg_kernel(const FusedTable<FftArgs>* t,
                                                             const float2* __restrict__ tw_global, int lpb_log2,
                                                             int pad) {
  const uint32_t item = fused_item(t, blockIdx.x);
  const FftArgs a = fused_args(t, item);
  const uint32_t g = static_cast<uint32_t>(blockIdx.x - fused_first(t, item));
  fft_reg_group<N, V, INV, false>(a, ColumnOutputs{}, tw_global, lpb_log2, pad, g);
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
      const uint2 zs = plane_zero_span(a, line / a.src.lines);
      if (!in_span(src, zs)) v = a.in[line_base(a.src, line) + src * a.src.elem_stride];
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
    if (line >= total || plane_skipped(a, line / a.src.lines)) continue;
    uint32_t bin = static_cast<uint32_t>(e) + a.out_offset;
    if (bin >= static_cast<uint32_t>(n)) bin -= n;
    float w = a.scale;
    if (a.out_weight) w *= a.out_weight[e];
    if (a.line_weight)
      w *= a.line_weight[(a.line_weight_by_plane ? line / a.src.lines : line % a.src.lines) / a.line_weight_div];
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
size_t reg_shmem(int lpb_log2) {
  return (static_cast<size_t>(N) + reg_lds_elems<N>(1 << lpb_log2, static_cast<int>(tune_get(kTuneFftLdsPad)))) *
         sizeof(float2);
}

template <typename K>
void raise_lds(K kernel, size_t shmem) {
  if (shmem > 65536)
    KSJ_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(kernel), hipFuncAttributeMaxDynamicSharedMemorySize,
                                static_cast<int>(shmem)));
}

// The register-staged form: both mappings across lines, a schedule of two or more passes.
template <int N, int V>
constexpr bool reg_form_compiled() {
  return Shape<N, V>::P >= 2;
}

template <int N, int V>
bool reg_form(const FftArgs& a) {
  return reg_form_compiled<N, V>() && a.load_lines_fast && a.store_lines_fast && tune_get(kTuneFftRegister) == 1;
}

// Lines per block (log2) of a register-form launch: the staged form's choice, reduced
// until the element-major LDS (N (lpb + pad) next to the twiddles and tables) fits.
template <int N, int V>
int reg_lines_log2(const FftArgs& a, int lpb) {
  int lpb_log2 = lines_per_block_log2<N, V>(a, lpb);
  constexpr int kStatic = 64 * 2 * 8 + N * 4 + 64 * 4;
  while (lpb_log2 > 0 && static_cast<int>(reg_shmem<N, V>(lpb_log2)) + kStatic > max_dynamic_lds()) --lpb_log2;
  return lpb_log2;
}

// a with lines_magic set for line indices below `span` (the launch's groups x lines per
// block): ceil(2^32 / lines) gives l / lines exactly by a multiply-high whenever
// l * lines < 2^32 (the rounding error of the magic, below l / 2^32, stays under 1 / lines).
inline FftArgs with_lines_magic(const FftArgs& a, uint64_t span) {
  FftArgs b = a;
  const uint64_t d = a.src.lines;
  b.lines_magic = (d >= 2 && span * d < (1ull << 32)) ? static_cast<uint32_t>(((1ull << 32) + d - 1) / d) : 0u;
  return b;
}

template <int N, int V, bool EPI>
void launch_reg(const FftArgs& a, const ColumnOutputs& co, const float2* tw, hipStream_t stream, int lpb) {
  constexpr int T = Shape<N, V>::T;
  const int lpb_log2 = reg_lines_log2<N, V>(a, lpb);
  const int lines = 1 << lpb_log2;
  if constexpr (EPI) require(lines % co.channels == 0, "column epilogue needs whole coil groups per block");
  const uint64_t total = static_cast<uint64_t>(a.src.planes) * a.src.lines;
  require(total < (1ull << 32), "register-form FFT launches take fewer than 2^32 lines");
  const int pad = static_cast<int>(tune_get(kTuneFftLdsPad));
  const size_t shmem = reg_shmem<N, V>(lpb_log2);
  const uint32_t groups = static_cast<uint32_t>(ceil_div(total, static_cast<uint64_t>(lines)));
  const FftArgs b = with_lines_magic(a, static_cast<uint64_t>(groups) * lines);
  auto go = [&](auto kernel) {
    raise_lds(kernel, shmem);
    hipLaunchKernelGGL(kernel, dim3(groups), dim3(lines * T), shmem, stream, b, co, tw, lpb_log2, pad);
  };
  if (a.inverse)
    go(&fft_reg_kernel<N, V, true, EPI>);
  else
    go(&fft_reg_kernel<N, V, false, EPI>);
  KSJ_LAUNCH_CHECK();
}

template <int N, int V>
void launch_fast(const FftArgs& a, const float2* tw, hipStream_t stream, int lpb) {
  constexpr int T = Shape<N, V>::T;
  if constexpr (reg_form_compiled<N, V>()) {
    if (reg_form<N, V>(a)) return launch_reg<N, V, false>(a, ColumnOutputs{}, tw, stream, lpb);
  }
  const int lpb_log2 = lines_per_block_log2<N, V>(a, lpb);
  lpb = 1 << lpb_log2;
  const uint64_t total = static_cast<uint64_t>(a.src.planes) * a.src.lines;
  const size_t shmem = fft_shmem<N, V>(lpb_log2);
  const unsigned grid = ceil_div(total, static_cast<uint64_t>(lpb));
  if (a.inverse) {
    raise_lds(&fft_lds_kernel<N, V, true>, shmem);
    hipLaunchKernelGGL((fft_lds_kernel<N, V, true>), dim3(grid), dim3(lpb * T), shmem, stream, a, tw, lpb_log2);
  } else {
    raise_lds(&fft_lds_kernel<N, V, false>, shmem);
    hipLaunchKernelGGL((fft_lds_kernel<N, V, false>), dim3(grid), dim3(lpb * T), shmem, stream, a, tw, lpb_log2);
  }
  KSJ_LAUNCH_CHECK();
}

template <int N, int V>
void launch_fused(const FftArgs* const* items, uint32_t count, const float2* tw, void* d_table, hipStream_t stream) {
  constexpr int T = Shape<N, V>::T;
  bool reg = false;
  if constexpr (reg_form_compiled<N, V>()) reg = reg_form<N, V>(*items[0]);
  const int lpb_log2 = reg ? reg_lines_log2<N, V>(*items[0], 0) : lines_per_block_log2<N, V>(*items[0], 0);
  const uint64_t lpb = 1ull << lpb_log2;
  FusedTableBuilder<FftArgs> b;
  for (uint32_t i = 0; i < count; ++i) {
    require(items[i]->inverse == items[0]->inverse && items[i]->load_lines_fast == items[0]->load_lines_fast &&
              items[i]->store_lines_fast == items[0]->store_lines_fast,
            "fused FFT items must share direction and thread mapping");
    const uint64_t total = static_cast<uint64_t>(items[i]->src.planes) * items[i]->src.lines;
    require(!reg || total < (1ull << 32), "register-form FFT launches take fewer than 2^32 lines");
    const uint64_t groups = ceil_div(total, lpb);
    b.add(reg ? with_lines_magic(*items[i], groups * lpb) : *items[i], groups);
  }
  b.upload(d_table, stream);
  const auto* t = static_cast<const FusedTable<FftArgs>*>(d_table);
  const dim3 grid(static_cast<uint32_t>(b.blocks())), block(static_cast<uint32_t>(lpb * T));
  if constexpr (reg_form_compiled<N, V>()) {
    if (reg) {
      const size_t rshmem = reg_shmem<N, V>(lpb_log2);
      const int pad = static_cast<int>(tune_get(kTuneFftLdsPad));
      if (items[0]->inverse) {
        raise_lds(&fft_fused_reg_kernel<N, V, true>, rshmem);
        hipLaunchKernelGGL((fft_fused_reg_kernel<N, V, true>), grid, block, rshmem, stream, t, tw, lpb_log2, pad);
      } else {
        raise_lds(&fft_fused_reg_kernel<N, V, false>, rshmem);
        hipLaunchKernelGGL((fft_fused_reg_kernel<N, V, false>), grid, block, rshmem, stream, t, tw, lpb_log2, pad);
      }
      KSJ_LAUNCH_CHECK();
      return;
    }
  }
  const size_t shmem = fft_shmem<N, V>(lpb_log2);
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
// threads per line (1) at 512 and schedule 0 at every other length.  (With per-block line
// tables, 3 served the 256x256 radial stream 1.4 % faster; with each thread finding its own
// line, schedule 0 serves it 4 % faster than 3.)
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

namespace {

template <int N>
int columns_lines_per_block(const FftArgs& a) {
  const int v = fft_schedule<N>();
  auto lines = [&](auto vv) {
    constexpr int V = decltype(vv)::value;
    if constexpr (!reg_form_compiled<N, V>()) {
      return 0;
    } else {
      return 1 << reg_lines_log2<N, V>(a, 0);
    }
  };
  switch (v) {
    case 1: return lines(std::integral_constant<int, 1>{});
    case 2: return lines(std::integral_constant<int, 2>{});
    default: return lines(std::integral_constant<int, 0>{});
  }
}

template <int N>
void launch_columns(const FftArgs& a, const ColumnOutputs& co, const float2* tw, hipStream_t stream) {
  switch (fft_schedule<N>()) {
    case 1: return launch_reg<N, 1, true>(a, co, tw, stream, 0);
    case 2: return launch_reg<N, 2, true>(a, co, tw, stream, 0);
    default: return launch_reg<N, 0, true>(a, co, tw, stream, 0);
  }
}

}  // namespace

bool fft_columns_supported(const FftArgs& a, uint32_t channels) {
  if (!a.load_lines_fast || !a.store_lines_fast || channels == 0 || tune_get(kTuneFftRegister) != 1) return false;
  if (a.src.lines % channels != 0) return false;
  int lines = 0;
  switch (a.n) {
    case 256: lines = columns_lines_per_block<256>(a); break;
    case 320: lines = columns_lines_per_block<320>(a); break;
    case 384: lines = columns_lines_per_block<384>(a); break;
    case 512: lines = columns_lines_per_block<512>(a); break;
    case 640: lines = columns_lines_per_block<640>(a); break;
    case 768: lines = columns_lines_per_block<768>(a); break;
    case 1024: lines = columns_lines_per_block<1024>(a); break;
    default: return false;
  }
  return lines > 0 && lines % channels == 0;
}

void fft_launch_columns(const FftArgs& a, const ColumnOutputs& out, hipStream_t stream) {
  require(fft_columns_supported(a, out.channels), "column epilogue not available for this pass");
  require(a.n_out >= 1 && a.n_out <= a.n && a.in_offset < a.n && a.out_offset < a.n, "invalid column pass");
  const float2* tw = fft_twiddles(a.n, stream);
  switch (a.n) {
    case 256: return launch_columns<256>(a, out, tw, stream);
    case 320: return launch_columns<320>(a, out, tw, stream);
    case 384: return launch_columns<384>(a, out, tw, stream);
    case 512: return launch_columns<512>(a, out, tw, stream);
    case 640: return launch_columns<640>(a, out, tw, stream);
    case 768: return launch_columns<768>(a, out, tw, stream);
    default: return launch_columns<1024>(a, out, tw, stream);
  }
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
