// Calibration estimators (providers/kspacejet-calibration/src/operators/*.cpp) on the GPU.
//
// covariance: workgroups stride over chunks of 64 samples staged in LDS (means removed
//   for the noise kind); each thread owns a 4 x 4 tile of the C x C matrix (and, for
//   fewer than 64 channels, every phases-th sample of the chunk), forms the chunk's
//   products in FP32 and adds them to FP64 accumulators.  A second kernel sums the
//   per-thread partials in a fixed order.  C <= 64, so the Gram work is small next to
//   reading the samples once.
// eigen: one workgroup, FP64, matrix and eigenvectors in LDS.  Cyclic Jacobi in the
//   round-robin order, n/2 disjoint rotations per step, each rotation built from the
//   pivot as in the reference's sequential solver
//   (libs/numerics/kspacejet-linalg/include/kspacejet/linalg/decompositions.hpp:503-626:
//   tau, tangent, cosine, sine, pivot phase) and applied two-sided; eigenvalues sorted
//   ascending by the same selection sort.
#include "calib.hpp"

#include "common.hpp"

#include <cmath>

namespace ksj::gpu {
namespace {

constexpr int kChunk = 64;         // samples per LDS chunk
constexpr int kThreads = 256;
constexpr int kMaxCovBlocks = 128; // workgroups of the covariance pass
constexpr int kMeanBlocks = 64;    // sample ranges per channel of the mean pass

__host__ __device__ inline uint32_t pad4(uint32_t c) { return (c + 3) / 4 * 4; }

struct CovLayout {
  uint32_t cp, ntiles, phases, blocks;
};
CovLayout cov_layout(uint32_t channels, uint64_t samples) {
  CovLayout l{};
  l.cp = pad4(channels);
  l.ntiles = (l.cp / 4) * (l.cp / 4);
  l.phases = kThreads / l.ntiles > 0 ? kThreads / l.ntiles : 1;
  const uint64_t chunks = (samples + kChunk - 1) / kChunk;
  l.blocks = static_cast<uint32_t>(chunks < kMaxCovBlocks ? (chunks > 0 ? chunks : 1) : kMaxCovBlocks);
  return l;
}

struct Workspace {
  float2* means;       // [64]
  double2* mean_part;  // [64][kMeanBlocks]
  double2* cov_part;   // [blocks][256][16]
  double2* cov;        // [64][64]
  double* values;      // [64]
  double2* vectors;    // [64][64]
  uint32_t* status;    // [1]
};
size_t align256(size_t b) { return (b + 255) / 256 * 256; }
Workspace carve(void* ws, const CovLayout& l) {
  char* p = static_cast<char*>(ws);
  Workspace w{};
  w.means = reinterpret_cast<float2*>(p);
  p += align256(64 * sizeof(float2));
  w.mean_part = reinterpret_cast<double2*>(p);
  p += align256(64 * kMeanBlocks * sizeof(double2));
  w.cov_part = reinterpret_cast<double2*>(p);
  p += align256(static_cast<size_t>(l.blocks) * kThreads * 16 * sizeof(double2));
  w.cov = reinterpret_cast<double2*>(p);
  p += align256(64 * 64 * sizeof(double2));
  w.values = reinterpret_cast<double*>(p);
  p += align256(64 * sizeof(double));
  w.vectors = reinterpret_cast<double2*>(p);
  p += align256(64 * 64 * sizeof(double2));
  w.status = reinterpret_cast<uint32_t*>(p);
  return w;
}

// ---------------------------------------------------------------- means
__global__ __launch_bounds__(kThreads) void mean_partial_kernel(const float2* x, uint64_t samples, double2* part) {
  const uint32_t c = blockIdx.y, b = blockIdx.x;
  const uint64_t per = (samples + kMeanBlocks - 1) / kMeanBlocks;
  const uint64_t s0 = b * per, s1 = s0 + per < samples ? s0 + per : samples;
  double sr = 0.0, si = 0.0;
  for (uint64_t s = s0 + threadIdx.x; s < s1; s += kThreads) {
    const float2 v = x[c * samples + s];
    sr += v.x;
    si += v.y;
  }
  __shared__ double red[2][kThreads];
  red[0][threadIdx.x] = sr;
  red[1][threadIdx.x] = si;
  __syncthreads();
  for (int w = kThreads / 2; w > 0; w /= 2) {
    if (static_cast<int>(threadIdx.x) < w) {
      red[0][threadIdx.x] += red[0][threadIdx.x + w];
      red[1][threadIdx.x] += red[1][threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) part[c * kMeanBlocks + b] = make_double2(red[0][0], red[1][0]);
}

__global__ void mean_final_kernel(const double2* part, uint32_t channels, uint64_t samples, float2* means) {
  const uint32_t c = threadIdx.x;
  if (c >= 64) return;
  if (c >= channels) {
    means[c] = make_float2(0.f, 0.f);
    return;
  }
  double sr = 0.0, si = 0.0;
  for (int b = 0; b < kMeanBlocks; ++b) {
    sr += part[c * kMeanBlocks + b].x;
    si += part[c * kMeanBlocks + b].y;
  }
  means[c] = make_float2(static_cast<float>(sr / static_cast<double>(samples)),
                         static_cast<float>(si / static_cast<double>(samples)));
}

// ---------------------------------------------------------------- covariance
template <int KIND>
__global__ __launch_bounds__(kThreads) void cov_partial_kernel(const float2* x, const float2* means, uint32_t channels,
                                                               uint64_t samples, CovLayout l, double2* part) {
  __shared__ float2 tile[64][kChunk + 1];
  const uint32_t t = threadIdx.x;
  const uint32_t tt = t % l.ntiles, phase = t / l.ntiles;
  const bool active = phase < l.phases;
  const uint32_t tr = 4 * (tt / (l.cp / 4)), tc = 4 * (tt % (l.cp / 4));
  double2 acc[4][4];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) acc[a][b] = make_double2(0.0, 0.0);
  const uint64_t chunks = (samples + kChunk - 1) / kChunk;
  for (uint64_t ch = blockIdx.x; ch < chunks; ch += gridDim.x) {
    const uint64_t s0 = ch * kChunk;
    for (uint32_t i = t; i < l.cp * kChunk; i += kThreads) {
      const uint32_t c = i / kChunk, s = i % kChunk;
      float2 v = make_float2(0.f, 0.f);
      if (c < channels && s0 + s < samples) {
        v = x[c * samples + s0 + s];
        if (KIND == kCovarianceNoise) v = make_float2(v.x - means[c].x, v.y - means[c].y);
      }
      tile[c][s] = v;
    }
    __syncthreads();
    if (active) {
      float2 f[4][4];
#pragma unroll
      for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) f[a][b] = make_float2(0.f, 0.f);
      for (uint32_t s = phase; s < kChunk; s += l.phases) {
        float2 u[4], v[4];
#pragma unroll
        for (int a = 0; a < 4; ++a) {
          u[a] = tile[tr + a][s];
          v[a] = tile[tc + a][s];
        }
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
          for (int b = 0; b < 4; ++b) {
            // noise: conj(u) v; second moment: u conj(v)
            const float sgn = KIND == kCovarianceNoise ? 1.f : -1.f;
            f[a][b].x += u[a].x * v[b].x + u[a].y * v[b].y;
            f[a][b].y += sgn * (u[a].x * v[b].y - u[a].y * v[b].x);
          }
      }
#pragma unroll
      for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) {
          acc[a][b].x += f[a][b].x;
          acc[a][b].y += f[a][b].y;
        }
    }
    __syncthreads();
  }
  double2* out = part + (static_cast<size_t>(blockIdx.x) * kThreads + t) * 16;
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) out[4 * a + b] = active ? acc[a][b] : make_double2(0.0, 0.0);
}

__global__ void cov_final_kernel(const double2* part, uint32_t channels, uint64_t samples, CovLayout l, double2* cov) {
  const uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= channels * channels) return;
  const uint32_t r = e / channels, c = e % channels;
  const uint32_t tt = (r / 4) * (l.cp / 4) + c / 4, slot = 4 * (r % 4) + c % 4;
  double sr = 0.0, si = 0.0;
  for (uint32_t b = 0; b < l.blocks; ++b)
    for (uint32_t ph = 0; ph < l.phases; ++ph) {
      const double2 v = part[(static_cast<size_t>(b) * kThreads + tt + ph * l.ntiles) * 16 + slot];
      sr += v.x;
      si += v.y;
    }
  const double inv = 1.0 / static_cast<double>(samples);
  cov[r * channels + c] = make_double2(sr * inv, si * inv);
}

// ---------------------------------------------------------------- eigen
__device__ inline double2 dmul(double2 a, double2 b) { return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }
__device__ inline double2 dscale(double2 a, double s) { return make_double2(a.x * s, a.y * s); }
__device__ inline double2 dsub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__device__ inline double2 dadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ inline double2 dconj(double2 a) { return make_double2(a.x, -a.y); }

constexpr int kEigMax = 64;
constexpr int kMaxSweeps = 60;

// Round-robin pairing: in step k of m - 1, position 0 is fixed and the others rotate;
// pair j meets positions j and m - 1 - j.
__device__ inline uint32_t rr_player(uint32_t pos, uint32_t k, uint32_t m) {
  return pos == 0 ? 0u : 1u + (pos - 1u + k) % (m - 1u);
}

__global__ __launch_bounds__(kThreads) void eigen_kernel(const double2* in, uint32_t n, double* values,
                                                         double2* vectors, uint32_t* status) {
  __shared__ double2 A[kEigMax][kEigMax + 1];
  __shared__ double2 V[kEigMax][kEigMax + 1];
  __shared__ double rot_c[kEigMax / 2], rot_s[kEigMax / 2];
  __shared__ double2 rot_phase[kEigMax / 2];
  __shared__ uint32_t rot_p[kEigMax / 2], rot_q[kEigMax / 2];
  __shared__ double red[kThreads];
  __shared__ int bad;
  __shared__ uint32_t perm[kEigMax];
  const uint32_t t = threadIdx.x;
  if (t == 0) bad = 0;
  __syncthreads();
  for (uint32_t i = t; i < n * n; i += kThreads) {
    const uint32_t r = i / n, c = i % n;
    const double2 v = in[i];
    if (!isfinite(v.x) || !isfinite(v.y)) bad = 1;
    A[r][c] = v;
    V[r][c] = make_double2(r == c ? 1.0 : 0.0, 0.0);
  }
  __syncthreads();
  if (bad) {
    if (t == 0) *status = 1;
    return;
  }
  // Hermitian part as the solver sees it: real diagonal (the reference reads real parts).
  for (uint32_t i = t; i < n; i += kThreads) A[i][i].y = 0.0;
  __syncthreads();
  // Frobenius norm^2 and the convergence threshold on the off-diagonal norm^2.
  double fro = 0.0;
  for (uint32_t i = t; i < n * n; i += kThreads) {
    const double2 v = A[i / n][i % n];
    fro += v.x * v.x + v.y * v.y;
  }
  red[t] = fro;
  __syncthreads();
  for (int w = kThreads / 2; w > 0; w /= 2) {
    if (static_cast<int>(t) < w) red[t] += red[t + w];
    __syncthreads();
  }
  const double tol = red[0] * 1e-30;  // (1e-15 ||A||_F)^2
  __syncthreads();
  const uint32_t m = n + (n & 1u);
  const uint32_t pairs = m / 2;
  bool converged = n == 1;
  for (int sweep = 0; sweep < kMaxSweeps && !converged; ++sweep) {
    for (uint32_t k = 0; k + 1 < m; ++k) {
      if (t < pairs) {
        uint32_t p = rr_player(t, k, m), q = rr_player(m - 1 - t, k, m);
        if (p > q) {
          const uint32_t tmp = p;
          p = q;
          q = tmp;
        }
        double c = 1.0, s = 0.0;
        double2 ph = make_double2(1.0, 0.0);
        if (q < n) {
          const double2 apq = A[p][q];
          const double mag = sqrt(apq.x * apq.x + apq.y * apq.y);
          if (mag > 0.0) {
            const double app = A[p][p].x, aqq = A[q][q].x;
            const double tau = (aqq - app) / (2.0 * mag);
            const double tg = tau >= 0.0 ? 1.0 / (tau + sqrt(1.0 + tau * tau)) : -1.0 / (-tau + sqrt(1.0 + tau * tau));
            c = 1.0 / sqrt(1.0 + tg * tg);
            s = tg * c;
            ph = make_double2(apq.x / mag, apq.y / mag);
          }
        }
        rot_p[t] = p;
        rot_q[t] = q;
        rot_c[t] = c;
        rot_s[t] = s;
        rot_phase[t] = ph;
      }
      __syncthreads();
      // Columns: A <- A J and V <- V J, J = [[c, s], [-s conj(ph), c conj(ph)]] on (p, q).
      for (uint32_t i = t; i < n * pairs; i += kThreads) {
        const uint32_t j = i / n, r = i % n;
        const uint32_t p = rot_p[j], q = rot_q[j];
        if (q >= n || rot_s[j] == 0.0) continue;
        const double c = rot_c[j], s = rot_s[j];
        const double2 phc = dconj(rot_phase[j]);
        const double2 ap = A[r][p], aq = dmul(A[r][q], phc);
        A[r][p] = dsub(dscale(ap, c), dscale(aq, s));
        A[r][q] = dadd(dscale(ap, s), dscale(aq, c));
        const double2 vp = V[r][p], vq = dmul(V[r][q], phc);
        V[r][p] = dsub(dscale(vp, c), dscale(vq, s));
        V[r][q] = dadd(dscale(vp, s), dscale(vq, c));
      }
      __syncthreads();
      // Rows: A <- J^H A.
      for (uint32_t i = t; i < n * pairs; i += kThreads) {
        const uint32_t j = i / n, col = i % n;
        const uint32_t p = rot_p[j], q = rot_q[j];
        if (q >= n || rot_s[j] == 0.0) continue;
        const double c = rot_c[j], s = rot_s[j];
        const double2 ph = rot_phase[j];
        const double2 ap = A[p][col], aq = dmul(A[q][col], ph);
        A[p][col] = dsub(dscale(ap, c), dscale(aq, s));
        A[q][col] = dadd(dscale(ap, s), dscale(aq, c));
      }
      __syncthreads();
      // The pivots are annihilated, the diagonal kept real.
      if (t < pairs && rot_q[t] < n && rot_s[t] != 0.0) {
        const uint32_t p = rot_p[t], q = rot_q[t];
        A[p][q] = make_double2(0.0, 0.0);
        A[q][p] = make_double2(0.0, 0.0);
        A[p][p].y = 0.0;
        A[q][q].y = 0.0;
      }
      __syncthreads();
    }
    double off = 0.0;
    for (uint32_t i = t; i < n * n; i += kThreads) {
      const uint32_t r = i / n, c = i % n;
      if (r != c) off += A[r][c].x * A[r][c].x + A[r][c].y * A[r][c].y;
    }
    red[t] = off;
    __syncthreads();
    for (int w = kThreads / 2; w > 0; w /= 2) {
      if (static_cast<int>(t) < w) red[t] += red[t + w];
      __syncthreads();
    }
    converged = red[0] <= tol;
    __syncthreads();
  }
  if (!converged) {
    if (t == 0) *status = 2;
    return;
  }
  // Ascending order by the reference's selection sort (strictly smaller wins, swaps).
  if (t == 0) {
    double ev[kEigMax];
    for (uint32_t i = 0; i < n; ++i) {
      ev[i] = A[i][i].x;
      perm[i] = i;
    }
    for (uint32_t first = 0; first < n; ++first) {
      uint32_t mn = first;
      for (uint32_t cand = first + 1; cand < n; ++cand)
        if (ev[cand] < ev[mn]) mn = cand;
      if (mn == first) continue;
      const double te = ev[first];
      ev[first] = ev[mn];
      ev[mn] = te;
      const uint32_t tp = perm[first];
      perm[first] = perm[mn];
      perm[mn] = tp;
    }
    for (uint32_t i = 0; i < n; ++i) values[i] = ev[i];
    *status = 0;
  }
  __syncthreads();
  for (uint32_t i = t; i < n * n; i += kThreads) {
    const uint32_t r = i / n, k = i % n;
    vectors[i] = V[r][perm[k]];
  }
}

// ---------------------------------------------------------------- outputs
__global__ void whitening_kernel(const double* values, const double2* vectors, uint32_t n, double floor_value,
                                 float2* w) {
  const uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n * n) return;
  const uint32_t r = e / n, c = e % n;
  double sr = 0.0, si = 0.0;
  for (uint32_t k = 0; k < n; ++k) {
    const double inv = 1.0 / sqrt(fmax(values[k], floor_value));
    const double2 a = vectors[r * n + k], b = vectors[c * n + k];
    // a * inv * conj(b)
    sr += (a.x * b.x + a.y * b.y) * inv;
    si += (a.y * b.x - a.x * b.y) * inv;
  }
  w[e] = make_float2(static_cast<float>(sr), static_cast<float>(si));
}

__global__ void basis_kernel(const double2* vectors, uint32_t n, uint32_t nv, float2* basis) {
  const uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nv * n) return;
  const uint32_t v = e / n, p = e % n;
  const double2 x = vectors[p * n + (n - 1 - v)];
  basis[e] = make_float2(static_cast<float>(x.x), static_cast<float>(-x.y));
}

__global__ void phase_estimate_kernel(const float2* ref, uint32_t channels, uint32_t lines, uint32_t readout,
                                      float2* out) {
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x, c = blockIdx.y;
  if (k >= readout || c >= channels) return;
  float sr = 0.f, si = 0.f;
  for (uint32_t l = 0; l < lines; ++l) {
    const float2 v = ref[(static_cast<uint64_t>(c) * lines + l) * readout + k];
    sr += v.x;
    si += v.y;
  }
  const double m2 = static_cast<double>(sr) * sr + static_cast<double>(si) * si;
  const float mag = static_cast<float>(sqrt(m2));
  out[static_cast<uint64_t>(c) * readout + k] = mag == 0.f ? make_float2(1.f, 0.f) : make_float2(sr / mag, -si / mag);
}

uint32_t read_status(const uint32_t* d_status, hipStream_t stream) {
  uint32_t s = 0;
  KSJ_HIP(hipMemcpyAsync(&s, d_status, sizeof(s), hipMemcpyDeviceToHost, stream));
  KSJ_HIP(hipStreamSynchronize(stream));
  return s;
}

// Walsh's adaptive sensitivity maps: CP lanes per pixel (lane c holds coil c; lanes past
// the coil count hold zeros), 64 / CP pixels per wave.  The local covariance is never
// formed: R v = sum over the window of I (I^H v), with I^H v summed across the pixel's
// lanes by an xor butterfly (every lane ends with the same bits: each step adds the
// same two values in either order).  Window offsets run in the same order for every
// pixel, with out-of-image neighbours read as zero, so the shuffles run on full waves.
template <int CP>
__device__ inline float group_sum(float s) {
#pragma unroll
  for (int m = 1; m < CP; m <<= 1) s += __shfl_xor(s, m, CP);
  return s;
}

template <int CP>
__device__ inline float2 group_normalise(float2 v) {
  const float n2 = group_sum<CP>(fmaf(v.x, v.x, v.y * v.y));
  const float inv = n2 > 0.f ? 1.0f / sqrtf(n2) : 0.f;
  return make_float2(v.x * inv, v.y * inv);
}

template <int CP>
__global__ __launch_bounds__(256) void walsh_kernel(const float2* __restrict__ img, float2* __restrict__ maps,
                                                    uint32_t rows, uint32_t cols, uint32_t coils, int r,
                                                    uint32_t iterations) {
  constexpr uint32_t kPerWave = 64 / CP;
  const uint32_t lane = threadIdx.x & 63, c = lane % CP, g = lane / CP;
  const uint64_t p = (static_cast<uint64_t>(blockIdx.x) * (blockDim.x / 64) + threadIdx.x / 64) * kPerWave + g;
  const uint64_t npix = static_cast<uint64_t>(rows) * cols;
  const bool live = p < npix;
  const int y = live ? static_cast<int>(p / cols) : 0, x = live ? static_cast<int>(p % cols) : 0;
  auto load = [&](int dy, int dx) {
    const int yy = y + dy, xx = x + dx;
    const bool ok = live && c < coils && yy >= 0 && yy < static_cast<int>(rows) && xx >= 0 &&
                    xx < static_cast<int>(cols);
    return ok ? img[(static_cast<uint64_t>(yy) * cols + xx) * coils + c] : make_float2(0.f, 0.f);
  };
  // v = R e_0 = sum I conj(I_0)
  float2 v = make_float2(0.f, 0.f);
  for (int dy = -r; dy <= r; ++dy)
    for (int dx = -r; dx <= r; ++dx) {
      const float2 a = load(dy, dx);
      const float2 a0 = make_float2(__shfl(a.x, 0, CP), __shfl(a.y, 0, CP));
      v.x = fmaf(a.x, a0.x, fmaf(a.y, a0.y, v.x));
      v.y = fmaf(a.y, a0.x, fmaf(-a.x, a0.y, v.y));
    }
  v = group_normalise<CP>(v);
  for (uint32_t it = 0; it < iterations; ++it) {
    float2 u = make_float2(0.f, 0.f);
    for (int dy = -r; dy <= r; ++dy)
      for (int dx = -r; dx <= r; ++dx) {
        const float2 a = load(dy, dx);
        // d = I^H v, then u += I d
        const float dr = group_sum<CP>(fmaf(a.x, v.x, a.y * v.y));
        const float di = group_sum<CP>(fmaf(a.x, v.y, -(a.y * v.x)));
        u.x = fmaf(a.x, dr, fmaf(-a.y, di, u.x));
        u.y = fmaf(a.x, di, fmaf(a.y, dr, u.y));
      }
    v = group_normalise<CP>(u);
  }
  // coil 0 real and >= 0
  const float2 v0 = make_float2(__shfl(v.x, 0, CP), __shfl(v.y, 0, CP));
  const float m = sqrtf(fmaf(v0.x, v0.x, v0.y * v0.y));
  if (m > 0.f) v = make_float2(fmaf(v.x, v0.x, v.y * v0.y) / m, fmaf(v.y, v0.x, -(v.x * v0.y)) / m);
  if (live && c < coils) maps[p * coils + c] = v;
}

void eigen_of_covariance(const float2* d_x, uint32_t channels, uint64_t samples, int kind, void* d_ws,
                         size_t ws_bytes, hipStream_t stream, Workspace& w) {
  channel_covariance(d_x, channels, samples, kind, nullptr, d_ws, ws_bytes, stream);
  w = carve(d_ws, cov_layout(channels, samples));
  hermitian_eigen(w.cov, channels, w.values, w.vectors, w.status, stream);
  const uint32_t s = read_status(w.status, stream);
  if (s == 1) throw Error(kErrArgument, "calibration samples hold a non-finite value");
  if (s != 0) throw Error(kErrArgument, "Hermitian eigen-decomposition did not converge");
}

}  // namespace

size_t calibration_workspace_bytes(uint32_t channels, uint64_t samples) {
  const CovLayout l = cov_layout(channels, samples);
  return align256(64 * sizeof(float2)) + align256(64 * kMeanBlocks * sizeof(double2)) +
         align256(static_cast<size_t>(l.blocks) * kThreads * 16 * sizeof(double2)) +
         align256(64 * 64 * sizeof(double2)) + align256(64 * sizeof(double)) + align256(64 * 64 * sizeof(double2)) +
         align256(sizeof(uint32_t));
}

void channel_covariance(const float2* d_x, uint32_t channels, uint64_t samples, int kind, double2* d_cov,
                        void* d_ws, size_t ws_bytes, hipStream_t stream) {
  require(channels >= 1 && channels <= kCalibMaxChannels, "calibration takes 1..64 channels");
  require(samples >= 1, "calibration needs at least one sample");
  require(kind == kCovarianceNoise || kind == kCovarianceSecondMoment, "unknown covariance kind");
  require(d_x != nullptr && d_ws != nullptr, "null argument");
  require(ws_bytes >= calibration_workspace_bytes(channels, samples), "calibration workspace too small");
  const CovLayout l = cov_layout(channels, samples);
  Workspace w = carve(d_ws, l);
  if (kind == kCovarianceNoise) {
    hipLaunchKernelGGL(mean_partial_kernel, dim3(kMeanBlocks, channels), dim3(kThreads), 0, stream, d_x, samples,
                       w.mean_part);
    KSJ_LAUNCH_CHECK();
    hipLaunchKernelGGL(mean_final_kernel, dim3(1), dim3(64), 0, stream, w.mean_part, channels, samples, w.means);
    KSJ_LAUNCH_CHECK();
    hipLaunchKernelGGL(cov_partial_kernel<kCovarianceNoise>, dim3(l.blocks), dim3(kThreads), 0, stream, d_x, w.means,
                       channels, samples, l, w.cov_part);
  } else {
    hipLaunchKernelGGL(cov_partial_kernel<kCovarianceSecondMoment>, dim3(l.blocks), dim3(kThreads), 0, stream, d_x,
                       w.means, channels, samples, l, w.cov_part);
  }
  KSJ_LAUNCH_CHECK();
  double2* out = d_cov ? d_cov : w.cov;
  hipLaunchKernelGGL(cov_final_kernel, dim3(ceil_div(static_cast<uint64_t>(channels) * channels, 256)), dim3(256), 0,
                     stream, w.cov_part, channels, samples, l, out);
  KSJ_LAUNCH_CHECK();
}

void hermitian_eigen(const double2* d_matrix, uint32_t n, double* d_values, double2* d_vectors, uint32_t* d_status,
                     hipStream_t stream) {
  require(n >= 1 && n <= static_cast<uint32_t>(kEigMax), "Hermitian eigen-decomposition takes 1..64 rows");
  require(d_matrix && d_values && d_vectors && d_status, "null argument");
  hipLaunchKernelGGL(eigen_kernel, dim3(1), dim3(kThreads), 0, stream, d_matrix, n, d_values, d_vectors, d_status);
  KSJ_LAUNCH_CHECK();
}

void noise_model_estimate(const float2* d_x, uint32_t channels, uint64_t samples, float eigenvalue_floor,
                          float2* d_whitening, void* d_ws, size_t ws_bytes, hipStream_t stream) {
  require(samples >= 2, "a noise model needs at least two samples per channel");
  require(eigenvalue_floor > 0.f, "eigenvalue floor must be positive");
  require(d_whitening != nullptr, "null argument");
  Workspace w{};
  eigen_of_covariance(d_x, channels, samples, kCovarianceNoise, d_ws, ws_bytes, stream, w);
  hipLaunchKernelGGL(whitening_kernel, dim3(ceil_div(static_cast<uint64_t>(channels) * channels, 256)), dim3(256), 0,
                     stream, w.values, w.vectors, channels, static_cast<double>(eigenvalue_floor), d_whitening);
  KSJ_LAUNCH_CHECK();
}

void coil_compression_basis_estimate(const float2* d_x, uint32_t physical, uint32_t virtual_channels,
                                     uint64_t samples, float2* d_basis, void* d_ws, size_t ws_bytes,
                                     hipStream_t stream) {
  require(virtual_channels >= 1 && virtual_channels <= physical, "virtual channels must be in [1, physical]");
  require(d_basis != nullptr, "null argument");
  Workspace w{};
  eigen_of_covariance(d_x, physical, samples, kCovarianceSecondMoment, d_ws, ws_bytes, stream, w);
  hipLaunchKernelGGL(basis_kernel, dim3(ceil_div(static_cast<uint64_t>(virtual_channels) * physical, 256)), dim3(256),
                     0, stream, w.vectors, physical, virtual_channels, d_basis);
  KSJ_LAUNCH_CHECK();
}

void phase_correction_estimate(const float2* d_ref, uint32_t channels, uint32_t lines, uint32_t readout,
                               float2* d_phase, hipStream_t stream) {
  require(channels >= 1 && channels <= kCalibMaxChannels, "phase estimation takes 1..64 channels");
  require(lines >= 1 && readout >= 1, "phase estimation needs reference lines");
  require(d_ref != nullptr && d_phase != nullptr, "null argument");
  hipLaunchKernelGGL(phase_estimate_kernel, dim3(ceil_div(readout, 256), channels), dim3(256), 0, stream, d_ref,
                     channels, lines, readout, d_phase);
  KSJ_LAUNCH_CHECK();
}

void sensitivity_maps_walsh(const float2* d_images, uint32_t rows, uint32_t cols, uint32_t coils, uint32_t window,
                            uint32_t iterations, float2* d_maps, hipStream_t stream) {
  require(d_images != nullptr && d_maps != nullptr, "null argument");
  require(rows >= 1 && cols >= 1 && static_cast<uint64_t>(rows) * cols <= (uint64_t{1} << 26),
          "sensitivity maps take 1 .. 2^26 pixels");
  require(coils >= 1 && coils <= kCalibMaxChannels, "sensitivity maps take 1..64 coils");
  require(window % 2 == 1 && window <= 31, "the sensitivity window is an odd side of at most 31 pixels");
  require(iterations <= 64, "at most 64 power iterations");
  uint32_t cp = 1;
  while (cp < coils) cp *= 2;
  const uint64_t npix = static_cast<uint64_t>(rows) * cols;
  const dim3 block(256), grid(static_cast<uint32_t>((npix + 4 * (64 / cp) - 1) / (4 * (64 / cp))));
  const int r = static_cast<int>(window / 2);
  switch (cp) {
    case 1: hipLaunchKernelGGL(walsh_kernel<1>, grid, block, 0, stream, d_images, d_maps, rows, cols, coils, r, iterations); break;
    case 2: hipLaunchKernelGGL(walsh_kernel<2>, grid, block, 0, stream, d_images, d_maps, rows, cols, coils, r, iterations); break;
    case 4: hipLaunchKernelGGL(walsh_kernel<4>, grid, block, 0, stream, d_images, d_maps, rows, cols, coils, r, iterations); break;
    case 8: hipLaunchKernelGGL(walsh_kernel<8>, grid, block, 0, stream, d_images, d_maps, rows, cols, coils, r, iterations); break;
    case 16: hipLaunchKernelGGL(walsh_kernel<16>, grid, block, 0, stream, d_images, d_maps, rows, cols, coils, r, iterations); break;
    case 32: hipLaunchKernelGGL(walsh_kernel<32>, grid, block, 0, stream, d_images, d_maps, rows, cols, coils, r, iterations); break;
    default: hipLaunchKernelGGL(walsh_kernel<64>, grid, block, 0, stream, d_images, d_maps, rows, cols, coils, r, iterations); break;
  }
  KSJ_LAUNCH_CHECK();
}

}  // namespace ksj::gpu
