// Shared device/host helpers for the KSpaceJet MI355X (gfx950) kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>

namespace ksj::gpu {

constexpr double kPi = 3.14159265358979323846;
constexpr double kTwoPi = 6.28318530717958647692;

struct cf {
  float x, y;
};

__host__ __device__ inline float2 cmul(float2 a, float2 b) { return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }
__host__ __device__ inline float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__host__ __device__ inline float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__host__ __device__ inline float2 cscale(float2 a, float s) { return make_float2(a.x * s, a.y * s); }
// |a|^2 with one explicit fma, so every kernel that sums it rounds the same way.
__host__ __device__ inline float norm2(float2 a) { return fmaf(a.x, a.x, a.y * a.y); }
// a kept (keep = ~0) or replaced by +0 (keep = 0) through its bit pattern: an integer mask,
// which the compiler cannot turn back into a branch around the load that produced a (a
// branch per load serialises a thread's loads).
__device__ inline float2 keep_bits(float2 a, uint32_t keep) {
  return make_float2(__builtin_bit_cast(float, __builtin_bit_cast(uint32_t, a.x) & keep),
                     __builtin_bit_cast(float, __builtin_bit_cast(uint32_t, a.y) & keep));
}
__host__ __device__ inline float2 cconj(float2 a) { return make_float2(a.x, -a.y); }

// Thrown by host code on a HIP error or a violated precondition; the C ABI
// layer turns it into a status code and a last-error message.
enum ErrorCode : int {
  kErrArgument = 1,   // a violated precondition (KSJ_GPU_INVALID_ARGUMENT)
  kErrDevice = 3,     // any other HIP error (KSJ_GPU_DEVICE_ERROR)
  kErrExhausted = 4,  // device memory or an HBM budget ran out (KSJ_GPU_RESOURCE_EXHAUSTED)
};

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

// KSJ_GPU_* status of an Error code (the values of include/ksj_gpu.h).
inline int status_of(int error_code) {
  switch (error_code) {
    case kErrArgument: return 1;
    case kErrExhausted: return 4;
    default: return 3;
  }
}

inline void check(hipError_t e, const char* what) {
  if (e == hipSuccess) return;
  throw Error(e == hipErrorOutOfMemory ? kErrExhausted : kErrDevice, std::string(what) + ": " + hipGetErrorString(e));
}

inline void require(bool ok, const char* what) {
  if (!ok) throw Error(kErrArgument, what);
}

#define KSJ_HIP(call) ::ksj::gpu::check((call), #call)
#define KSJ_LAUNCH_CHECK() ::ksj::gpu::check(hipGetLastError(), "kernel launch")

inline unsigned ceil_div(uint64_t a, uint64_t b) { return static_cast<unsigned>((a + b - 1) / b); }

// Process-wide launch parameters the autotuner may change (tune.hip).  Each
// value has a fixed set of legal settings; every setting gives bit-identical
// results except where a parameter's comment says otherwise.
enum TuneKey : int {
  kTuneGatherMap = 0,     // 0: items in launch order; 1: equal item ranges per XCD; 2: equal-cost ranges
  kTuneItemVisits,        // default gather work-item size of new plans (visits; summation order changes)
  kTuneFftLinesPerBlock,  // 0: per-length default; else lines per LDS FFT block (power of two)
  kTuneGatherBatch,       // gather visits per batch (sample-row loads in flight per lane): 4, 8 or 16
  kTunePassLanes,         // streams over which the passes of a batched adjoint overlap: 1..3
  kTuneMultiLanes,        // streams over which the items of a multi-plan adjoint are spread: 1..8
  kTuneGatherCoilsPerLane,  // coils each gather lane loads per visit for 32+ coils: 1 or 2
  kTuneGatherScheme,      // schedule of new plans: 0 = 2x2 blocks, 1 = LDS tiles, 2 = 4x4 blocks on the matrix cores
                          // (summation order changes)
  kTuneTileItemSamples,   // window samples per tile work item of new plans (summation order changes)
  kTuneTileWidth,         // tile columns of new plans: 4 (16 x 4 tiles) or 8 (8 x 8) (summation order changes)
  kTuneBlockGroup,        // 2x2-block group side of new plans: 1, 2 or 4 (work-item order; results unchanged)
  kTuneFftSchedule,       // FFT radix schedule: 0 (default), 1 (half the threads), 2 (radix-4 heavy) or 3 (per length:
                          // 1 at 512, else 0); rounding changes
  kTuneMultiFused,        // uneven batches: 0 = items over lanes, 1 = one launch per stage and shape class,
                          // 2 = as 1 with prep, gather and reduce in one work-queue launch (same bits)
  kTuneMatrixRunCost,     // matrix schedule of new plans: cost (visit groups of four + 1 per item) per wave run, 1..64
  kTuneMatrixPipeline,    // matrix gather: batches of sample rows in flight per wave, 1 or 2
  kTuneMatrixNtStore,     // matrix gather: 1 = grid cells written with non-temporal stores
  kTuneMatrixSlab,        // matrix gather, gather_map 0: consecutive workgroups per XCD slab (1 = launch order)
  kTuneMatrixItemVisits,  // matrix schedule: work-item size of new plans (visits; summation order changes)
  kTuneMatrixBatch,       // matrix gather: visits per batch, 4, 8 or 16
  kTuneMatrixBlockWidth,  // matrix schedule of new plans: 4 x 4 or 4 x 8 cell blocks (summation order changes)
  kTuneGather3Scheme,     // 3-D schedule of new plans: 0 = 2x2x2 blocks, 2 = 4x4x4 on the matrix cores
  kTuneMatrix3ItemVisits, // 3-D matrix schedule: work-item size of new plans (visits; summation order changes)
  kTuneMatrix3RunCost,    // 3-D matrix schedule of new plans: cost per wave run, 1..64
  kTuneMixScheme,         // channel mix: 0 = vector FMAs, 1 = matrix cores (summation order changes)
  kTuneMatrixWeights,     // 2-D matrix schedule of new plans: 0 = 16 (or 32) cell weights stored per visit,
                          // 1 = packed (sample, window position) visits, weights formed from the taps (same bits)
  kTuneMatrixPersistent,  // 2-D matrix gather: 0 = one wave per run; n = n resident workgroups per CU looping over
                          // their XCD's runs (same bits)
  kTuneMatrixTapShare,    // 2-D matrix gather, packed visits: 1 = one tap load per lane shared by bpermute (same bits)
  kTuneServeCopyThreads,  // servers created from now on: memcpy worker threads (beside the caller) for host requests
  kTuneFftRegister,       // 1 = register-staged FFT where loads and stores both run across lines (same bits),
                          // 0 = the LDS-staged form everywhere
  kTuneFftLdsPad,         // register-form FFT: LDS element stride = lines per block + this (1..4; same bits)
  kTuneMatrix3Map,        // 3-D matrix gather: 0 launch order, 1 XCD-contiguous eighths, s > 1 XCD slabs (same bits)
  kTuneMatrix3TapShare,   // 3-D matrix gather: 1 = one or two tap words per lane shared by bpermute (same bits)
  kTuneGridPrune,         // 2-D matrix schedule: 1 = cells no sample reaches (around column G/2) are neither written
                          // by the gather nor read by the row FFT, which takes them as zero (same bits)
  kTuneRowOrder,          // coil-fastest row-pass output of the column-epilogue pipeline: 0 = grid-row major
                          // [G_r][N_c][Cp], 1 = column major [N_c][G_r][Cp] (same bits)
  kTuneAssembleRows,      // acquisition assembly: channel rows of one acquisition per workgroup (1, 2, 4; same bits)
  kTuneKeys
};
int64_t tune_get(TuneKey key);
uint32_t device_cu_count();  // compute units of the current device (cached per device)
void tune_set(TuneKey key, int64_t value);  // throws Error on an illegal value
int tune_key(const char* name);             // -1 for an unknown name

}  // namespace ksj::gpu
