// ISMRMRD acquisitions -> the k-space frame [C][S] of a non-Cartesian route.
//
// The route's replay (libs/recon/kspacejet-recon-runtime/src/noncartesian_rss_hdf5.cpp:1077-1083)
// copies, for acquisition a of ns_a samples and each channel c, the ISMRMRD row
// samples[c ns_a, c ns_a + ns_a) (dataset_reader.hpp:57: sample + channel * ns) to
// frame[c S + o_a, c S + o_a + ns_a), where o_a counts the samples of the acquisitions
// before a and S all of them.  Here the acquisitions lie in HBM one after another, each in
// its own layout, and one launch moves them all: a workgroup per (acquisition, channel,
// chunk of the row) reads one contiguous run and writes one contiguous run, so both sides
// stream whole cache lines and no thread searches for its acquisition.
#include "common.hpp"
#include "plan.hpp"

#include <vector>

namespace ksj::gpu {
namespace {

// One wave per workgroup, each lane EPT accesses of 8 VEC bytes: a 512-sample row is one
// workgroup with the whole 4 KB row in flight at once (G rows of one acquisition: G times that).
constexpr uint32_t kAssembleThreads = 64;
// The reference route's channel limit (noncartesian_rss_hdf5.cpp:49, checked at :880-884).
constexpr uint32_t kAssembleMaxChannels = 64;
// Workgroups per launch: the launch's work items stay below 2^32.
constexpr uint64_t kAssembleMaxBlocks = (1ull << 32) / kAssembleThreads - 1;

struct AssembleArgs {
  const float2* packed;
  float2* frame;
  const uint32_t* offsets;  // [A + 1] first sample of each acquisition; null: ns samples each
  uint32_t acq0;            // first acquisition of the launch
  uint32_t channels, ns, parts;
  uint64_t total;  // S
};

// Workgroup b: rows [G r, G r + G) of the launch, r = b / parts (channels G r mod C of
// acquisition G r div C), chunk b % parts of EPT * 64 accesses of VEC complex values (VEC = 2
// only when every count is even and both buffers are 16-byte aligned, so every row starts on
// a 16-byte boundary).  The G rows of one acquisition lie one after another in its ISMRMRD
// block, so the workgroup reads one contiguous run of G n accesses; access i belongs to row
// i div n.  Every load is issued before the first store; an access past the run's end reads
// the run's last one (in range) and is not stored.
template <int EPT, int VEC, int G>
__global__ __launch_bounds__(kAssembleThreads) void assemble_kernel(AssembleArgs a) {
  typedef float FV __attribute__((ext_vector_type(2 * VEC)));
  const uint32_t b = blockIdx.x;
  const uint32_t rg = b / a.parts, part = b - rg * a.parts;
  const uint32_t groups = a.channels / G;
  const uint32_t acq = a.acq0 + rg / groups, c0 = (rg % groups) * G;
  uint64_t off;
  uint32_t n;
  if (a.offsets) {
    off = a.offsets[acq];
    n = static_cast<uint32_t>(a.offsets[acq + 1] - off);
  } else {
    off = static_cast<uint64_t>(acq) * a.ns;
    n = a.ns;
  }
  n /= VEC;  // accesses per row
  const uint32_t span = G * n;
  const uint32_t first = part * (kAssembleThreads * EPT);
  if (first >= span) return;  // the whole workgroup: a chunk past a short acquisition's rows
  const FV* src = reinterpret_cast<const FV*>(a.packed + off * a.channels) + static_cast<uint64_t>(c0) * n;
  FV v[EPT];
#pragma unroll
  for (int k = 0; k < EPT; ++k) {
    const uint32_t i = first + k * kAssembleThreads + threadIdx.x;
    // Read once: non-temporal, out of the caches' way (the frame is what the next kernel reads).
    v[k] = __builtin_nontemporal_load(src + (i < span ? i : span - 1));
  }
#pragma unroll
  for (int k = 0; k < EPT; ++k) {
    const uint32_t i = first + k * kAssembleThreads + threadIdx.x;
    if (i < span) {
      uint32_t r = 0;
#pragma unroll
      for (int j = 1; j < G; ++j) r += i >= j * n ? 1u : 0u;
      FV* dst = reinterpret_cast<FV*>(a.frame + static_cast<uint64_t>(c0 + r) * a.total + off);
      dst[i - r * n] = v[k];
    }
  }
}

template <int EPT, int VEC, int G>
void launch_assemble(AssembleArgs a, uint32_t acquisitions, hipStream_t stream) {
  const uint64_t per_acq = static_cast<uint64_t>(a.channels / G) * a.parts;
  const uint32_t chunk = static_cast<uint32_t>(std::min<uint64_t>(acquisitions, kAssembleMaxBlocks / per_acq));
  for (uint32_t a0 = 0; a0 < acquisitions; a0 += chunk) {
    a.acq0 = a0;
    const uint32_t count = std::min(chunk, acquisitions - a0);
    hipLaunchKernelGGL((assemble_kernel<EPT, VEC, G>), dim3(static_cast<unsigned>(count * per_acq)),
                       dim3(kAssembleThreads), 0, stream, a);
    KSJ_LAUNCH_CHECK();
  }
}

template <int VEC, int G>
void launch_assemble_ept(int ept, AssembleArgs a, uint32_t acquisitions, hipStream_t stream) {
  switch (ept) {
    case 1: launch_assemble<1, VEC, G>(a, acquisitions, stream); break;
    case 2: launch_assemble<2, VEC, G>(a, acquisitions, stream); break;
    case 4: launch_assemble<4, VEC, G>(a, acquisitions, stream); break;
    case 8: launch_assemble<8, VEC, G>(a, acquisitions, stream); break;
    default: launch_assemble<16, VEC, G>(a, acquisitions, stream); break;
  }
}

template <int VEC>
void launch_assemble_vec(int g, int ept, AssembleArgs a, uint32_t acquisitions, hipStream_t stream) {
  switch (g) {
    case 1: launch_assemble_ept<VEC, 1>(ept, a, acquisitions, stream); break;
    case 2: launch_assemble_ept<VEC, 2>(ept, a, acquisitions, stream); break;
    default: launch_assemble_ept<VEC, 4>(ept, a, acquisitions, stream); break;
  }
}

}  // namespace

size_t assemble_workspace_bytes(uint32_t acquisitions) {
  return 4 * (static_cast<size_t>(acquisitions) + 1);
}

void assemble_acquisitions(const float2* d_packed, const uint32_t* samples_per_acquisition, uint32_t acquisitions,
                           uint32_t channels, float2* d_frame, void* d_workspace, size_t workspace_bytes,
                           hipStream_t stream) {
  require(acquisitions >= 1, "a frame needs at least one acquisition");
  require(samples_per_acquisition != nullptr, "null sample counts");
  require(channels >= 1 && channels <= kAssembleMaxChannels, "acquisitions must have [1, 64] active channels");
  uint64_t total = 0;
  uint32_t max_ns = 0;
  bool uniform = true, all_even = true;
  for (uint32_t i = 0; i < acquisitions; ++i) {
    const uint32_t n = samples_per_acquisition[i];
    require(n >= 1, "every acquisition must have samples");
    total += n;
    max_ns = std::max(max_ns, n);
    uniform = uniform && n == samples_per_acquisition[0];
    all_even = all_even && n % 2 == 0;
  }
  // The route counts the frame's samples in 32 bits (noncartesian_rss_hdf5.cpp:1053, :1104).
  require(total <= 0xFFFFFFFFull, "the frame's samples must fit 32 bits");
  require(d_packed != nullptr && d_frame != nullptr, "null k-space");
  AssembleArgs a{};
  a.packed = d_packed;
  a.frame = d_frame;
  a.channels = channels;
  a.total = total;
  a.ns = samples_per_acquisition[0];
  const bool vec2 = all_even && reinterpret_cast<uintptr_t>(d_packed) % 16 == 0 &&
                    reinterpret_cast<uintptr_t>(d_frame) % 16 == 0;
  int g = static_cast<int>(tune_get(kTuneAssembleRows));  // rows per workgroup: the largest that divides C
  while (channels % g != 0) g /= 2;
  // Accesses in the longest run of g rows; up to 8 (g = 1) or 16 (g > 1) per lane cover it whole.
  const uint64_t max_span = static_cast<uint64_t>(g) * (vec2 ? max_ns / 2 : max_ns);
  int ept = 1;
  while (ept < (g > 1 ? 16 : 8) && static_cast<uint64_t>(ept) * kAssembleThreads < max_span) ept *= 2;
  const uint64_t parts = (max_span + static_cast<uint64_t>(kAssembleThreads) * ept - 1) / (kAssembleThreads * ept);
  require(parts <= 0xFFFFFFFFull && static_cast<uint64_t>(channels / g) * parts <= kAssembleMaxBlocks &&
              max_span < (1ull << 31),  // access indices of a run stay below 2^32
          "an acquisition is too long for one launch");
  a.parts = static_cast<uint32_t>(parts);
  std::vector<uint32_t> offsets;
  if (!uniform) {
    require(d_workspace != nullptr && workspace_bytes >= assemble_workspace_bytes(acquisitions),
            "workspace smaller than ksj_gpu_assemble_workspace_bytes");
    require(reinterpret_cast<uintptr_t>(d_workspace) % 4 == 0, "workspace must be 4-byte aligned");
    offsets.resize(static_cast<size_t>(acquisitions) + 1);
    offsets[0] = 0;
    for (uint32_t i = 0; i < acquisitions; ++i) offsets[i + 1] = offsets[i] + samples_per_acquisition[i];
    KSJ_HIP(hipMemcpyAsync(d_workspace, offsets.data(), offsets.size() * sizeof(uint32_t), hipMemcpyHostToDevice,
                           stream));
    a.offsets = static_cast<const uint32_t*>(d_workspace);
  }
  try {
    if (vec2)
      launch_assemble_vec<2>(g, ept, a, acquisitions, stream);
    else
      launch_assemble_vec<1>(g, ept, a, acquisitions, stream);
  } catch (...) {
    if (!uniform) (void)hipStreamSynchronize(stream);  // the copy reads `offsets`
    throw;
  }
  // The offsets were copied from host memory that ends with this call.
  if (!uniform) KSJ_HIP(hipStreamSynchronize(stream));
}

}  // namespace ksj::gpu
