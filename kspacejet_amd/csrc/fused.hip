#include "fused.hpp"

#include <algorithm>
#include <cstring>

namespace ksj::gpu {
namespace {

constexpr uint32_t kChunkWords = 768;  // 3 KiB of kernel arguments per launch

struct Chunk {
  uint32_t words;
  uint32_t data[kChunkWords];
};

__global__ void upload_kernel(uint32_t* dst, Chunk c) {
  for (uint32_t i = threadIdx.x; i < c.words; i += blockDim.x) dst[i] = c.data[i];
}

}  // namespace

void upload_table(void* dst, const void* src, size_t bytes, hipStream_t stream) {
  require(bytes % 4 == 0 && reinterpret_cast<uintptr_t>(dst) % 4 == 0, "table uploads move whole 32-bit words");
  const auto* s = static_cast<const uint32_t*>(src);
  auto* d = static_cast<uint32_t*>(dst);
  size_t words = bytes / 4;
  while (words > 0) {
    Chunk c;
    c.words = static_cast<uint32_t>(std::min<size_t>(words, kChunkWords));
    std::memcpy(c.data, s, c.words * sizeof(uint32_t));
    hipLaunchKernelGGL(upload_kernel, dim3(1), dim3(256), 0, stream, d, c);
    KSJ_LAUNCH_CHECK();
    s += c.words;
    d += c.words;
    words -= c.words;
  }
}

}  // namespace ksj::gpu
