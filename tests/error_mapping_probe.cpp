// Host-only probe of the library's error mapping (built with g++ by
// test_error_mapping.py; no GPU call is made): prints the KSJ_GPU_* status the C
// ABI returns for a HIP out-of-memory error, another HIP error and a violated
// precondition.
#include <cstdio>

#include "common.hpp"
#include "ksj_gpu.h"

static int status_after(void (*f)()) {
  try {
    f();
  } catch (const ksj::gpu::Error& e) {
    return ksj::gpu::status_of(e.code);
  }
  return KSJ_GPU_OK;
}

int main() {
  std::printf("%d %d %d %d\n",
              status_after([] { ksj::gpu::check(hipErrorOutOfMemory, "hipMalloc"); }),
              status_after([] { ksj::gpu::check(hipErrorInvalidValue, "hipMemcpy"); }),
              status_after([] { ksj::gpu::require(false, "precondition"); }),
              status_after([] { ksj::gpu::check(hipSuccess, "ok"); }));
  std::printf("%d %d %d %d\n", KSJ_GPU_RESOURCE_EXHAUSTED, KSJ_GPU_DEVICE_ERROR, KSJ_GPU_INVALID_ARGUMENT, KSJ_GPU_OK);
  return 0;
}
