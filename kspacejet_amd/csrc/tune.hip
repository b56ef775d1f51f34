// Process-wide launch parameters (see TuneKey in common.hpp).
#include "common.hpp"

#include <atomic>
#include <cstring>

namespace ksj::gpu {

namespace {

struct Param {
  const char* name;
  int64_t value;
  bool (*legal)(int64_t);
};

bool is_pow2_or_zero(int64_t v) { return v == 0 || (v > 0 && v <= 64 && (v & (v - 1)) == 0); }

Param g_params[kTuneKeys] = {
    {"gather_map", 0, [](int64_t v) { return v >= 0 && v <= 2; }},
    {"item_visits", 64, [](int64_t v) { return v >= 1 && v <= 4096; }},
    {"fft_lines_per_block", 0, &is_pow2_or_zero},
    {"gather_batch", 8, [](int64_t v) { return v == 4 || v == 8 || v == 16; }},
    {"pass_lanes", 2, [](int64_t v) { return v >= 1 && v <= 3; }},
    {"multi_lanes", 4, [](int64_t v) { return v >= 1 && v <= 8; }},
    {"gather_coils_per_lane", 1, [](int64_t v) { return v == 1 || v == 2; }},
    {"gather_scheme", 2, [](int64_t v) { return v >= 0 && v <= 2; }},
    {"tile_item_samples", 256, [](int64_t v) { return v >= 1 && v <= 65536; }},
    {"tile_width", 4, [](int64_t v) { return v == 4 || v == 8; }},
    {"block_group", 2, [](int64_t v) { return v == 1 || v == 2 || v == 4; }},
    {"fft_schedule", 0, [](int64_t v) { return v >= 0 && v <= 3; }},
    {"multi_fused", 0, [](int64_t v) { return v >= 0 && v <= 2; }},
    {"matrix_run_cost", 12, [](int64_t v) { return v >= 1 && v <= 64; }},
    {"matrix_pipeline", 1, [](int64_t v) { return v == 1 || v == 2; }},
    {"matrix_nt_store", 1, [](int64_t v) { return v == 0 || v == 1; }},
    {"matrix_slab", 16, [](int64_t v) { return v >= 1 && v <= 4096; }},
    {"matrix_item_visits", 128, [](int64_t v) { return v >= 1 && v <= 4096; }},
    {"matrix_batch", 16, [](int64_t v) { return v == 4 || v == 8 || v == 16; }},
    {"matrix_block_width", 4, [](int64_t v) { return v == 4 || v == 8; }},
    {"gather3_scheme", 2, [](int64_t v) { return v == 0 || v == 2; }},
    {"matrix3_item_visits", 256, [](int64_t v) { return v >= 1 && v <= 4096; }},
    {"matrix3_run_cost", 32, [](int64_t v) { return v >= 1 && v <= 64; }},
    {"mix_scheme", 1, [](int64_t v) { return v == 0 || v == 1; }},
    {"matrix_weights", 1, [](int64_t v) { return v == 0 || v == 1; }},
    {"matrix_persistent", 0, [](int64_t v) { return v >= 0 && v <= 8; }},
    {"matrix_tap_share", 1, [](int64_t v) { return v == 0 || v == 1; }},
    {"serve_copy_threads", 7, [](int64_t v) { return v >= 0 && v <= 15; }},
    {"fft_register", 1, [](int64_t v) { return v == 0 || v == 1; }},
    {"fft_lds_pad", 1, [](int64_t v) { return v >= 1 && v <= 4; }},
    {"matrix3_map", 0, [](int64_t v) { return v >= 0 && v <= 4096; }},
    {"matrix3_tap_share", 0, [](int64_t v) { return v == 0 || v == 1; }},
    {"grid_prune", 1, [](int64_t v) { return v == 0 || v == 1; }},
    {"row_order", 0, [](int64_t v) { return v == 0 || v == 1; }},
    {"assemble_rows", 4, [](int64_t v) { return v == 1 || v == 2 || v == 4; }},
};

}  // namespace

int64_t tune_get(TuneKey key) { return g_params[key].value; }

void tune_set(TuneKey key, int64_t value) {
  require(key >= 0 && key < kTuneKeys, "unknown tuning key");
  require(g_params[key].legal(value), "illegal tuning value");
  g_params[key].value = value;
}

uint32_t device_cu_count() {
  constexpr int kMaxDevices = 64;
  static std::atomic<uint32_t> cached[kMaxDevices] = {};
  int device = 0;
  KSJ_HIP(hipGetDevice(&device));
  require(device >= 0 && device < kMaxDevices, "device ordinal beyond the CU-count table");
  uint32_t n = cached[device].load(std::memory_order_relaxed);
  if (n == 0) {
    int cus = 0;
    KSJ_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device));
    n = static_cast<uint32_t>(cus > 0 ? cus : 1);
    cached[device].store(n, std::memory_order_relaxed);
  }
  return n;
}

int tune_key(const char* name) {
  for (int k = 0; k < kTuneKeys; ++k)
    if (name != nullptr && std::strcmp(name, g_params[k].name) == 0) return k;
  return -1;
}

}  // namespace ksj::gpu
