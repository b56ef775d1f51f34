// extern "C" boundary of libksj_gpu.so: argument checks, exception -> status mapping.
#include "../../include/ksj_gpu.h"

#include "calib.hpp"
#include "combine.hpp"
#include "condition.hpp"
#include "fft.hpp"
#include "grid.hpp"
#include "nudft.hpp"
#include "plan.hpp"
#include "plan3.hpp"
#include "prof.hpp"
#include "sim3.hpp"
#include "recon.hpp"
#include "recon3.hpp"
#include "serve.hpp"

#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <vector>

struct ksj_gpu_plan2 {
  ksj::gpu::Plan2* impl;
};

struct ksj_gpu_plan3 {
  ksj::gpu::Plan3* impl;
};

struct ksj_gpu_server {
  ksj::gpu::Server* impl;
};

struct ksj_gpu_graph {
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

namespace {

thread_local std::string g_last_error;

int fail(int code, const std::string& message) {
  g_last_error = message;
  return code;
}

template <typename F> int guarded(F&& f) {
  try {
    f();
    return KSJ_GPU_OK;
  } catch (const ksj::gpu::Error& e) {
    return fail(ksj::gpu::status_of(e.code), e.what());
  } catch (const std::bad_alloc&) {
    return fail(KSJ_GPU_RESOURCE_EXHAUSTED, "host allocation failed");
  } catch (const std::exception& e) {
    return fail(KSJ_GPU_DEVICE_ERROR, e.what());
  } catch (...) {
    return fail(KSJ_GPU_DEVICE_ERROR, "unknown error");
  }
}

hipStream_t as_stream(void* s) { return static_cast<hipStream_t>(s); }

// Workspace of ksj_gpu_adjoint2_execute_static: a file-scope device array, one copy per
// GPU the library runs on, held until the process ends.  In host code `scratch` names a
// host-side stand-in, never device memory: its device address on each GPU comes from
// hipGetSymbolAddress (at ksj_gpu_init, or at the entry point's first call on a device)
// and only that address is handed to kernels.
constexpr int kMaxScratchDevices = 64;
__device__ __align__(256) unsigned char scratch[KSJ_GPU_STATIC_SCRATCH_BYTES];
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxScratchDevices] = {};

// Device address of `scratch` on the current device (looked up once per device).
void* scratch_for_current_device() {
  int device = 0;
  KSJ_HIP(hipGetDevice(&device));
  ksj::gpu::require(device >= 0 && device < kMaxScratchDevices, "device ordinal beyond the scratch table");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_ptr[device] == nullptr) {
    void* ptr = nullptr;
    KSJ_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    ksj::gpu::require(ptr != nullptr, "scratch array has no device address");
    g_scratch_ptr[device] = ptr;
  }
  return g_scratch_ptr[device];
}

}  // namespace

extern "C" {

int ksj_gpu_init(int device_ordinal) {
  return guarded([&] {
    int count = 0;
    KSJ_HIP(hipGetDeviceCount(&count));
    ksj::gpu::require(device_ordinal >= 0 && device_ordinal < count, "device ordinal out of range");
    KSJ_HIP(hipSetDevice(device_ordinal));
    KSJ_HIP(hipFree(nullptr));
    scratch_for_current_device();
    ksj::gpu::grid_chain_prepare(nullptr);
  });
}

const char* ksj_gpu_last_error(void) { return g_last_error.c_str(); }

int ksj_gpu_stream_synchronize(void* stream) {
  return guarded([&] { KSJ_HIP(hipStreamSynchronize(as_stream(stream))); });
}

size_t ksj_gpu_ifft2_workspace_bytes(uint32_t channels, uint32_t rows, uint32_t cols) {
  return ksj::gpu::ifft2_workspace_bytes(channels, rows, cols);
}

int ksj_gpu_ifft2_coil_images(const void* d_kspace, void* d_images, uint32_t channels, uint32_t rows, uint32_t cols,
                              int centred, int layout, void* d_workspace, size_t workspace_bytes, void* stream) {
  return guarded([&] {
    ksj::gpu::require(d_kspace && d_images && d_workspace, "null device pointer");
    ksj::gpu::ifft2_coil_images(static_cast<const float2*>(d_kspace), static_cast<float2*>(d_images), channels, rows,
                                cols, centred != 0, layout, d_workspace, workspace_bytes, as_stream(stream));
  });
}

int ksj_gpu_rss(const void* d_coil_images, float* d_image, uint64_t pixels, uint32_t channels, void* stream) {
  return guarded([&] {
    ksj::gpu::require(d_coil_images && d_image, "null device pointer");
    ksj::gpu::rss_coil_last(static_cast<const float2*>(d_coil_images), d_image, pixels, channels, as_stream(stream));
  });
}

int ksj_gpu_sense_combine(const void* d_coil_images, const void* d_sensitivities, void* d_image, uint64_t pixels,
                          uint32_t channels, void* stream) {
  return guarded([&] {
    ksj::gpu::require(d_coil_images && d_sensitivities && d_image, "null device pointer");
    ksj::gpu::sense_combine(static_cast<const float2*>(d_coil_images), static_cast<const float2*>(d_sensitivities),
                            static_cast<float2*>(d_image), pixels, channels, as_stream(stream));
  });
}

int ksj_gpu_ramp_dcf2(const float* d_traj, float* d_dcf, uint64_t samples, void* stream) {
  return guarded([&] {
    ksj::gpu::require(d_traj && d_dcf, "null device pointer");
    ksj::gpu::ramp_dcf2(d_traj, d_dcf, samples, as_stream(stream));
  });
}

size_t ksj_gpu_nudft2_adjoint_workspace_bytes(uint32_t channels, uint32_t rows, uint32_t cols) {
  return ksj::gpu::nudft2_adjoint_workspace_bytes(channels, rows, cols);
}

int ksj_gpu_nudft2_adjoint(const float* d_traj, const void* d_samples, void* d_images, uint32_t channels,
                           uint32_t samples, uint32_t rows, uint32_t cols, double row_origin, double col_origin,
                           int layout, void* d_workspace, size_t workspace_bytes, void* stream) {
  return guarded([&] {
    ksj::gpu::require(d_traj && d_samples && d_images, "null device pointer");
    ksj::gpu::require(layout == KSJ_GPU_LAYOUT_COIL_MAJOR || layout == KSJ_GPU_LAYOUT_COIL_LAST, "unknown layout");
    if (layout == KSJ_GPU_LAYOUT_COIL_MAJOR) {
      ksj::gpu::nudft2_adjoint(d_traj, static_cast<const float2*>(d_samples), static_cast<float2*>(d_images), samples,
                               channels, rows, cols, row_origin, col_origin, as_stream(stream));
      return;
    }
    ksj::gpu::require(d_workspace && workspace_bytes >= ksj::gpu::nudft2_adjoint_workspace_bytes(channels, rows, cols),
                      "direct adjoint workspace too small");
    auto* tmp = static_cast<float2*>(d_workspace);
    ksj::gpu::nudft2_adjoint(d_traj, static_cast<const float2*>(d_samples), tmp, samples, channels, rows, cols,
                             row_origin, col_origin, as_stream(stream));
    ksj::gpu::coil_major_to_last(tmp, static_cast<float2*>(d_images), nullptr, static_cast<uint64_t>(rows) * cols,
                                 channels, as_stream(stream));
  });
}

int ksj_gpu_nudft2_forward_f64(const float* d_traj, const void* d_image, void* d_samples, uint32_t channels,
                               uint32_t samples, uint32_t rows, uint32_t cols, double row_origin, double col_origin,
                               void* stream) {
  return guarded([&] {
    ksj::gpu::require(d_traj && d_image && d_samples, "null device pointer");
    ksj::gpu::nudft2_forward_f64(d_traj, static_cast<const double2*>(d_image), static_cast<double2*>(d_samples),
                                 samples, channels, rows, cols, row_origin, col_origin, as_stream(stream));
  });
}

int ksj_gpu_golden_angle2(float* d_traj, uint32_t spokes, uint32_t readout, void* stream) {
  return guarded([&] {
    ksj::gpu::require(d_traj != nullptr, "null device pointer");
    ksj::gpu::golden_angle2(d_traj, spokes, readout, as_stream(stream));
  });
}

int ksj_gpu_plan2_create(const ksj_gpu_plan2_desc* desc, ksj_gpu_plan2** out_plan) {
  return guarded([&] {
    ksj::gpu::require(desc && out_plan, "null argument");
    ksj::gpu::require(desc->struct_size >= sizeof(ksj_gpu_plan2_desc), "plan descriptor too small");
    *out_plan = nullptr;
    ksj::gpu::Plan2Desc d;
    d.image_rows = desc->image_rows;
    d.image_cols = desc->image_cols;
    d.grid_rows = desc->grid_rows;
    d.grid_cols = desc->grid_cols;
    d.width = desc->kernel_width;
    d.kernel = desc->kernel;
    d.beta = desc->beta;
    d.row_origin = desc->row_origin;
    d.col_origin = desc->col_origin;
    d.density = desc->density;
    d.centred = desc->centred;
    d.samples = desc->samples;
    d.item_visits = desc->item_visits;  // 0: the schedule's tuning key
    ksj::gpu::require(d.density <= KSJ_GPU_DENSITY_GIVEN, "unknown density mode");
    auto* handle = new ksj_gpu_plan2{nullptr};
    try {
      handle->impl = ksj::gpu::plan2_create(d);
    } catch (...) {
      delete handle;
      throw;
    }
    *out_plan = handle;
  });
}

int ksj_gpu_plan2_set_trajectory(ksj_gpu_plan2* plan, const float* d_traj, const float* d_dcf, void* stream) {
  return guarded([&] {
    ksj::gpu::require(plan && plan->impl && d_traj, "null argument");
    ksj::gpu::plan2_set_trajectory(plan->impl, d_traj, d_dcf, as_stream(stream));
    // The FFT tables of the plan's grid lengths exist from here on, so executing the
    // plan allocates nothing (ksj_gpu_adjoint2_execute_static relies on it).
    ksj::gpu::fft_twiddles(plan->impl->d.grid_rows, as_stream(stream));
    ksj::gpu::fft_twiddles(plan->impl->d.grid_cols, as_stream(stream));
  });
}

int ksj_gpu_plan2_get_info(const ksj_gpu_plan2* plan, ksj_gpu_plan2_info* out_info) {
  return guarded([&] {
    ksj::gpu::require(plan && plan->impl && out_info, "null argument");
    ksj::gpu::require(out_info->struct_size >= sizeof(ksj_gpu_plan2_info), "info struct too small");
    out_info->items = plan->impl->item_count;
    out_info->visits = plan->impl->visit_count;
    const uint64_t cells =
      plan->impl->packed_visits ? 0 : static_cast<uint64_t>(plan->impl->block_h) * plan->impl->block_w;
    out_info->device_bytes = plan->impl->device_bytes +
                             plan->impl->visit_capacity * (sizeof(uint32_t) + cells * sizeof(float)) +
                             static_cast<uint64_t>(plan->impl->item_capacity) * (sizeof(uint4) + sizeof(uint32_t));
    out_info->beta = plan->impl->d.beta;
  });
}

int ksj_gpu_plan2_export_bins(const ksj_gpu_plan2* plan, uint32_t* d_cell_start, uint32_t* d_perm, void* stream) {
  return guarded([&] {
    ksj::gpu::require(plan && plan->impl && plan->impl->has_trajectory, "plan has no trajectory");
    const auto* p = plan->impl;
    if (d_cell_start)
      KSJ_HIP(hipMemcpyAsync(d_cell_start, p->cell_start, (p->nbins + 1) * sizeof(uint32_t), hipMemcpyDeviceToDevice,
                             as_stream(stream)));
    if (d_perm)
      KSJ_HIP(hipMemcpyAsync(d_perm, p->src, static_cast<size_t>(p->d.samples) * sizeof(uint32_t),
                             hipMemcpyDeviceToDevice, as_stream(stream)));
  });
}

void ksj_gpu_plan2_destroy(ksj_gpu_plan2* plan) {
  if (!plan) return;
  ksj::gpu::plan2_destroy(plan->impl);
  delete plan;
}

size_t ksj_gpu_adjoint2_workspace_bytes(const ksj_gpu_plan2* plan, uint32_t channels) {
  if (!plan || !plan->impl) return 0;
  return ksj::gpu::adjoint2_workspace_bytes(*plan->impl, channels, 1);
}

int ksj_gpu_adjoint2_execute(const ksj_gpu_plan2* plan, const void* d_kspace, uint32_t channels, void* d_images,
                             int layout, float* d_rss, void* d_workspace, size_t workspace_bytes, void* stream) {
  return guarded([&] {
    ksj::gpu::require(plan && plan->impl && d_kspace && d_workspace, "null argument");
    ksj::gpu::require(d_images || d_rss, "adjoint needs an output");
    ksj::gpu::AdjointScan scan;
    scan.kspace = static_cast<const float2*>(d_kspace);
    scan.images = static_cast<float2*>(d_images);
    scan.rss = d_rss;
    ksj::gpu::adjoint2_execute(*plan->impl, &scan, 1, channels, layout, d_workspace, workspace_bytes,
                               as_stream(stream));
  });
}

int ksj_gpu_adjoint2_execute_static(const ksj_gpu_plan2* plan, const void* d_kspace, uint32_t channels,
                                    void* d_images, int layout, float* d_rss, void* stream) {
  return guarded([&] {
    ksj::gpu::require(plan && plan->impl && d_kspace, "null argument");
    ksj::gpu::require(d_images || d_rss, "adjoint needs an output");
    ksj::gpu::require(channels >= 1 && channels <= 64, "adjoint supports 1..64 channels");
    const size_t need = ksj::gpu::adjoint2_workspace_bytes(*plan->impl, channels, 1);
    if (need > KSJ_GPU_STATIC_SCRATCH_BYTES)
      throw ksj::gpu::Error(ksj::gpu::kErrExhausted,
                            "workspace of " + std::to_string(need) + " bytes exceeds the static scratch of " +
                              std::to_string(static_cast<size_t>(KSJ_GPU_STATIC_SCRATCH_BYTES)) + " bytes");
    void* ws = scratch_for_current_device();
    ksj::gpu::AdjointScan scan;
    scan.kspace = static_cast<const float2*>(d_kspace);
    scan.images = static_cast<float2*>(d_images);
    scan.rss = d_rss;
    ksj::gpu::adjoint2_execute(*plan->impl, &scan, 1, channels, layout, ws, KSJ_GPU_STATIC_SCRATCH_BYTES,
                               as_stream(stream));
  });
}

size_t ksj_gpu_calibration_workspace_bytes(uint32_t channels, uint64_t samples) {
  return ksj::gpu::calibration_workspace_bytes(channels, samples);
}

int ksj_gpu_channel_covariance(const void* d_samples, uint32_t channels, uint64_t samples, int centred,
                               void* d_covariance, void* d_workspace, size_t workspace_bytes, void* stream) {
  return guarded([&] {
    ksj::gpu::require(d_covariance != nullptr, "null argument");
    ksj::gpu::require(centred == 0 || centred == 1, "centred must be 0 or 1");
    ksj::gpu::channel_covariance(static_cast<const float2*>(d_samples), channels, samples,
                                 centred ? ksj::gpu::kCovarianceNoise : ksj::gpu::kCovarianceSecondMoment,
                                 static_cast<double2*>(d_covariance), d_workspace, workspace_bytes, as_stream(stream));
  });
}

int ksj_gpu_hermitian_eigen(const void* d_matrix, uint32_t n, double* d_values, void* d_vectors, uint32_t* d_status,
                            void* stream) {
  return guarded([&] {
    ksj::gpu::hermitian_eigen(static_cast<const double2*>(d_matrix), n, d_values, static_cast<double2*>(d_vectors),
                              d_status, as_stream(stream));
  });
}

int ksj_gpu_noise_model_estimate(const void* d_samples, uint32_t channels, uint64_t samples, float eigenvalue_floor,
                                 void* d_whitening, void* d_workspace, size_t workspace_bytes, void* stream) {
  return guarded([&] {
    ksj::gpu::noise_model_estimate(static_cast<const float2*>(d_samples), channels, samples, eigenvalue_floor,
                                   static_cast<float2*>(d_whitening), d_workspace, workspace_bytes, as_stream(stream));
  });
}

int ksj_gpu_coil_compression_basis_estimate(const void* d_samples, uint32_t physical_channels,
                                            uint32_t virtual_channels, uint64_t samples, void* d_basis,
                                            void* d_workspace, size_t workspace_bytes, void* stream) {
  return guarded([&] {
    ksj::gpu::coil_compression_basis_estimate(static_cast<const float2*>(d_samples), physical_channels,
                                              virtual_channels, samples, static_cast<float2*>(d_basis), d_workspace,
                                              workspace_bytes, as_stream(stream));
  });
}

int ksj_gpu_sensitivity_maps_walsh(const void* d_coil_images, uint32_t rows, uint32_t cols, uint32_t channels,
                                   uint32_t window, uint32_t iterations, void* d_maps, void* stream) {
  return guarded([&] {
    ksj::gpu::sensitivity_maps_walsh(static_cast<const float2*>(d_coil_images), rows, cols, channels, window,
                                     iterations, static_cast<float2*>(d_maps), as_stream(stream));
  });
}

int ksj_gpu_phase_correction_estimate(const void* d_reference, uint32_t channels, uint32_t lines, uint32_t readout,
                                      void* d_phase, void* stream) {
  return guarded([&] {
    ksj::gpu::phase_correction_estimate(static_cast<const float2*>(d_reference), channels, lines, readout,
                                        static_cast<float2*>(d_phase), as_stream(stream));
  });
}

int ksj_gpu_plan3_create(const ksj_gpu_plan3_desc* desc, ksj_gpu_plan3** out_plan) {
  return guarded([&] {
    ksj::gpu::require(desc && out_plan, "null argument");
    ksj::gpu::require(desc->struct_size >= sizeof(ksj_gpu_plan3_desc), "plan descriptor too small");
    ksj::gpu::Plan3Desc d;
    d.image[0] = desc->image_z;
    d.image[1] = desc->image_y;
    d.image[2] = desc->image_x;
    d.grid[0] = desc->grid_z;
    d.grid[1] = desc->grid_y;
    d.grid[2] = desc->grid_x;
    d.width = desc->kernel_width;
    d.density = desc->density;
    d.samples = desc->samples;
    d.item_visits = desc->item_visits;  // 0: the schedule's tuning key
    d.beta = desc->beta;
    d.origin[0] = desc->origin_z;
    d.origin[1] = desc->origin_y;
    d.origin[2] = desc->origin_x;
    auto* handle = new ksj_gpu_plan3{nullptr};
    try {
      handle->impl = ksj::gpu::plan3_create(d);
    } catch (...) {
      delete handle;
      throw;
    }
    *out_plan = handle;
  });
}

int ksj_gpu_plan3_set_trajectory(ksj_gpu_plan3* plan, const float* d_traj, const float* d_dcf, void* stream) {
  return guarded([&] {
    ksj::gpu::require(plan && plan->impl, "null plan");
    ksj::gpu::plan3_set_trajectory(plan->impl, d_traj, d_dcf, as_stream(stream));
  });
}

int ksj_gpu_plan3_get_info(const ksj_gpu_plan3* plan, ksj_gpu_plan3_info* out_info) {
  return guarded([&] {
    ksj::gpu::require(plan && plan->impl && out_info, "null argument");
    ksj::gpu::require(out_info->struct_size >= sizeof(ksj_gpu_plan3_info), "info struct too small");
    const auto* p = plan->impl;
    out_info->items = p->item_count;
    out_info->visits = p->visit_count;
    const uint64_t per_visit = p->scheme == ksj::gpu::kSchemeMatrix ? 2 * sizeof(uint32_t)
                                                                    : sizeof(uint32_t) + 2 * sizeof(float4);
    const uint64_t run_tables = p->scheme == ksj::gpu::kSchemeMatrix
                                  ? (static_cast<uint64_t>(p->item_capacity) + p->run_capacity + 2) * sizeof(uint32_t)
                                  : 0;
    out_info->device_bytes = p->device_bytes + p->visit_capacity * per_visit +
                             static_cast<uint64_t>(p->item_capacity) * sizeof(uint4) + run_tables;
    out_info->beta = p->d.beta;
  });
}

int ksj_gpu_plan3_export_bins(const ksj_gpu_plan3* plan, uint32_t* d_cell_start, uint32_t* d_perm, void* stream) {
  return guarded([&] {
    ksj::gpu::require(plan && plan->impl && plan->impl->has_trajectory, "plan has no trajectory");
    const auto* p = plan->impl;
    if (d_cell_start)
      KSJ_HIP(hipMemcpyAsync(d_cell_start, p->cell_start, (p->nbins + 1) * sizeof(uint32_t), hipMemcpyDeviceToDevice,
                             as_stream(stream)));
    if (d_perm)
      KSJ_HIP(hipMemcpyAsync(d_perm, p->src, static_cast<size_t>(p->d.samples) * sizeof(uint32_t),
                             hipMemcpyDeviceToDevice, as_stream(stream)));
  });
}

void ksj_gpu_plan3_destroy(ksj_gpu_plan3* plan) {
  if (!plan) return;
  ksj::gpu::plan3_destroy(plan->impl);
  delete plan;
}

size_t ksj_gpu_adjoint3_workspace_bytes(const ksj_gpu_plan3* plan, uint32_t channels, uint32_t coils_per_pass) {
  size_t bytes = 0;
  const int rc = guarded([&] {
    ksj::gpu::require(plan && plan->impl, "null plan");
    bytes = ksj::gpu::adjoint3_workspace_bytes(*plan->impl, channels, coils_per_pass);
  });
  return rc == KSJ_GPU_OK ? bytes : 0;
}

int ksj_gpu_adjoint3_execute(const ksj_gpu_plan3* plan, const void* d_kspace, uint32_t channels,
                             uint32_t coils_per_pass, void* d_images, float* d_rss, void* d_workspace,
                             size_t workspace_bytes, void* stream) {
  return guarded([&] {
    ksj::gpu::require(plan && plan->impl && d_kspace && d_workspace, "null argument");
    ksj::gpu::adjoint3_execute(*plan->impl, static_cast<const float2*>(d_kspace), channels, coils_per_pass,
                               static_cast<float2*>(d_images), d_rss, d_workspace, workspace_bytes,
                               as_stream(stream));
  });
}

int ksj_gpu_normalize_trajectory2(const float* d_raw, uint64_t samples, int units, uint32_t encoded_rows,
                                  uint32_t encoded_cols, float* d_out, void* stream) {
  return guarded([&] {
    ksj::gpu::require(samples == 0 || (d_raw && d_out), "null trajectory");
    ksj::gpu::normalize_trajectory2(d_raw, samples, units, encoded_rows, encoded_cols, d_out, as_stream(stream));
  });
}

size_t ksj_gpu_assemble_workspace_bytes(uint32_t acquisitions) {
  return ksj::gpu::assemble_workspace_bytes(acquisitions);
}

int ksj_gpu_assemble_acquisitions(const void* d_acquisitions, const uint32_t* samples_per_acquisition,
                                  uint32_t acquisitions, uint32_t channels, void* d_kspace, void* d_workspace,
                                  size_t workspace_bytes, void* stream) {
  return guarded([&] {
    ksj::gpu::assemble_acquisitions(static_cast<const float2*>(d_acquisitions), samples_per_acquisition,
                                    acquisitions, channels, static_cast<float2*>(d_kspace), d_workspace,
                                    workspace_bytes, as_stream(stream));
  });
}

int ksj_gpu_trajectory_check(const float* d_traj, const float* d_resident, uint64_t values, float bound,
                             uint32_t* d_flags, void* stream) {
  return guarded([&] {
    ksj::gpu::require(d_flags != nullptr && (values == 0 || d_traj != nullptr), "null argument");
    ksj::gpu::require(bound >= 0.f, "bound must be non-negative");
    ksj::gpu::trajectory_check(d_traj, d_resident, values, bound, d_flags, as_stream(stream));
  });
}

int ksj_gpu_channel_mix(const void* d_matrix, const void* d_kspace, void* d_output, uint32_t out_channels,
                        uint32_t in_channels, uint64_t samples, void* stream) {
  return guarded([&] {
    ksj::gpu::require(samples == 0 || (d_matrix && d_kspace && d_output), "null argument");
    ksj::gpu::require(samples == 0 || d_kspace != d_output, "output must not overlap the input");
    ksj::gpu::channel_mix(static_cast<const float2*>(d_matrix), static_cast<const float2*>(d_kspace),
                          static_cast<float2*>(d_output), out_channels, in_channels, samples, as_stream(stream));
  });
}

int ksj_gpu_phase_correct(const void* d_kspace, const void* d_phase, void* d_output, uint32_t channels,
                          uint32_t rows, uint32_t cols, void* stream) {
  return guarded([&] {
    ksj::gpu::require(uint64_t{channels} * rows * cols == 0 || (d_kspace && d_phase && d_output), "null argument");
    ksj::gpu::phase_correct(static_cast<const float2*>(d_kspace), static_cast<const float2*>(d_phase),
                            static_cast<float2*>(d_output), channels, rows, cols, as_stream(stream));
  });
}

int ksj_gpu_readout_crop(const void* d_kspace, void* d_output, uint64_t lines, uint32_t in_cols, uint32_t out_cols,
                         uint32_t offset, void* stream) {
  return guarded([&] {
    ksj::gpu::require(lines == 0 || (d_kspace && d_output), "null argument");
    ksj::gpu::readout_crop(static_cast<const float2*>(d_kspace), static_cast<float2*>(d_output), lines, in_cols,
                           out_cols, offset, as_stream(stream));
  });
}

int ksj_gpu_golden_means3(float* d_traj, uint32_t spokes, uint32_t readout, void* stream) {
  return guarded([&] {
    ksj::gpu::require(d_traj != nullptr, "null trajectory");
    ksj::gpu::golden_means3(d_traj, spokes, readout, as_stream(stream));
  });
}

size_t ksj_gpu_adjoint2_batch_workspace_bytes(const ksj_gpu_plan2* plan, uint32_t channels, uint32_t scans) {
  if (!plan || !plan->impl) return 0;
  return ksj::gpu::adjoint2_workspace_bytes(*plan->impl, channels, scans);
}

int ksj_gpu_adjoint2_execute_batch(const ksj_gpu_plan2* plan, uint32_t count, const void* const* d_kspace,
                                   uint32_t channels, void* const* d_images, int layout, float* const* d_rss,
                                   void* d_workspace, size_t workspace_bytes, void* stream) {
  return guarded([&] {
    ksj::gpu::require(plan && plan->impl && d_kspace && d_workspace && count >= 1, "null argument");
    std::vector<ksj::gpu::AdjointScan> scans(count);
    for (uint32_t i = 0; i < count; ++i) {
      scans[i].kspace = static_cast<const float2*>(d_kspace[i]);
      scans[i].images = d_images ? static_cast<float2*>(d_images[i]) : nullptr;
      scans[i].rss = d_rss ? d_rss[i] : nullptr;
      ksj::gpu::require(scans[i].images || scans[i].rss, "every scan needs an output");
    }
    ksj::gpu::adjoint2_execute(*plan->impl, scans.data(), count, channels, layout, d_workspace, workspace_bytes,
                               as_stream(stream));
  });
}

}  // extern "C"

namespace {
std::vector<ksj::gpu::MultiItem> multi_items(uint32_t count, const ksj_gpu_plan2* const* plans,
                                             const uint32_t* channels, const void* const* d_kspace,
                                             void* const* d_images, float* const* d_rss) {
  ksj::gpu::require(count >= 1 && plans && channels, "multi adjoint needs plans and channels");
  std::vector<ksj::gpu::MultiItem> items(count);
  for (uint32_t i = 0; i < count; ++i) {
    ksj::gpu::require(plans[i] && plans[i]->impl, "null plan in a multi adjoint");
    ksj::gpu::require(channels[i] >= 1 && channels[i] <= 64, "multi adjoint supports 1..64 channels per item");
    items[i].plan = plans[i]->impl;
    items[i].channels = channels[i];
    if (d_kspace) {
      ksj::gpu::require(d_kspace[i] != nullptr, "null k-space in a multi adjoint");
      items[i].scan.kspace = static_cast<const float2*>(d_kspace[i]);
      items[i].scan.images = d_images ? static_cast<float2*>(d_images[i]) : nullptr;
      items[i].scan.rss = d_rss ? d_rss[i] : nullptr;
      ksj::gpu::require(items[i].scan.images || items[i].scan.rss, "every multi item needs an output");
    }
  }
  return items;
}
}  // namespace

extern "C" {

size_t ksj_gpu_adjoint2_multi_workspace_bytes(uint32_t count, const ksj_gpu_plan2* const* plans,
                                              const uint32_t* channels) {
  size_t out = 0;
  const int rc = guarded([&] {
    const auto items = multi_items(count, plans, channels, nullptr, nullptr, nullptr);
    out = ksj::gpu::adjoint2_multi_workspace_bytes(items.data(), count);
  });
  return rc == KSJ_GPU_OK ? out : 0;
}

int ksj_gpu_adjoint2_execute_multi(uint32_t count, const ksj_gpu_plan2* const* plans, const void* const* d_kspace,
                                   const uint32_t* channels, void* const* d_images, int layout, float* const* d_rss,
                                   void* d_workspace, size_t workspace_bytes, void* stream) {
  return guarded([&] {
    ksj::gpu::require(d_kspace && d_workspace, "null argument");
    const auto items = multi_items(count, plans, channels, d_kspace, d_images, d_rss);
    ksj::gpu::adjoint2_execute_multi(items.data(), count, layout, d_workspace, workspace_bytes, as_stream(stream));
  });
}

int ksj_gpu_chain_faults(uint64_t* out_count) {
  return guarded([&] {
    ksj::gpu::require(out_count != nullptr, "null argument");
    *out_count = ksj::gpu::grid_chain_faults();
  });
}

size_t ksj_gpu_nufft3_forward_f64_workspace_bytes(const uint32_t* image) {
  if (image == nullptr) return 0;
  return ksj::gpu::nufft3_forward_f64_workspace_bytes(image);
}

int ksj_gpu_nufft3_forward_f64(const void* d_image, const uint32_t* image, const double* origin, const float* d_traj,
                               uint64_t samples, uint32_t width, void* d_out, void* d_workspace,
                               size_t workspace_bytes, void* stream) {
  return guarded([&] {
    ksj::gpu::require(image != nullptr && origin != nullptr, "null argument");
    ksj::gpu::nufft3_forward_f64(static_cast<const double2*>(d_image), image, origin, d_traj, samples,
                                 static_cast<int>(width), static_cast<double2*>(d_out), d_workspace, workspace_bytes,
                                 as_stream(stream));
  });
}

int ksj_gpu_graph_begin(void* stream) {
  return guarded([&] {
    ksj::gpu::require(stream != nullptr, "graph capture needs a stream other than the null stream");
    KSJ_HIP(hipStreamBeginCapture(as_stream(stream), hipStreamCaptureModeThreadLocal));
  });
}

int ksj_gpu_graph_end(void* stream, ksj_gpu_graph** out_graph) {
  return guarded([&] {
    ksj::gpu::require(stream != nullptr && out_graph != nullptr, "null argument");
    auto* g = new ksj_gpu_graph{};
    try {
      KSJ_HIP(hipStreamEndCapture(as_stream(stream), &g->graph));
      KSJ_HIP(hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0));
    } catch (...) {
      if (g->graph) (void)hipGraphDestroy(g->graph);
      delete g;
      throw;
    }
    *out_graph = g;
  });
}

int ksj_gpu_graph_launch(ksj_gpu_graph* graph, void* stream) {
  return guarded([&] {
    ksj::gpu::require(graph && graph->exec, "null graph");
    KSJ_HIP(hipGraphLaunch(graph->exec, as_stream(stream)));
  });
}

void ksj_gpu_graph_destroy(ksj_gpu_graph* graph) {
  if (!graph) return;
  if (graph->exec) (void)hipGraphExecDestroy(graph->exec);
  if (graph->graph) (void)hipGraphDestroy(graph->graph);
  delete graph;
}

int ksj_gpu_serve_create(const ksj_gpu_serve_config* config, ksj_gpu_server** out_server) {
  return guarded([&] {
    ksj::gpu::require(config && out_server, "null argument");
    ksj::gpu::require(config->struct_size >= sizeof(ksj_gpu_serve_config), "config struct too small");
    ksj::gpu::ServeConfig c;
    c.device = config->device;
    c.hbm_budget_bytes = config->hbm_budget_bytes;
    c.lanes = config->lanes ? config->lanes : 2;
    c.max_batch_scans = config->max_batch_scans ? config->max_batch_scans : 8;
    c.plan_cache_entries = config->plan_cache_entries ? config->plan_cache_entries : 16;
    c.host_staging_bytes = uint64_t{config->host_staging_mib ? config->host_staging_mib : 1024u} << 20;
    auto* h = new ksj_gpu_server{nullptr};
    try {
      h->impl = new ksj::gpu::Server(c);
    } catch (...) {
      delete h;
      throw;
    }
    *out_server = h;
  });
}

}  // extern "C"

namespace {
ksj::gpu::ServeRequest to_request(const ksj_gpu_serve_request* request) {
  // A struct that ends before `memory` (the layout without host-memory requests) is a
  // device-memory request.
  ksj::gpu::require(request->struct_size >= offsetof(ksj_gpu_serve_request, memory), "request struct too small");
  const uint32_t memory = request->struct_size >= offsetof(ksj_gpu_serve_request, reserved1)
                              ? request->memory
                              : KSJ_GPU_SERVE_DEVICE_MEMORY;
  ksj::gpu::require(memory == KSJ_GPU_SERVE_DEVICE_MEMORY || memory == KSJ_GPU_SERVE_HOST_MEMORY,
                    "request memory must be device or host");
  ksj::gpu::ServeRequest r;
  r.kind = request->kind;
  r.image_rows = request->image_rows;
  r.image_cols = request->image_cols;
  r.channels = request->channels;
  r.samples = request->samples;
  r.kernel_width = request->kernel_width;
  r.density = request->density;
  r.centred = request->centred;
  r.image_depth = request->image_depth;
  r.trajectory_key = request->trajectory_key;
  r.d_traj = request->d_traj;
  r.d_kspace = static_cast<const float2*>(request->d_kspace);
  r.d_images = static_cast<float2*>(request->d_images);
  r.d_rss = request->d_rss;
  r.host = memory == KSJ_GPU_SERVE_HOST_MEMORY;
  return r;
}
}  // namespace

extern "C" {

int ksj_gpu_serve_submit(ksj_gpu_server* server, const ksj_gpu_serve_request* request, uint64_t* ticket) {
  return guarded([&] {
    ksj::gpu::require(server && server->impl && request && ticket, "null argument");
    *ticket = server->impl->submit(to_request(request));
  });
}

int ksj_gpu_serve_submit_batch(ksj_gpu_server* server, uint32_t count, const ksj_gpu_serve_request* requests,
                               uint64_t* tickets) {
  return guarded([&] {
    ksj::gpu::require(server && server->impl && requests && tickets && count >= 1, "null argument");
    std::vector<ksj::gpu::ServeRequest> rs;
    rs.reserve(count);
    for (uint32_t i = 0; i < count; ++i) rs.push_back(to_request(&requests[i]));
    server->impl->submit_many(rs.data(), count, tickets);
  });
}

int ksj_gpu_serve_wait(ksj_gpu_server* server, uint64_t ticket, double* latency_ms) {
  int status = KSJ_GPU_OK;
  const int rc = guarded([&] {
    ksj::gpu::require(server && server->impl, "null argument");
    std::string error;
    status = server->impl->wait(ticket, latency_ms, &error);
    if (status != KSJ_GPU_OK) g_last_error = error;
  });
  return rc != KSJ_GPU_OK ? rc : status;
}

int ksj_gpu_serve_drain(ksj_gpu_server* server) {
  return guarded([&] {
    ksj::gpu::require(server && server->impl, "null argument");
    server->impl->drain();
  });
}

int ksj_gpu_serve_hold(ksj_gpu_server* server, int on) {
  return guarded([&] {
    ksj::gpu::require(server && server->impl, "null argument");
    server->impl->hold(on != 0);
  });
}

int ksj_gpu_serve_get_stats(ksj_gpu_server* server, ksj_gpu_serve_stats* out) {
  return guarded([&] {
    ksj::gpu::require(server && server->impl && out, "null argument");
    ksj::gpu::require(out->struct_size >= sizeof(ksj_gpu_serve_stats), "stats struct too small");
    const ksj::gpu::ServeStats s = server->impl->stats();
    out->submitted = s.submitted;
    out->completed = s.completed;
    out->failed = s.failed;
    out->batches = s.batches;
    out->plan_builds = s.plan_builds;
    out->plan_hits = s.plan_hits;
    out->plan_evictions = s.plan_evictions;
    out->peak_device_bytes = s.peak_workspace_bytes;
    out->plan_bytes = s.plan_bytes;
    out->max_batch_seen = s.max_batch_seen;
  });
}

void ksj_gpu_serve_destroy(ksj_gpu_server* server) {
  if (!server) return;
  delete server->impl;
  delete server;
}

int ksj_gpu_tune_set(const char* key, int64_t value) {
  return guarded([&] {
    const int k = ksj::gpu::tune_key(key);
    ksj::gpu::require(k >= 0, "unknown tuning key");
    ksj::gpu::tune_set(static_cast<ksj::gpu::TuneKey>(k), value);
  });
}

int ksj_gpu_tune_get(const char* key, int64_t* value) {
  return guarded([&] {
    const int k = ksj::gpu::tune_key(key);
    ksj::gpu::require(k >= 0 && value != nullptr, "unknown tuning key or null output");
    *value = ksj::gpu::tune_get(static_cast<ksj::gpu::TuneKey>(k));
  });
}

int ksj_gpu_profile_enable(int on) {
  return guarded([&] { ksj::gpu::prof_enable(on != 0); });
}

int ksj_gpu_profile_reset(void) {
  return guarded([&] { ksj::gpu::prof_reset(); });
}

int ksj_gpu_profile_read(const char* stage, double* total_ms, uint64_t* count) {
  return guarded([&] {
    ksj::gpu::require(stage && total_ms && count, "null argument");
    ksj::gpu::prof_read(stage, total_ms, count);
  });
}

}  // extern "C"
