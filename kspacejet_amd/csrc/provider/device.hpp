// Device-side state of one provider operator instance: its HIP device and
// stream, grow-only device buffers, pinned staging for the host-pageable
// payloads of the Provider ABI, and the resident gridding plan of the last
// trajectory it saw.
#pragma once

#include "ksj_gpu.h"

#include <hip/hip_runtime_api.h>

#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

namespace ksj::prov {

// A failure of the GPU library or the HIP runtime; `exhausted` marks an
// out-of-memory condition (reported as KSJ_STATUS_RESOURCE_EXHAUSTED).
struct DeviceFailure : std::runtime_error {
  bool exhausted;
  DeviceFailure(const std::string& what, bool oom) : std::runtime_error(what), exhausted(oom) {}
};

void check_gpu(int rc, const char* what);           // ksj_gpu_* status
void check_hip(hipError_t e, const char* what);     // HIP runtime status

class DeviceSession {
 public:
  enum Slot : int { kInput = 0, kInput2, kOutput, kWorkspace, kTrajectory, kFlags, kSlots };

  explicit DeviceSession(uint32_t device_ordinal);
  ~DeviceSession();
  DeviceSession(const DeviceSession&) = delete;
  DeviceSession& operator=(const DeviceSession&) = delete;

  // Makes this session's device current on the calling thread.
  void activate() const;
  hipStream_t stream() const { return stream_; }
  // Device buffer of at least `bytes` for `slot` (grown, never shrunk).
  void* buffer(Slot slot, size_t bytes);

  // Host-pageable <-> device copies through two pinned staging buffers, so the
  // host memcpy of one chunk overlaps the DMA of the previous one.  Both
  // return once the data has arrived.
  void upload(void* d_dst, const void* h_src, size_t bytes);
  void download(void* h_dst, const void* d_src, size_t bytes);
  void synchronize();

  // True when a plan for `desc` is resident and the host trajectory [samples][2]
  // at `h_traj` holds its trajectory's bytes.
  bool resident(const ksj_gpu_plan2_desc& desc, const float* h_traj) const;
  // The same question for a trajectory in this session's device memory, answered
  // on the GPU together with the value check (every value finite and within
  // [-bound, bound]): returns KSJ_GPU_TRAJ_DIFFERS / KSJ_GPU_TRAJ_INVALID bits.
  uint32_t check_device_trajectory(const ksj_gpu_plan2_desc& desc, const float* d_traj, float bound);
  // Plan for `desc` over the trajectory at `traj` (device memory when `on_device`,
  // else host): the resident plan when `is_resident` (the caller has found the
  // bytes equal with one of the two calls above), otherwise rebuilt on the device.
  // The trajectory itself stays resident in the kTrajectory slot, with a host copy
  // for later comparisons.
  ksj_gpu_plan2* plan(const ksj_gpu_plan2_desc& desc, const float* traj, bool on_device, bool is_resident);
  // Device copy of the trajectory of the resident plan (valid after plan()).
  const float* trajectory() const { return static_cast<const float*>(buffers_[kTrajectory]); }
  uint64_t plan_builds() const { return plan_builds_; }

 private:
  static constexpr size_t kStageBytes = size_t{8} << 20;
  uint32_t device_;
  hipStream_t stream_ = nullptr;
  void* buffers_[kSlots] = {};
  size_t sizes_[kSlots] = {};
  void* stage_[2] = {};
  hipEvent_t stage_done_[2] = {};
  ksj_gpu_plan2* plan_ = nullptr;
  ksj_gpu_plan2_desc plan_desc_{};
  std::vector<float> plan_traj_;
  uint64_t plan_builds_ = 0;
};

}  // namespace ksj::prov
