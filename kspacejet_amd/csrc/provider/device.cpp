#include "device.hpp"

#include <algorithm>
#include <cstring>

namespace ksj::prov {

void check_gpu(int rc, const char* what) {
  if (rc == KSJ_GPU_OK) return;
  std::string msg = std::string(what) + ": " + ksj_gpu_last_error();
  throw DeviceFailure(msg, rc == KSJ_GPU_RESOURCE_EXHAUSTED);
}

void check_hip(hipError_t e, const char* what) {
  if (e == hipSuccess) return;
  throw DeviceFailure(std::string(what) + ": " + hipGetErrorString(e), e == hipErrorOutOfMemory);
}

DeviceSession::DeviceSession(uint32_t device_ordinal) : device_(device_ordinal) {
  check_gpu(ksj_gpu_init(static_cast<int>(device_ordinal)), "device initialisation");
  check_hip(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking), "hipStreamCreate");
  for (int i = 0; i < 2; ++i) {
    check_hip(hipHostMalloc(&stage_[i], kStageBytes, hipHostMallocDefault), "hipHostMalloc");
    check_hip(hipEventCreateWithFlags(&stage_done_[i], hipEventDisableTiming), "hipEventCreate");
  }
}

DeviceSession::~DeviceSession() {
  (void)hipSetDevice(static_cast<int>(device_));
  if (stream_) (void)hipStreamSynchronize(stream_);
  if (plan_) ksj_gpu_plan2_destroy(plan_);
  for (void* b : buffers_)
    if (b) (void)hipFree(b);
  for (int i = 0; i < 2; ++i) {
    if (stage_[i]) (void)hipHostFree(stage_[i]);
    if (stage_done_[i]) (void)hipEventDestroy(stage_done_[i]);
  }
  if (stream_) (void)hipStreamDestroy(stream_);
}

void DeviceSession::activate() const { check_hip(hipSetDevice(static_cast<int>(device_)), "hipSetDevice"); }

void* DeviceSession::buffer(Slot slot, size_t bytes) {
  bytes = std::max<size_t>(bytes, 256);
  if (sizes_[slot] >= bytes) return buffers_[slot];
  if (buffers_[slot]) {
    check_hip(hipStreamSynchronize(stream_), "hipStreamSynchronize");
    check_hip(hipFree(buffers_[slot]), "hipFree");
    buffers_[slot] = nullptr;
    sizes_[slot] = 0;
  }
  check_hip(hipMalloc(&buffers_[slot], bytes), "hipMalloc");
  sizes_[slot] = bytes;
  return buffers_[slot];
}

void DeviceSession::upload(void* d_dst, const void* h_src, size_t bytes) {
  const auto* src = static_cast<const unsigned char*>(h_src);
  auto* dst = static_cast<unsigned char*>(d_dst);
  int k = 0;
  for (size_t off = 0; off < bytes; off += kStageBytes, k ^= 1) {
    const size_t n = std::min(kStageBytes, bytes - off);
    // The staging buffer is free once its previous DMA has completed.
    check_hip(hipEventSynchronize(stage_done_[k]), "hipEventSynchronize");
    std::memcpy(stage_[k], src + off, n);
    check_hip(hipMemcpyAsync(dst + off, stage_[k], n, hipMemcpyHostToDevice, stream_), "hipMemcpyAsync H2D");
    check_hip(hipEventRecord(stage_done_[k], stream_), "hipEventRecord");
  }
  check_hip(hipStreamSynchronize(stream_), "hipStreamSynchronize");
}

void DeviceSession::download(void* h_dst, const void* d_src, size_t bytes) {
  const auto* src = static_cast<const unsigned char*>(d_src);
  auto* dst = static_cast<unsigned char*>(h_dst);
  const size_t chunks = (bytes + kStageBytes - 1) / kStageBytes;
  auto issue = [&](size_t c) {
    const size_t off = c * kStageBytes, n = std::min(kStageBytes, bytes - off);
    check_hip(hipMemcpyAsync(stage_[c & 1], src + off, n, hipMemcpyDeviceToHost, stream_), "hipMemcpyAsync D2H");
    check_hip(hipEventRecord(stage_done_[c & 1], stream_), "hipEventRecord");
  };
  if (chunks > 0) issue(0);
  for (size_t c = 0; c < chunks; ++c) {
    if (c + 1 < chunks) issue(c + 1);  // the DMA of the next chunk overlaps this memcpy
    check_hip(hipEventSynchronize(stage_done_[c & 1]), "hipEventSynchronize");
    const size_t off = c * kStageBytes, n = std::min(kStageBytes, bytes - off);
    std::memcpy(dst + off, stage_[c & 1], n);
  }
  check_hip(hipStreamSynchronize(stream_), "hipStreamSynchronize");
}

void DeviceSession::synchronize() { check_hip(hipStreamSynchronize(stream_), "hipStreamSynchronize"); }

bool DeviceSession::resident(const ksj_gpu_plan2_desc& desc, const float* h_traj) const {
  const size_t values = size_t{desc.samples} * 2;
  return plan_ != nullptr && std::memcmp(&plan_desc_, &desc, sizeof(desc)) == 0 && plan_traj_.size() == values &&
         std::memcmp(plan_traj_.data(), h_traj, values * 4) == 0;
}

uint32_t DeviceSession::check_device_trajectory(const ksj_gpu_plan2_desc& desc, const float* d_traj, float bound) {
  const bool same_desc = plan_ != nullptr && std::memcmp(&plan_desc_, &desc, sizeof(desc)) == 0 &&
                         plan_traj_.size() == size_t{desc.samples} * 2;
  auto* flags = static_cast<uint32_t*>(buffer(kFlags, sizeof(uint32_t)));
  check_gpu(ksj_gpu_trajectory_check(d_traj, same_desc ? trajectory() : nullptr, uint64_t{desc.samples} * 2, bound,
                                     flags, stream_),
            "trajectory check");
  // The staging buffers are idle between firings' transfers.
  check_hip(hipMemcpyAsync(stage_[0], flags, sizeof(uint32_t), hipMemcpyDeviceToHost, stream_), "hipMemcpyAsync D2H");
  check_hip(hipStreamSynchronize(stream_), "hipStreamSynchronize");
  uint32_t result = 0;
  std::memcpy(&result, stage_[0], sizeof(result));
  return result;
}

ksj_gpu_plan2* DeviceSession::plan(const ksj_gpu_plan2_desc& desc, const float* traj, bool on_device,
                                   bool is_resident) {
  if (is_resident && plan_ != nullptr) return plan_;
  if (plan_) {
    ksj_gpu_plan2_destroy(plan_);
    plan_ = nullptr;
  }
  plan_traj_.clear();
  const size_t values = size_t{desc.samples} * 2;
  void* d_traj = buffer(kTrajectory, values * 4);
  std::vector<float> host(values);
  if (on_device) {
    check_hip(hipMemcpyAsync(d_traj, traj, values * 4, hipMemcpyDeviceToDevice, stream_), "hipMemcpyAsync D2D");
    download(host.data(), d_traj, values * 4);
  } else {
    std::memcpy(host.data(), traj, values * 4);
    upload(d_traj, traj, values * 4);
  }
  ksj_gpu_plan2* p = nullptr;
  check_gpu(ksj_gpu_plan2_create(&desc, &p), "plan create");
  const int rc = ksj_gpu_plan2_set_trajectory(p, static_cast<const float*>(d_traj), nullptr, stream_);
  if (rc != KSJ_GPU_OK) {
    ksj_gpu_plan2_destroy(p);
    check_gpu(rc, "plan trajectory");
  }
  plan_ = p;
  plan_desc_ = desc;
  plan_traj_ = std::move(host);
  ++plan_builds_;
  return plan_;
}

}  // namespace ksj::prov
