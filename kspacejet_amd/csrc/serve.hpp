// Continuous-batching reconstruction server on one GPU.
//
// Requests (2-D radial KB adjoint, 3-D radial KB adjoint of a volume, or 2-D
// Cartesian inverse FFT; inputs and outputs in HBM, or for 2-D kinds in host memory)
// are queued by submit() and run by one scheduler thread:
//  * admission: a request's batch is launched only when a lane (stream) is free and
//    the HBM budget covers its workspace next to the resident plans;
//  * plans (FP64 coordinates, bin sort, visit schedule, KB and deapodization tables)
//    stay resident per (geometry, trajectory key), least recently used evicted first;
//  * batching: pending radial requests with the same plan are stacked along the coil
//    axis into one launch (up to max_batch_scans), Cartesian requests of one shape
//    share a launch; requests arriving while batches run join the next free lane;
//  * lanes: up to `lanes` batches are in flight on separate default-priority streams,
//    so their kernels overlap on the device;
//  * host memory: submit() copies a host request's k-space into a pinned staging region
//    (blocking while the staging is full); its lane sends it to HBM ahead of the batch's
//    kernels and its outputs back after them, so one lane's copies overlap another lane's
//    kernels; the scheduler copies the outputs to the caller's buffers on completion.
// Every request's result equals its result when run alone (each coil lane's
// arithmetic is independent of what it is stacked with).
#pragma once

#include "plan.hpp"
#include "plan3.hpp"

#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstdint>
#include <deque>
#include <list>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

namespace ksj::gpu {

enum ServeKind : uint32_t { kServeRadial2D = 0, kServeCartesian2D = 1, kServeRadial3D = 2 };
constexpr uint32_t kServeCoilsPerPass3D = 16;  // a 512^3 grid of 16 coils is 17.2 GB

struct ServeConfig {
  int device = 0;
  uint64_t hbm_budget_bytes = 0;  // 0: 80% of the device's free memory at creation
  uint32_t lanes = 2;             // batches in flight at once (streams)
  uint32_t max_batch_scans = 8;   // scans stacked into one radial launch
  uint32_t plan_cache_entries = 16;
  uint64_t host_staging_bytes = 1ull << 30;  // pinned staging of host-memory requests
};

struct ServeRequest {
  uint32_t kind = kServeRadial2D;
  uint32_t image_rows = 0, image_cols = 0, channels = 0, samples = 0;
  uint32_t image_depth = 0;  // 3-D: N_z (image_rows = N_y, image_cols = N_x)
  uint32_t kernel_width = 4;
  uint32_t density = 1;  // KSJ_GPU_DENSITY_NONE / _RAMP
  uint32_t centred = 0;  // Cartesian only
  uint64_t trajectory_key = 0;  // radial: identifies the trajectory bytes; 0 = hash them
  const float* d_traj = nullptr;  // radial: [S][2] (ky, kx) or 3-D [S][3] (kz, ky, kx)
  const float2* d_kspace = nullptr;
  float2* d_images = nullptr;  // coil-last, may be null
  float* d_rss = nullptr;      // may be null
  bool host = false;           // the four buffers above are host memory (2-D kinds)
};

struct ServeStats {
  uint64_t submitted = 0, completed = 0, failed = 0;
  uint64_t batches = 0, plan_builds = 0, plan_hits = 0, plan_evictions = 0;
  uint64_t peak_workspace_bytes = 0, plan_bytes = 0;
  uint64_t max_batch_seen = 0;
};

// Parallel memcpy on a few worker threads plus the caller (copies of host requests into
// and out of pinned staging, where one thread reaches only part of the host's copy rate).
class CopyPool {
 public:
  explicit CopyPool(unsigned workers);
  ~CopyPool();
  void copy(void* dst, const void* src, size_t bytes);  // returns when every byte is copied

 private:
  struct Job {
    char* dst;
    const char* src;
    size_t bytes, chunks;
    std::atomic<size_t> next{0}, done{0};
  };
  void run(Job& j);
  std::mutex mu_, job_mu_;
  std::condition_variable cv_, cv_done_;
  std::shared_ptr<Job> job_;
  uint64_t generation_ = 0;
  bool stop_ = false;
  std::vector<std::thread> workers_;
};

class Server {
 public:
  explicit Server(const ServeConfig& cfg);
  ~Server();
  Server(const Server&) = delete;
  Server& operator=(const Server&) = delete;

  // Queues a request; returns its ticket (> 0).  Throws Error on an invalid request.
  uint64_t submit(const ServeRequest& r);
  // Queues several requests atomically (the scheduler sees all of them at once).
  void submit_many(const ServeRequest* r, uint32_t count, uint64_t* tickets);
  // Blocks until the ticket completes; returns 0 or an error code (ksj_gpu status)
  // and the submit-to-completion latency in ms.  A ticket can be waited on once.
  int wait(uint64_t ticket, double* latency_ms, std::string* error);
  // Blocks until every submitted request has completed.
  void drain();
  // While held, nothing new is launched (requests queue up and are batched on release).
  void hold(bool on);
  ServeStats stats();

 private:
  struct Pending {
    uint64_t ticket;
    ServeRequest req;
    uint64_t plan_key;
    std::chrono::steady_clock::time_point submitted;
    // Host-memory requests: k-space, RSS and coil images.  Each buffer has `len` bytes,
    // `dev` bytes (len rounded to 256) in the lane's device staging, and either goes by
    // DMA to or from the caller's buffer itself (`direct`: the caller's memory is pinned)
    // or through offset `at` of the server's pinned staging (region `slot`).
    struct HostBuf {
      void* user = nullptr;
      size_t len = 0, dev = 0, at = 0;
      bool direct = false;
    };
    HostBuf in, rss, img;
    size_t slot = SIZE_MAX;
    size_t staged_bytes() const {
      size_t b = 0;
      for (const HostBuf* h : {&in, &rss, &img}) b += h->direct ? 0 : h->dev;
      return b;
    }
    size_t device_bytes() const { return in.dev + rss.dev + img.dev; }
  };
  // First-fit allocator over one pinned host buffer (the staging of host-memory requests).
  struct Staging {
    std::mutex mu;  // taken after mu_ when both are held
    std::condition_variable cv;
    char* base = nullptr;
    size_t bytes = 0;
    std::map<size_t, size_t> used;  // offset -> bytes of a staged request's region
    bool stop = false;
  };
  struct Done {
    int status;
    double latency_ms;
    std::string error;
  };
  using PlanPtr = std::unique_ptr<Plan2, void (*)(Plan2*)>;
  using Plan3Ptr = std::unique_ptr<Plan3, void (*)(Plan3*)>;
  struct PlanEntry {
    PlanPtr plan{nullptr, &plan2_destroy};
    Plan3Ptr plan3{nullptr, &plan3_destroy};
    uint64_t bytes = 0;
    uint32_t users = 0;  // batches in flight that use it
  };
  struct Lane {
    hipStream_t stream = nullptr;
    hipEvent_t done = nullptr;
    void* ws = nullptr;
    size_t ws_bytes = 0;
    bool busy = false;
    std::vector<Pending> batch;
    uint64_t plan_key = 0;
    void* stage = nullptr;  // device copies of the batch's host-memory inputs and outputs
    size_t stage_bytes = 0;
  };

  Pending prepare(const ServeRequest& r);  // validates and keys a request (hashes a trajectory if needed)
  bool fit_slot(size_t bytes, size_t* offset);        // under staging_.mu
  bool try_reserve_slot(size_t bytes, size_t* offset);
  size_t reserve_slot(size_t bytes);                  // blocks while the staging is full
  void release_slot(size_t slot);
  void run();
  bool launch_next(Lane& lane);  // under mu_; returns false when nothing could be launched
  void complete(Lane& lane, int status, const std::string& error);
  Plan2* plan_for(const Pending& p, hipStream_t stream);
  Plan3* plan3_for(const Pending& p, hipStream_t stream);
  PlanEntry* cached_plan(uint64_t key);
  bool evict_one();
  void evict_until(uint64_t need_bytes);
  uint64_t used_bytes() const;
  static uint64_t key_of(const ServeRequest& r);

  ServeConfig cfg_;
  std::mutex mu_;
  std::condition_variable cv_work_, cv_done_;
  std::deque<Pending> pending_;
  std::unordered_map<uint64_t, Done> done_;
  std::vector<Lane> lanes_;
  std::map<uint64_t, PlanEntry> plans_;
  std::list<uint64_t> lru_;  // front = most recently used
  ServeStats stats_;
  Staging staging_;
  std::unique_ptr<CopyPool> copies_;  // parallel memcpy into and out of the staging
  uint64_t next_ticket_ = 1;
  uint64_t inflight_ = 0;
  bool stop_ = false;
  bool hold_ = false;
  std::thread thread_;
};

// 64-bit hash of a device trajectory of `values` floats (FNV-1a over its bytes,
// computed on the device and read back); synchronises `stream`.
uint64_t trajectory_hash(const float* d_traj, uint64_t values, hipStream_t stream);
// The same hash of a trajectory in host memory.
uint64_t trajectory_hash_host(const float* h_traj, uint64_t values);

}  // namespace ksj::gpu
