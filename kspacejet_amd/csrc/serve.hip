#include "serve.hpp"

#include "combine.hpp"
#include "grid.hpp"
#include "recon.hpp"
#include "recon3.hpp"

#include <algorithm>
#include <cstring>

namespace ksj::gpu {

namespace {

constexpr int kHashThreads = 1024;

// FNV-1a over 32-bit words: each thread hashes one contiguous chunk, thread 0
// folds the chunk hashes in order, so the value is fixed for given bytes.
__global__ __launch_bounds__(kHashThreads) void hash_kernel(const uint32_t* words, uint64_t n, uint64_t* out) {
  __shared__ uint64_t part[kHashThreads];
  const uint64_t per = (n + kHashThreads - 1) / kHashThreads;
  const uint64_t b = threadIdx.x * per, e = b + per < n ? b + per : n;
  uint64_t h = 1469598103934665603ull;
  for (uint64_t i = b; i < e; ++i) h = (h ^ words[i]) * 1099511628211ull;
  part[threadIdx.x] = h;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint64_t f = 1469598103934665603ull ^ n;
    for (int t = 0; t < kHashThreads; ++t) f = (f ^ part[t]) * 1099511628211ull;
    *out = f;
  }
}

uint64_t mix(uint64_t h, uint64_t v) {
  h ^= v + 0x9E3779B97F4A7C15ull + (h << 6) + (h >> 2);
  return h;
}

uint64_t round_up(uint64_t v, uint64_t a) { return (v + a - 1) / a * a; }

// True when [p, p + len) lies in one pinned (page-locked) host allocation, which the
// copy engines read and write directly.
bool pinned_host(const void* p, uint64_t len) {
  hipPointerAttribute_t a{}, b{};
  const char* first = static_cast<const char*>(p);
  const char* last = first + len - 1;
  for (auto [attr, q] : {std::pair<hipPointerAttribute_t*, const char*>{&a, first}, {&b, last}}) {
    if (hipPointerGetAttributes(attr, q) != hipSuccess) {
      (void)hipGetLastError();  // pageable memory: clear the error this thread now holds
      return false;
    }
    if (attr->type != hipMemoryTypeHost || attr->hostPointer == nullptr) return false;
  }
  // Both ends in the same allocation: the same offset from the start of the range.
  return static_cast<const char*>(b.hostPointer) - static_cast<const char*>(a.hostPointer) == last - first;
}

constexpr size_t kCopyChunk = 2u << 20;  // bytes per memcpy task
constexpr size_t kCopyParallelMin = 4u << 20;

double ms_since(std::chrono::steady_clock::time_point t) {
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t).count();
}

}  // namespace

uint64_t trajectory_hash(const float* d_traj, uint64_t values, hipStream_t stream) {
  uint64_t* d_out = nullptr;
  uint64_t h = 0;
  KSJ_HIP(hipMallocAsync(reinterpret_cast<void**>(&d_out), sizeof(uint64_t), stream));
  hipLaunchKernelGGL(hash_kernel, dim3(1), dim3(kHashThreads), 0, stream, reinterpret_cast<const uint32_t*>(d_traj),
                     values, d_out);
  KSJ_LAUNCH_CHECK();
  KSJ_HIP(hipMemcpyAsync(&h, d_out, sizeof(h), hipMemcpyDeviceToHost, stream));
  KSJ_HIP(hipFreeAsync(d_out, stream));
  KSJ_HIP(hipStreamSynchronize(stream));
  return h;
}

uint64_t trajectory_hash_host(const float* h_traj, uint64_t values) {
  // The same chunking and fold as hash_kernel, so a trajectory keys one plan whether it
  // arrives in host or device memory.
  const auto* words = reinterpret_cast<const uint32_t*>(h_traj);
  const uint64_t per = (values + kHashThreads - 1) / kHashThreads;
  uint64_t f = 1469598103934665603ull ^ values;
  for (uint64_t t = 0; t < kHashThreads; ++t) {
    const uint64_t b = t * per, e = std::min(b + per, values);
    uint64_t h = 1469598103934665603ull;
    for (uint64_t i = b; i < e; ++i) h = (h ^ words[i]) * 1099511628211ull;
    f = (f ^ h) * 1099511628211ull;
  }
  return f;
}

CopyPool::CopyPool(unsigned workers) {
  for (unsigned i = 0; i < workers; ++i)
    workers_.emplace_back([this] {
      uint64_t seen = 0;
      while (true) {
        std::shared_ptr<Job> j;
        {
          std::unique_lock<std::mutex> lock(mu_);
          cv_.wait(lock, [&] { return stop_ || generation_ != seen; });
          if (stop_) return;
          seen = generation_;
          j = job_;
        }
        if (j) run(*j);
      }
    });
}

CopyPool::~CopyPool() {
  {
    std::lock_guard<std::mutex> lock(mu_);
    stop_ = true;
  }
  cv_.notify_all();
  for (std::thread& t : workers_) t.join();
}

void CopyPool::run(Job& j) {
  while (true) {
    const size_t c = j.next.fetch_add(1);
    if (c >= j.chunks) return;
    const size_t b = c * kCopyChunk, n = std::min(kCopyChunk, j.bytes - b);
    std::memcpy(j.dst + b, j.src + b, n);
    if (j.done.fetch_add(1) + 1 == j.chunks) {
      std::lock_guard<std::mutex> lock(mu_);
      cv_done_.notify_all();
    }
  }
}

void CopyPool::copy(void* dst, const void* src, size_t bytes) {
  if (bytes < kCopyParallelMin || workers_.empty()) {
    std::memcpy(dst, src, bytes);
    return;
  }
  std::lock_guard<std::mutex> one(job_mu_);  // one parallel copy at a time
  auto j = std::make_shared<Job>();
  j->dst = static_cast<char*>(dst);
  j->src = static_cast<const char*>(src);
  j->bytes = bytes;
  j->chunks = (bytes + kCopyChunk - 1) / kCopyChunk;
  {
    std::lock_guard<std::mutex> lock(mu_);
    job_ = j;
    ++generation_;
  }
  cv_.notify_all();
  run(*j);
  std::unique_lock<std::mutex> lock(mu_);
  cv_done_.wait(lock, [&] { return j->done.load() == j->chunks; });
  job_.reset();
}

uint64_t Server::key_of(const ServeRequest& r) {
  uint64_t h = mix(0, r.kind);
  for (uint64_t v : {uint64_t{r.image_rows}, uint64_t{r.image_cols}, uint64_t{r.image_depth}, uint64_t{r.samples},
                     uint64_t{r.kernel_width}, uint64_t{r.density}, uint64_t{r.centred}, r.trajectory_key})
    h = mix(h, v);
  return h;
}

Server::Server(const ServeConfig& cfg) : cfg_(cfg) {
  require(cfg_.lanes >= 1 && cfg_.lanes <= 8, "server lanes must be in [1, 8]");
  require(cfg_.max_batch_scans >= 1 && cfg_.max_batch_scans <= 64, "max batch scans must be in [1, 64]");
  require(cfg_.plan_cache_entries >= 1, "the plan cache needs at least one entry");
  KSJ_HIP(hipSetDevice(cfg_.device));
  if (cfg_.hbm_budget_bytes == 0) {
    size_t free_b = 0, total_b = 0;
    KSJ_HIP(hipMemGetInfo(&free_b, &total_b));
    cfg_.hbm_budget_bytes = free_b / 5 * 4;
  }
  lanes_.resize(cfg_.lanes);
  for (Lane& l : lanes_) {
    KSJ_HIP(hipStreamCreateWithFlags(&l.stream, hipStreamNonBlocking));
    KSJ_HIP(hipEventCreateWithFlags(&l.done, hipEventDisableTiming));
  }
  copies_ = std::make_unique<CopyPool>(static_cast<unsigned>(tune_get(kTuneServeCopyThreads)));
  thread_ = std::thread([this] { run(); });
}

Server::~Server() {
  {
    std::lock_guard<std::mutex> lock(mu_);
    stop_ = true;
  }
  {
    std::lock_guard<std::mutex> lock(staging_.mu);
    staging_.stop = true;
  }
  staging_.cv.notify_all();
  cv_work_.notify_all();
  if (thread_.joinable()) thread_.join();
  (void)hipSetDevice(cfg_.device);
  for (Lane& l : lanes_) {
    if (l.stream) (void)hipStreamSynchronize(l.stream);
    if (l.ws) (void)hipFree(l.ws);
    if (l.stage) (void)hipFree(l.stage);
    if (l.done) (void)hipEventDestroy(l.done);
    if (l.stream) (void)hipStreamDestroy(l.stream);
  }
  plans_.clear();
  if (staging_.base) (void)hipHostFree(staging_.base);
}

Server::Pending Server::prepare(const ServeRequest& r) {
  require(r.kind == kServeRadial2D || r.kind == kServeCartesian2D || r.kind == kServeRadial3D,
          "unknown request kind");
  require(r.channels >= 1 && (r.kind == kServeRadial3D || r.channels <= 64),
          "request channels must be in [1, 64] (3-D volumes: any count, taken in passes)");
  require(r.image_rows >= 2 && r.image_cols >= 2 && r.image_rows <= 4096 && r.image_cols <= 4096,
          "request image dimensions out of range");
  require(r.d_kspace != nullptr && (r.d_images != nullptr || r.d_rss != nullptr), "request needs k-space and an output");
  require(!r.host || r.kind != kServeRadial3D, "host-memory requests are 2-D (3-D volumes are served from HBM)");
  Pending p{};
  p.req = r;
  if (r.kind == kServeRadial2D || r.kind == kServeRadial3D) {
    const bool three = r.kind == kServeRadial3D;
    require(r.samples >= 1 && r.d_traj != nullptr, "radial request needs a trajectory");
    require(r.kernel_width == 4 || r.kernel_width == 6 || r.kernel_width == 8, "kernel width must be 4, 6 or 8");
    require(three ? (r.density == kDensityNone || r.density == kDensityRampSquared)
                  : (r.density == kDensityNone || r.density == kDensityRamp),
            "density must be none or ramp (2-D) / ramp squared (3-D)");
    require(!three || (r.image_depth >= 2 && r.image_depth <= 1024), "3-D request needs an image depth in [2, 1024]");
    if (!three) p.req.image_depth = 0;
    if (p.req.trajectory_key == 0 && r.host) {
      p.req.trajectory_key = trajectory_hash_host(r.d_traj, static_cast<uint64_t>(r.samples) * 2) | 1ull;
    } else if (p.req.trajectory_key == 0) {
      KSJ_HIP(hipSetDevice(cfg_.device));
      hipStream_t s = nullptr;
      KSJ_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
      p.req.trajectory_key = trajectory_hash(r.d_traj, static_cast<uint64_t>(r.samples) * (three ? 3 : 2), s) | 1ull;
      KSJ_HIP(hipStreamDestroy(s));
    }
  } else {
    p.req.samples = 0;
    p.req.image_depth = 0;
    p.req.trajectory_key = 0;
  }
  p.plan_key = key_of(p.req);
  if (r.host) {
    const uint64_t pixels = static_cast<uint64_t>(r.image_rows) * r.image_cols;
    const uint64_t values = r.kind == kServeRadial2D ? uint64_t{r.samples} : pixels;
    auto set = [](Pending::HostBuf& h, const void* user, uint64_t len) {
      if (!user) return;
      h.user = const_cast<void*>(user);
      h.len = len;
      h.dev = round_up(len, 256);
      h.direct = pinned_host(user, len);
    };
    set(p.in, r.d_kspace, values * r.channels * sizeof(float2));
    set(p.rss, r.d_rss, pixels * sizeof(float));
    set(p.img, r.d_images, pixels * r.channels * sizeof(float2));
  }
  return p;
}

// First fit of `bytes` in the pinned staging (allocated at the first use), under
// staging_.mu; returns false when no gap holds it.
bool Server::fit_slot(size_t bytes, size_t* offset) {
  if (!staging_.base) {
    KSJ_HIP(hipSetDevice(cfg_.device));
    void* base = nullptr;
    KSJ_HIP(hipHostMalloc(&base, cfg_.host_staging_bytes, hipHostMallocDefault));
    staging_.base = static_cast<char*>(base);
    staging_.bytes = cfg_.host_staging_bytes;
  }
  require(bytes <= staging_.bytes, "host-memory request larger than the server's pinned staging");
  size_t at = 0;  // first gap in offset order that holds the region
  for (const auto& kv : staging_.used) {
    if (kv.first - at >= bytes) break;
    at = kv.first + kv.second;
  }
  if (at + bytes > staging_.bytes) return false;
  staging_.used[at] = bytes;
  *offset = at;
  return true;
}

bool Server::try_reserve_slot(size_t bytes, size_t* offset) {
  std::lock_guard<std::mutex> lock(staging_.mu);
  return fit_slot(bytes, offset);
}

// Reserves one region of the pinned staging, blocking while the staging is full.
size_t Server::reserve_slot(size_t bytes) {
  std::unique_lock<std::mutex> lock(staging_.mu);
  size_t at = 0;
  while (true) {
    require(!staging_.stop, "server is shutting down");
    if (fit_slot(bytes, &at)) return at;
    staging_.cv.wait(lock);
  }
}

void Server::release_slot(size_t slot) {
  {
    std::lock_guard<std::mutex> lock(staging_.mu);
    staging_.used.erase(slot);
  }
  staging_.cv.notify_all();
}

uint64_t Server::submit(const ServeRequest& r) {
  uint64_t t = 0;
  submit_many(&r, 1, &t);
  return t;
}

void Server::submit_many(const ServeRequest* r, uint32_t count, uint64_t* tickets) {
  std::vector<Pending> ps;
  ps.reserve(count);
  for (uint32_t i = 0; i < count; ++i) ps.push_back(prepare(r[i]));  // all validated before any is queued
  for (const Pending& p : ps)
    require(p.staged_bytes() <= cfg_.host_staging_bytes, "host-memory request larger than the server's pinned staging");
  // A call without host-memory requests is queued atomically (the scheduler sees all of
  // it at once).  Host-memory requests are staged one by one, each in its own region of
  // the pinned staging, and queued in flushes of up to max_batch_scans requests or
  // kFlushBytes of transfers, so the copies of later requests overlap the transfers and
  // kernels of earlier ones.  Before waiting for staging space the call queues what it
  // has staged, so it never waits on its own requests.
  constexpr size_t kFlushBytes = 64u << 20;
  bool any_host = false;
  for (const Pending& p : ps) any_host = any_host || p.req.host;
  std::vector<uint32_t> ready;
  size_t ready_bytes = 0;
  auto flush = [&] {
    if (ready.empty()) return;
    {
      std::lock_guard<std::mutex> lock(mu_);
      const auto now = std::chrono::steady_clock::now();
      for (uint32_t i : ready) {
        ps[i].ticket = next_ticket_++;
        ps[i].submitted = now;
        tickets[i] = ps[i].ticket;
        pending_.push_back(ps[i]);
      }
      stats_.submitted += ready.size();
      inflight_ += ready.size();
    }
    cv_work_.notify_one();
    ready.clear();
    ready_bytes = 0;
  };
  for (uint32_t i = 0; i < count; ++i) {
    Pending& p = ps[i];
    const size_t b = p.staged_bytes();
    if (b > 0) {
      size_t at = 0;
      if (!try_reserve_slot(b, &at)) {
        flush();
        at = reserve_slot(b);
      }
      p.slot = at;
      for (Pending::HostBuf* h : {&p.in, &p.rss, &p.img}) {
        if (h->direct || h->dev == 0) continue;
        h->at = at;
        at += h->dev;
      }
      if (!p.in.direct) copies_->copy(staging_.base + p.in.at, p.in.user, p.in.len);
    }
    ready_bytes += p.req.host ? p.device_bytes() : 0;
    ready.push_back(i);
    if (any_host && (ready.size() >= cfg_.max_batch_scans || ready_bytes >= kFlushBytes)) flush();
  }
  flush();
}

int Server::wait(uint64_t ticket, double* latency_ms, std::string* error) {
  std::unique_lock<std::mutex> lock(mu_);
  require(ticket > 0 && ticket < next_ticket_, "unknown ticket");
  cv_done_.wait(lock, [&] { return done_.count(ticket) != 0; });
  const Done d = done_[ticket];
  done_.erase(ticket);
  if (latency_ms) *latency_ms = d.latency_ms;
  if (error) *error = d.error;
  return d.status;
}

void Server::drain() {
  std::unique_lock<std::mutex> lock(mu_);
  cv_done_.wait(lock, [&] { return inflight_ == 0; });
}

void Server::hold(bool on) {
  {
    std::lock_guard<std::mutex> lock(mu_);
    hold_ = on;
  }
  cv_work_.notify_all();
}

ServeStats Server::stats() {
  std::lock_guard<std::mutex> lock(mu_);
  ServeStats s = stats_;
  s.plan_bytes = 0;
  for (auto& kv : plans_) s.plan_bytes += kv.second.bytes;
  return s;
}

uint64_t Server::used_bytes() const {
  uint64_t b = 0;
  for (const Lane& l : lanes_) b += l.ws_bytes + l.stage_bytes;
  for (const auto& kv : plans_) b += kv.second.bytes;
  return b;
}

bool Server::evict_one() {
  // The least recently used plan that no batch in flight uses.
  for (auto it = lru_.rbegin(); it != lru_.rend(); ++it) {
    auto pe = plans_.find(*it);
    if (pe != plans_.end() && pe->second.users == 0) {
      plans_.erase(pe);  // the entry's deleter frees the plan's device tables
      lru_.erase(std::next(it).base());
      ++stats_.plan_evictions;
      return true;
    }
  }
  return false;
}

void Server::evict_until(uint64_t need_bytes) {
  while (used_bytes() + need_bytes > cfg_.hbm_budget_bytes && evict_one()) {
  }
}

Server::PlanEntry* Server::cached_plan(uint64_t key) {
  auto it = plans_.find(key);
  if (it == plans_.end()) return nullptr;
  ++stats_.plan_hits;
  lru_.remove(key);
  lru_.push_front(key);
  return &it->second;
}

Plan3* Server::plan3_for(const Pending& p, hipStream_t stream) {
  if (PlanEntry* e = cached_plan(p.plan_key)) return e->plan3.get();
  while (plans_.size() >= cfg_.plan_cache_entries && evict_one()) {
  }
  const ServeRequest& r = p.req;
  Plan3Desc d;
  const uint32_t n[3] = {r.image_depth, r.image_rows, r.image_cols};
  for (int i = 0; i < 3; ++i) {
    d.image[i] = n[i];
    d.grid[i] = 2 * n[i];
    d.origin[i] = (n[i] - 1) / 2.0;
  }
  d.width = r.kernel_width;
  d.density = r.density;
  d.samples = r.samples;
  d.item_visits = 0;  // the schedule's tuning key
  Plan3Ptr plan(plan3_create(d), &plan3_destroy);
  plan3_set_trajectory(plan.get(), r.d_traj, nullptr, stream);
  PlanEntry e;
  e.bytes = plan->device_bytes + plan->visit_capacity * (sizeof(uint32_t) + 2 * sizeof(float4)) +
            static_cast<uint64_t>(plan->item_capacity) * sizeof(uint4);
  e.plan3 = std::move(plan);
  Plan3* raw = e.plan3.get();
  plans_.emplace(p.plan_key, std::move(e));
  lru_.push_front(p.plan_key);
  ++stats_.plan_builds;
  return raw;
}

Plan2* Server::plan_for(const Pending& p, hipStream_t stream) {
  if (PlanEntry* e = cached_plan(p.plan_key)) return e->plan.get();
  while (plans_.size() >= cfg_.plan_cache_entries && evict_one()) {
  }
  const ServeRequest& r = p.req;
  Plan2Desc d;
  d.image_rows = r.image_rows;
  d.image_cols = r.image_cols;
  d.grid_rows = 2 * r.image_rows;
  d.grid_cols = 2 * r.image_cols;
  d.width = r.kernel_width;
  d.kernel = kKernelKaiserBessel;
  d.row_origin = (r.image_rows - 1) / 2.0;
  d.col_origin = (r.image_cols - 1) / 2.0;
  d.density = r.density;
  d.centred = 1;
  d.samples = r.samples;
  d.item_visits = 0;  // the schedule's tuning key
  PlanPtr plan(plan2_create(d), &plan2_destroy);
  if (r.host) {  // the plan reads its trajectory once, while it is built
    const size_t bytes = static_cast<size_t>(r.samples) * 2 * sizeof(float);
    float* d_traj = nullptr;
    KSJ_HIP(hipMalloc(&d_traj, bytes));
    try {
      KSJ_HIP(hipMemcpyAsync(d_traj, r.d_traj, bytes, hipMemcpyHostToDevice, stream));
      plan2_set_trajectory(plan.get(), d_traj, nullptr, stream);
    } catch (...) {
      (void)hipStreamSynchronize(stream);
      (void)hipFree(d_traj);
      throw;
    }
    KSJ_HIP(hipFree(d_traj));
  } else {
    plan2_set_trajectory(plan.get(), r.d_traj, nullptr, stream);
  }
  PlanEntry e;
  const uint64_t cells = plan->packed_visits ? 0 : static_cast<uint64_t>(plan->block_h) * plan->block_w;
  e.bytes = plan->device_bytes + plan->visit_capacity * (sizeof(uint32_t) + cells * sizeof(float)) +
            static_cast<uint64_t>(plan->item_capacity) * (sizeof(uint4) + sizeof(uint32_t));
  e.plan = std::move(plan);
  Plan2* raw = e.plan.get();
  plans_.emplace(p.plan_key, std::move(e));
  lru_.push_front(p.plan_key);
  ++stats_.plan_builds;
  return raw;
}

// Takes the oldest pending request and every later one that can share its launch,
// sizes the workspace against the HBM budget, and enqueues the batch on `lane`.
bool Server::launch_next(Lane& lane) {
  if (pending_.empty()) return false;
  const Pending head = pending_.front();
  const ServeRequest& r = head.req;
  std::vector<Pending> batch;
  const uint32_t limit = r.kind == kServeRadial3D ? 1u : cfg_.max_batch_scans;
  // Host-memory requests are bound by their transfers, not the kernels: a launch takes at
  // most kHostBatchBytes of them, so one lane's transfers overlap another's kernels.
  constexpr size_t kHostBatchBytes = 64u << 20;
  size_t host_bytes = 0;
  for (auto it = pending_.begin(); it != pending_.end() && batch.size() < limit;) {
    const ServeRequest& q = it->req;
    const bool same = it->plan_key == head.plan_key && q.channels == r.channels;
    if (same && q.host && !batch.empty() && host_bytes + it->device_bytes() > kHostBatchBytes) break;
    if (same) {
      host_bytes += q.host ? it->device_bytes() : 0;
      batch.push_back(*it);
      it = pending_.erase(it);
    } else {
      ++it;
    }
  }
  KSJ_HIP(hipSetDevice(cfg_.device));
  size_t need = 0;
  Plan2* plan = nullptr;
  Plan3* plan3 = nullptr;
  try {
    if (r.kind == kServeRadial3D) {
      plan3 = plan3_for(head, lane.stream);
      plans_[head.plan_key].users++;
      lane.plan_key = head.plan_key;
      need = adjoint3_workspace_bytes(*plan3, r.channels, kServeCoilsPerPass3D);
    } else if (r.kind == kServeRadial2D) {
      plan = plan_for(head, lane.stream);
      plans_[head.plan_key].users++;  // pinned against eviction from here on; released by complete()
      lane.plan_key = head.plan_key;
      need = adjoint2_workspace_bytes(*plan, r.channels, static_cast<uint32_t>(batch.size()));
    } else {
      need = ifft2_workspace_bytes(r.channels, r.image_rows, r.image_cols) +
             256 + static_cast<size_t>(r.image_rows) * r.image_cols * r.channels * sizeof(float2);
    }
    // Admission against the budget: drop idle plans first, then give up on this lane's growth.
    auto grow = [&](void*& buf, size_t& have, size_t want) {
      if (want <= have) return;
      evict_until(want - have);
      if (used_bytes() - have + want > cfg_.hbm_budget_bytes)
        throw Error(kErrExhausted, "request does not fit in the HBM budget of the server");
      if (buf) {
        KSJ_HIP(hipStreamSynchronize(lane.stream));
        KSJ_HIP(hipFree(buf));
        buf = nullptr;
        have = 0;
      }
      KSJ_HIP(hipMalloc(&buf, want));
      have = want;
      stats_.peak_workspace_bytes = std::max<uint64_t>(stats_.peak_workspace_bytes, used_bytes());
    };
    grow(lane.ws, lane.ws_bytes, need);
    // Host-memory requests: their k-space goes to the lane's device staging ahead of the
    // kernels, which then read and write that staging.
    size_t stage = 0;
    for (const Pending& p : batch) stage += p.device_bytes();
    grow(lane.stage, lane.stage_bytes, stage);
    size_t off = 0;
    for (Pending& p : batch) {
      if (!p.req.host) continue;
      char* d = static_cast<char*>(lane.stage) + off;
      const void* src = p.in.direct ? p.in.user : staging_.base + p.in.at;
      KSJ_HIP(hipMemcpyAsync(d, src, p.in.len, hipMemcpyHostToDevice, lane.stream));
      p.req.d_kspace = reinterpret_cast<const float2*>(d);
      p.req.d_rss = p.rss.dev ? reinterpret_cast<float*>(d + p.in.dev) : nullptr;
      p.req.d_images = p.img.dev ? reinterpret_cast<float2*>(d + p.in.dev + p.rss.dev) : nullptr;
      off += p.device_bytes();
    }
    if (r.kind == kServeRadial3D) {
      adjoint3_execute(*plan3, r.d_kspace, r.channels, kServeCoilsPerPass3D, r.d_images, r.d_rss, lane.ws,
                       lane.ws_bytes, lane.stream);
    } else if (r.kind == kServeRadial2D) {
      std::vector<AdjointScan> scans(batch.size());
      for (size_t i = 0; i < batch.size(); ++i)
        scans[i] = AdjointScan{batch[i].req.d_kspace, batch[i].req.d_images, batch[i].req.d_rss};
      adjoint2_execute(*plan, scans.data(), static_cast<uint32_t>(scans.size()), r.channels, kLayoutCoilLast, lane.ws,
                       lane.ws_bytes, lane.stream);
    } else {
      const size_t fft_ws = ifft2_workspace_bytes(r.channels, r.image_rows, r.image_cols);
      auto* coils = reinterpret_cast<float2*>(static_cast<char*>(lane.ws) + ((fft_ws + 255) / 256) * 256);
      for (const Pending& p : batch) {
        float2* images = p.req.d_images ? p.req.d_images : coils;
        ifft2_coil_images(p.req.d_kspace, images, r.channels, r.image_rows, r.image_cols, p.req.centred != 0,
                          kLayoutCoilLast, lane.ws, fft_ws, lane.stream);
        if (p.req.d_rss)
          rss_coil_last(images, p.req.d_rss, static_cast<uint64_t>(r.image_rows) * r.image_cols, r.channels,
                        lane.stream);
      }
    }
    for (const Pending& p : batch) {  // outputs of host-memory requests back to the host
      if (!p.req.host) continue;
      const std::pair<const Pending::HostBuf*, const void*> outs[2] = {{&p.rss, p.req.d_rss}, {&p.img, p.req.d_images}};
      for (const auto& [h, d] : outs) {
        if (h->dev == 0) continue;
        void* dst = h->direct ? h->user : staging_.base + h->at;
        KSJ_HIP(hipMemcpyAsync(dst, d, h->len, hipMemcpyDeviceToHost, lane.stream));
      }
    }
    KSJ_HIP(hipEventRecord(lane.done, lane.stream));
  } catch (const Error& e) {
    (void)hipStreamSynchronize(lane.stream);  // copies already queued end before the staging is reused
    lane.batch = std::move(batch);
    complete(lane, status_of(e.code), e.what());
    return true;
  } catch (const std::exception& e) {
    (void)hipStreamSynchronize(lane.stream);
    lane.batch = std::move(batch);
    complete(lane, 3, e.what());
    return true;
  }
  lane.batch = std::move(batch);
  lane.busy = true;
  ++stats_.batches;
  stats_.max_batch_seen = std::max<uint64_t>(stats_.max_batch_seen, lane.batch.size());
  return true;
}

void Server::complete(Lane& lane, int status, const std::string& error) {
  for (const Pending& p : lane.batch) {
    if (p.req.host) {
      if (status == 0)
        for (const Pending::HostBuf* h : {&p.rss, &p.img})
          if (h->dev && !h->direct) copies_->copy(h->user, staging_.base + h->at, h->len);
      if (p.slot != SIZE_MAX) release_slot(p.slot);
    }
    done_[p.ticket] = Done{status, ms_since(p.submitted), error};
    if (status == 0)
      ++stats_.completed;
    else
      ++stats_.failed;
    --inflight_;
  }
  if (lane.plan_key) {
    auto it = plans_.find(lane.plan_key);
    if (it != plans_.end() && it->second.users > 0) it->second.users--;
  }
  lane.batch.clear();
  lane.plan_key = 0;
  lane.busy = false;
  cv_done_.notify_all();
}

void Server::run() {
  (void)hipSetDevice(cfg_.device);
  std::unique_lock<std::mutex> lock(mu_);
  while (true) {
    bool busy = false;
    for (Lane& l : lanes_) {
      if (l.busy) {
        const hipError_t q = hipEventQuery(l.done);
        if (q == hipSuccess) {
          complete(l, 0, std::string());
        } else if (q != hipErrorNotReady) {
          complete(l, 3, std::string("device error: ") + hipGetErrorString(q));
        } else {
          busy = true;
        }
      }
    }
    bool launched = false;
    for (Lane& l : lanes_)
      if (!l.busy && !pending_.empty() && !hold_) launched = launch_next(l) || launched;
    if (stop_ && (pending_.empty() || hold_) && !busy && !launched) break;
    if (launched) continue;
    if (busy) {
      // Work in flight: poll its completion events at a fine grain.
      lock.unlock();
      std::this_thread::sleep_for(std::chrono::microseconds(20));
      lock.lock();
    } else {
      cv_work_.wait(lock, [&] { return stop_ || (!pending_.empty() && !hold_); });
    }
  }
}

}  // namespace ksj::gpu
