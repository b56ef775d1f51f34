#include "prof.hpp"

#include <atomic>
#include <map>
#include <mutex>
#include <string>
#include <vector>

namespace ksj::gpu {
namespace {

struct Pending {
  hipEvent_t start, stop;
};

struct Stage {
  std::vector<Pending> pending;
  hipEvent_t open_start = nullptr;
  double total_ms = 0.0;
  uint64_t count = 0;
};

std::atomic<bool> g_enabled{false};
std::mutex g_mu;
std::map<std::string, Stage> g_stages;
std::vector<hipEvent_t> g_free;

hipEvent_t take_event() {
  if (!g_free.empty()) {
    hipEvent_t e = g_free.back();
    g_free.pop_back();
    return e;
  }
  hipEvent_t e;
  KSJ_HIP(hipEventCreate(&e));
  return e;
}

void resolve(Stage& st) {
  for (auto& p : st.pending) {
    KSJ_HIP(hipEventSynchronize(p.stop));
    float ms = 0.f;
    KSJ_HIP(hipEventElapsedTime(&ms, p.start, p.stop));
    st.total_ms += ms;
    st.count += 1;
    g_free.push_back(p.start);
    g_free.push_back(p.stop);
  }
  st.pending.clear();
}

}  // namespace

bool prof_enabled() { return g_enabled.load(std::memory_order_relaxed); }

void prof_enable(bool on) { g_enabled.store(on); }

void prof_reset() {
  std::lock_guard<std::mutex> lock(g_mu);
  for (auto& [name, st] : g_stages) {
    resolve(st);
    st.total_ms = 0.0;
    st.count = 0;
  }
}

void prof_begin(const char* stage, hipStream_t stream) {
  if (!prof_enabled()) return;
  std::lock_guard<std::mutex> lock(g_mu);
  Stage& st = g_stages[stage];
  st.open_start = take_event();
  KSJ_HIP(hipEventRecord(st.open_start, stream));
}

void prof_end(const char* stage, hipStream_t stream) {
  if (!prof_enabled()) return;
  std::lock_guard<std::mutex> lock(g_mu);
  Stage& st = g_stages[stage];
  if (!st.open_start) return;
  hipEvent_t stop = take_event();
  KSJ_HIP(hipEventRecord(stop, stream));
  st.pending.push_back({st.open_start, stop});
  st.open_start = nullptr;
}

void prof_read(const char* stage, double* total_ms, uint64_t* count) {
  std::lock_guard<std::mutex> lock(g_mu);
  auto it = g_stages.find(stage);
  if (it == g_stages.end()) {
    *total_ms = 0.0;
    *count = 0;
    return;
  }
  resolve(it->second);
  *total_ms = it->second.total_ms;
  *count = it->second.count;
}

}  // namespace ksj::gpu
