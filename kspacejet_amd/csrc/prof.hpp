// Optional per-stage timing with HIP events recorded on the launch stream.
// Disabled by default; when enabled every stage records a start/stop event pair
// on the stream it runs on, and reads resolve the pending pairs.
#pragma once

#include "common.hpp"

namespace ksj::gpu {

bool prof_enabled();
void prof_enable(bool on);
void prof_reset();
// Records the start/stop of `stage` on `stream` (no-ops when disabled).
void prof_begin(const char* stage, hipStream_t stream);
void prof_end(const char* stage, hipStream_t stream);
// Synchronises pending events and returns the accumulated milliseconds / count.
void prof_read(const char* stage, double* total_ms, uint64_t* count);

struct ProfScope {
  const char* stage;
  hipStream_t stream;
  ProfScope(const char* s, hipStream_t st) : stage(s), stream(st) { prof_begin(stage, stream); }
  ~ProfScope() { prof_end(stage, stream); }
};

}  // namespace ksj::gpu
