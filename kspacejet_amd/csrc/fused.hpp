// Fused launches over the items of an uneven batch: one kernel launch runs a stage
// for every item of a shape class.  Per-item parameters live in a device table
// {count, first block of each item, args}; the launch's blocks are split among the
// items by the prefix of their block counts, and each block runs the stage's
// per-block body with its item's arguments.
//
// Tables are written into caller memory by small kernels that carry the bytes as
// kernel arguments (upload_table), so nothing is read from host memory when a
// captured graph replays, and nothing is allocated.
#pragma once

#include "common.hpp"

#include <cstddef>

namespace ksj::gpu {

constexpr uint32_t kMaxFusedItems = 64;

template <typename Args>
struct FusedTable {
  uint32_t count;
  uint32_t pad;
  uint64_t first[kMaxFusedItems + 1];  // first block of each item; first[count] = total blocks
  Args args[kMaxFusedItems];
};

// The table is read through the constant address space: its loads are uniform per
// block, so they become scalar loads into SGPRs.
using ConstWord = const __attribute__((address_space(4))) uint32_t*;
using ConstDword = const __attribute__((address_space(4))) uint64_t*;

// Block -> item (the last item whose first block is <= block); uniform per block.
template <typename Args>
__device__ inline uint32_t fused_item(const FusedTable<Args>* t, uint64_t block) {
  const uint32_t count = *(ConstWord)(&t->count);
  const ConstDword first = (ConstDword)(t->first);
  uint32_t lo = 0, hi = count;  // first[lo] <= block < first[hi]
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (first[mid] <= block)
      lo = mid;
    else
      hi = mid;
  }
  return lo;
}

template <typename Args>
__device__ inline uint64_t fused_first(const FusedTable<Args>* t, uint32_t item) {
  return ((ConstDword)(t->first))[item];
}

template <typename Args>
__device__ inline Args fused_args(const FusedTable<Args>* t, uint32_t item) {
  static_assert(sizeof(Args) % 4 == 0, "table arguments are read as 32-bit words");
  const ConstWord src = (ConstWord)(&t->args[item]);
  uint32_t w[sizeof(Args) / 4];
#pragma unroll
  for (uint32_t i = 0; i < sizeof(Args) / 4; ++i) w[i] = src[i];
  Args a;
  __builtin_memcpy(&a, w, sizeof(Args));
  return a;
}

// Builds a table on the host: items' arguments and block counts.
template <typename Args>
struct FusedTableBuilder {
  FusedTable<Args> t{};
  void add(const Args& a, uint64_t blocks) {
    require(t.count < kMaxFusedItems, "too many items for one fused launch");
    t.args[t.count] = a;
    t.first[t.count + 1] = t.first[t.count] + blocks;
    ++t.count;
  }
  uint64_t blocks() const { return t.first[t.count]; }
  // Uploads the header and the used arguments into `dst` (fused_table_bytes<Args>() long).
  void upload(void* dst, hipStream_t stream) const {
    const size_t used = offsetof(FusedTable<Args>, args) + t.count * sizeof(Args);
    upload_table(dst, &t, (used + 3) & ~static_cast<size_t>(3), stream);
  }
};

// Copies `bytes` from host memory into device memory `dst` on `stream` through the
// arguments of one or more small kernels (no host pointer is kept by the launch).
void upload_table(void* dst, const void* src, size_t bytes, hipStream_t stream);

// Device bytes of a FusedTable<Args>, rounded to 256.
template <typename Args>
constexpr size_t fused_table_bytes() {
  return (sizeof(FusedTable<Args>) + 255) & ~static_cast<size_t>(255);
}

}  // namespace ksj::gpu
