// Deterministic device primitives used by sample binning: an exclusive scan of
// uint32 counts and a stable LSD radix sort of (key, value) pairs.  No float
// atomics; integer LDS atomics only build per-tile histograms, whose totals do
// not depend on thread order, so every output is bit-reproducible.
#pragma once

#include "common.hpp"

namespace ksj::gpu {

// Scratch bytes exclusive_scan_u32 needs for n elements.
size_t scan_scratch_bytes(uint64_t n);
// out[i] = sum_{j < i} in[i]; out may alias in.  Returns nothing; all work is on `stream`.
void exclusive_scan_u32(const uint32_t* in, uint32_t* out, uint64_t n, void* scratch, hipStream_t stream);

// Scratch bytes radix_sort_pairs needs for n elements.
size_t radix_sort_scratch_bytes(uint64_t n);
// Stable sort by key (ties keep input order).  vals_in == nullptr means values are
// the input positions 0..n-1.  key_bits bounds the keys (< 2^key_bits).
// keys_out/vals_out receive the result; the inputs are left untouched.
void radix_sort_pairs(const uint32_t* keys_in, const uint32_t* vals_in, uint32_t* keys_out, uint32_t* vals_out,
                      uint64_t n, int key_bits, void* scratch, hipStream_t stream);

}  // namespace ksj::gpu
