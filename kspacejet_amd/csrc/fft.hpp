// LDS Stockham FFT for gfx950: batched 1-D transforms along either the contiguous
// axis (rows) or a strided axis (columns) of a stack of 2-D planes, with fused
// load rotation (ifftshift), pruned/rotated store (crop, fftshift), scaling and
// separable output weights (deapodization).
#pragma once

#include "common.hpp"

namespace ksj::gpu {

// Addressing of one batched 1-D transform: line (plane p, index q), element e is at
//   base + p * plane_stride + q * line_stride + e * elem_stride   (in complex elements).
struct FftLines {
  uint32_t planes = 1;
  uint32_t lines = 1;           // lines per plane
  uint64_t plane_stride = 0;
  uint64_t line_stride = 0;
  uint64_t elem_stride = 1;
};

struct FftArgs {
  const float2* in = nullptr;
  float2* out = nullptr;
  FftLines src;
  FftLines dst;                 // planes/lines must equal src's
  uint32_t n = 0;               // transform length
  uint32_t n_out = 0;           // stored outputs (<= n); output e' reads bin (e' + out_offset) mod n
  uint32_t in_offset = 0;       // line[e] = in[(e + in_offset) mod n]
  uint32_t out_offset = 0;
  float scale = 1.0f;
  const float* out_weight = nullptr;   // per stored element e' (length n_out) or null
  const float* line_weight = nullptr;  // per line index q or null
  int inverse = 1;              // +1: sum x e^{+2 pi i jk/n}; 0: forward (e^{-...})
  // Thread mapping of the global loads / stores: consecutive threads take
  // consecutive elements of a line (0) or the same element of consecutive lines
  // (1).  Pick the one whose addresses are contiguous in memory.
  int load_lines_fast = 0;
  int store_lines_fast = 0;
  // line_weight index: line % lines (0) or the line's plane (1), divided by
  // line_weight_div (lines that share one weight, e.g. the coils of one column).
  int line_weight_by_plane = 0;
  uint32_t line_weight_div = 1;
  // Per source plane: elements [x, y) of each of its lines are zero and are not read
  // (null: every element is read).  With zero_span_period > 0 the span of plane p is
  // zero_span[p % zero_span_period].
  const uint2* zero_span = nullptr;
  uint32_t zero_span_period = 0;
  // Lines whose outputs are not stored (all zero; the next pass reads them as zero): plane
  // p is skipped when p % skip_div lies in skip_span[p / skip_div] (null: every line stored).
  const uint2* skip_span = nullptr;
  uint32_t skip_div = 1;
  // Set by the launch (callers leave 0): ceil(2^32 / src.lines) when every line index l
  // satisfies l * src.lines < 2^32, so l / src.lines = mulhi(l, lines_magic); 0: divide.
  uint32_t lines_magic = 0;
};

// Outputs of a column pass whose lines are (plane = image column c, line = stacked coil
// k), k in [0, groups * channels): group g = k / channels writes coil-last images
// coil_last[g] [n_out][cols][channels] and/or its RSS image rss[g] [n_out][cols]
// (sqrt of the sum of |x|^2 in coil order); null pointers skip.
struct ColumnOutputs {
  float2* coil_last[64] = {};
  float* rss[64] = {};
  uint32_t channels = 1;
  uint32_t cols = 0;
  // Coil stride of the coil-last images (0: channels), for images holding more coils
  // than one group (coil_last[g] then points at the group's first coil).
  uint32_t img_stride = 0;
  // Sum of squares carried across launches (one group per block position): the RSS sum
  // starts from sumsq[e cols + c] unless `first`, and is stored back there unless
  // `last`, when rss gets its square root (null: every launch is first and last).
  float* sumsq = nullptr;
  int first = 1, last = 1;
};

// Returns true when an LDS Stockham kernel is compiled for length n; other
// lengths run through a direct DFT kernel.
bool fft_fast_length(uint32_t n);

// Device table of exp(+2 pi i m / n), m in [0, n), generated in FP64 and rounded
// to float.  Cached per length for the life of the process (per device).
const float2* fft_twiddles(uint32_t n, hipStream_t stream);

void fft_launch(const FftArgs& args, hipStream_t stream, int lines_per_block = 0);

// True when fft_launch_columns can take this pass: an LDS length, both mappings across
// lines, and coil groups that fit whole in the launch's blocks.
bool fft_columns_supported(const FftArgs& args, uint32_t channels);
// The pass of `args` (dst unused) with the ColumnOutputs epilogue in place of its stores.
void fft_launch_columns(const FftArgs& args, const ColumnOutputs& out, hipStream_t stream);

// Transforms of up to 64 items of an uneven batch (same direction and thread
// mapping): one launch per LDS length, each item's blocks running that length's
// kernel on its own arguments (direct-DFT lengths are launched per item).  Tables
// go to d_tables, fft_fused_table_bytes(count) long.
size_t fft_fused_table_bytes(uint32_t count);
void fft_launch_fused(const FftArgs* args, uint32_t count, void* d_tables, hipStream_t stream);

}  // namespace ksj::gpu
