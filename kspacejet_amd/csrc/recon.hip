// Pipeline orchestration (host side; every step is enqueued on one stream).
#include "recon.hpp"

#include "combine.hpp"
#include "fft.hpp"
#include "fused.hpp"
#include "grid.hpp"
#include "prof.hpp"

#include <algorithm>
#include <map>
#include <mutex>
#include <tuple>
#include <utility>
#include <vector>

namespace ksj::gpu {
namespace {

constexpr size_t align256(size_t b) { return (b + 255) & ~size_t(255); }

FftLines rows_of(uint32_t planes, uint32_t rows, uint32_t row_len, uint32_t pitch) {
  FftLines L;
  L.planes = planes;
  L.lines = rows;
  L.plane_stride = static_cast<uint64_t>(rows) * pitch;
  L.line_stride = pitch;
  L.elem_stride = 1;
  (void)row_len;
  return L;
}

FftLines cols_of(uint32_t planes, uint32_t rows, uint32_t cols) {
  FftLines L;
  L.planes = planes;
  L.lines = cols;
  L.plane_stride = static_cast<uint64_t>(rows) * cols;
  L.line_stride = 1;
  L.elem_stride = cols;
  return L;
}

}  // namespace

size_t ifft2_workspace_bytes(uint32_t channels, uint32_t rows, uint32_t cols) {
  const size_t plane = static_cast<size_t>(rows) * cols * sizeof(float2) * channels;
  return 2 * align256(plane);
}

void ifft2_coil_images(const float2* d_kspace, float2* d_images, uint32_t channels, uint32_t rows, uint32_t cols,
                       bool centred, int layout, void* d_ws, size_t ws_bytes, hipStream_t stream) {
  require(channels >= 1 && rows >= 1 && cols >= 1, "ifft2 needs a non-empty frame");
  require(layout == kLayoutCoilLast || layout == kLayoutCoilMajor, "unknown image layout");
  require(ws_bytes >= ifft2_workspace_bytes(channels, rows, cols), "ifft2 workspace too small");
  const size_t plane_bytes = align256(static_cast<size_t>(rows) * cols * sizeof(float2) * channels);
  auto* tmp = static_cast<float2*>(d_ws);
  auto* img = reinterpret_cast<float2*>(static_cast<char*>(d_ws) + plane_bytes);
  FftArgs a;
  a.in = d_kspace;
  a.out = tmp;
  a.src = a.dst = rows_of(channels, rows, cols, cols);
  a.n = a.n_out = cols;
  a.in_offset = centred ? cols / 2 : 0;
  a.out_offset = centred ? (cols - cols / 2) % cols : 0;
  a.scale = 1.0f / static_cast<float>(cols);
  a.inverse = 1;
  fft_launch(a, stream);
  FftArgs b;
  b.in = tmp;
  b.out = layout == kLayoutCoilMajor ? d_images : img;
  b.src = b.dst = cols_of(channels, rows, cols);
  b.n = b.n_out = rows;
  b.in_offset = centred ? rows / 2 : 0;
  b.out_offset = centred ? (rows - rows / 2) % rows : 0;
  b.scale = 1.0f / static_cast<float>(rows);
  b.inverse = 1;
  b.load_lines_fast = b.store_lines_fast = 1;
  fft_launch(b, stream);
  if (layout == kLayoutCoilLast)
    coil_major_to_last(img, d_images, nullptr, static_cast<uint64_t>(rows) * cols, channels, stream);
}

uint32_t adjoint2_scans_per_pass(uint32_t channels) {
  return channels >= kMaxStackedCoils ? 1u : kMaxStackedCoils / channels;
}

namespace {

// Streams over which the passes of one batched call overlap (pass p runs on lane
// p % lanes), kept per (device, caller stream) and created at first use.  They are
// ordinary default-priority streams, joined to the caller's stream with events.
struct Lanes {
  hipStream_t stream[8] = {};
  hipEvent_t join[8] = {};
  hipEvent_t fork = nullptr;
};

// kind 0: pass lanes of one batched call; kind 1: item lanes of a multi-plan call
// (a multi-plan item runs a batched call of one scan on its lane, so the two pools
// must differ).
Lanes& lane_pool(hipStream_t caller, int kind) {
  static std::mutex mu;
  static std::map<std::tuple<int, hipStream_t, int>, Lanes> pool;
  int dev = 0;
  KSJ_HIP(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lock(mu);
  Lanes& l = pool[{dev, caller, kind}];
  if (l.fork == nullptr) {
    for (int i = 0; i < 8; ++i) {
      KSJ_HIP(hipStreamCreateWithFlags(&l.stream[i], hipStreamNonBlocking));
      KSJ_HIP(hipEventCreateWithFlags(&l.join[i], hipEventDisableTiming));
    }
    KSJ_HIP(hipEventCreateWithFlags(&l.fork, hipEventDisableTiming));
  }
  return l;
}

Lanes& lanes_for(hipStream_t caller) { return lane_pool(caller, 0); }
Lanes& multi_lanes_for(hipStream_t caller) { return lane_pool(caller, 1); }

// Row pass: lines (grid row y, coil c) of the coil-fastest grid [G_r][G_c][Cp],
// loaded coil-fastest, stored as coil planes [C][G_r][N_c] (only the N_c kept columns).
FftArgs row_pass_args(const Plan2& plan, uint32_t total, size_t cp, const float2* grid, float2* rowout) {
  const auto& d = plan.d;
  FftArgs a;
  a.in = grid;
  a.out = rowout;
  a.src.planes = d.grid_rows;
  a.src.lines = total;
  a.src.plane_stride = static_cast<uint64_t>(d.grid_cols) * cp;
  a.src.line_stride = 1;
  a.src.elem_stride = cp;
  a.dst.planes = d.grid_rows;
  a.dst.lines = total;
  a.dst.plane_stride = d.image_cols;
  a.dst.line_stride = static_cast<uint64_t>(d.grid_rows) * d.image_cols;
  a.dst.elem_stride = 1;
  a.n = d.grid_cols;
  a.n_out = d.image_cols;
  a.out_offset = d.centred ? (d.grid_cols - d.image_cols / 2) % d.grid_cols : 0;
  a.inverse = 1;
  a.load_lines_fast = 1;
  a.store_lines_fast = 0;
  // Columns the gather leaves unwritten (no sample reaches them) are read as zero.
  a.zero_span = tune_get(kTuneGridPrune) == 1 ? plan.zero_cols : nullptr;
  return a;
}

// Column pass with deapodization and crop -> coil planes [C][N_r][N_c].
FftArgs col_pass_args(const Plan2& plan, uint32_t total, const float2* rowout, float2* out) {
  const auto& d = plan.d;
  FftArgs b;
  b.in = rowout;
  b.out = out;
  b.src = cols_of(total, d.grid_rows, d.image_cols);
  b.dst = cols_of(total, d.image_rows, d.image_cols);
  b.n = d.grid_rows;
  b.n_out = d.image_rows;
  b.out_offset = d.centred ? (d.grid_rows - d.image_rows / 2) % d.grid_rows : 0;
  b.out_weight = plan.deapod_rows;
  b.line_weight = plan.deapod_cols;
  b.inverse = 1;
  b.load_lines_fast = b.store_lines_fast = 1;
  return b;
}

// Row pass of the column-epilogue pipeline: the same lines, stored coil-fastest (only
// the N_c kept columns), so both mappings run across coils.  Grid-row major
// [G_r][N_c][Cp] (row_order 0) or column major [N_c][G_r][Cp] (row_order 1, where
// each column-pass block reads one contiguous run per column).
FftArgs row_pass_coil_last_args(const Plan2& plan, uint32_t total, size_t cp, const float2* grid, float2* rowout) {
  FftArgs a = row_pass_args(plan, total, cp, grid, rowout);
  const bool by_column = tune_get(kTuneRowOrder) == 1;
  a.dst.planes = plan.d.grid_rows;
  a.dst.lines = total;
  a.dst.plane_stride = by_column ? cp : static_cast<uint64_t>(plan.d.image_cols) * cp;
  a.dst.line_stride = 1;
  a.dst.elem_stride = by_column ? static_cast<uint64_t>(plan.d.grid_rows) * cp : cp;
  a.store_lines_fast = 1;
  return a;
}

// Column pass over the coil-fastest row output: lines (plane = column c, coil k) along
// the grid rows, deapodized (rows by output element, columns by plane) and cropped; its
// epilogue writes each scan's coil-last images and RSS.
FftArgs col_pass_coil_last_args(const Plan2& plan, uint32_t total, size_t cp, const float2* rowout) {
  const auto& d = plan.d;
  FftArgs b;
  b.in = rowout;
  const bool by_column = tune_get(kTuneRowOrder) == 1;
  b.src.planes = d.image_cols;
  b.src.lines = total;
  b.src.plane_stride = by_column ? static_cast<uint64_t>(d.grid_rows) * cp : cp;
  b.src.line_stride = 1;
  b.src.elem_stride = by_column ? cp : static_cast<uint64_t>(d.image_cols) * cp;
  b.dst = b.src;
  b.n = d.grid_rows;
  b.n_out = d.image_rows;
  b.out_offset = d.centred ? (d.grid_rows - d.image_rows / 2) % d.grid_rows : 0;
  b.out_weight = plan.deapod_rows;
  b.line_weight = plan.deapod_cols;
  b.line_weight_by_plane = 1;
  b.inverse = 1;
  b.load_lines_fast = b.store_lines_fast = 1;
  return b;
}

uint32_t pass_count(uint32_t channels, uint32_t scans) {
  const uint32_t per = adjoint2_scans_per_pass(channels);
  return (std::max(scans, 1u) + per - 1) / per;
}

uint32_t lane_count(uint32_t channels, uint32_t scans) {
  return static_cast<uint32_t>(std::min<int64_t>(tune_get(kTunePassLanes), pass_count(channels, scans)));
}

}  // namespace

size_t adjoint2_workspace_bytes(const Plan2& plan, uint32_t channels, uint32_t scans) {
  return adjoint2_pass_workspace_bytes(plan, channels, scans) * lane_count(channels, scans);
}

size_t adjoint2_pass_workspace_bytes(const Plan2& plan, uint32_t channels, uint32_t scans) {
  const auto& d = plan.d;
  const uint32_t per = std::min(std::max(scans, 1u), adjoint2_scans_per_pass(channels));
  const uint32_t total = per * channels;
  const size_t cp = padded_coils(total);
  const size_t sorted = static_cast<size_t>(d.samples) * cp * sizeof(float2);
  const size_t grid = cp * d.grid_rows * d.grid_cols * sizeof(float2);
  const size_t rowout = cp * d.grid_rows * d.image_cols * sizeof(float2);
  const size_t img = static_cast<size_t>(total) * d.image_rows * d.image_cols * sizeof(float2);
  return align256(sorted) + align256(grid) + align256(rowout) + align256(img) +
         align256(grid_partial_bytes(plan, total));
}

void adjoint2_execute(const Plan2& plan, const AdjointScan* scans, uint32_t count, uint32_t channels, int layout,
                      void* d_ws, size_t ws_bytes, hipStream_t stream) {
  const auto& d = plan.d;
  require(plan.has_trajectory, "plan has no trajectory");
  require(channels >= 1 && channels <= kMaxStackedCoils, "adjoint supports 1..64 channels");
  require(layout == kLayoutCoilLast || layout == kLayoutCoilMajor, "unknown image layout");
  require(count >= 1, "adjoint needs at least one scan");
  require(ws_bytes >= adjoint2_workspace_bytes(plan, channels, count), "adjoint workspace too small");
  const uint32_t per = adjoint2_scans_per_pass(channels);
  const uint64_t pixels = static_cast<uint64_t>(d.image_rows) * d.image_cols;
  const uint32_t nlanes = lane_count(channels, count);
  const size_t lane_ws = adjoint2_pass_workspace_bytes(plan, channels, count);
  Lanes* lanes = nullptr;
  const hipStream_t caller = stream;
  if (nlanes > 1) {
    lanes = &lanes_for(caller);
    KSJ_HIP(hipEventRecord(lanes->fork, caller));
    for (uint32_t i = 0; i < nlanes; ++i) KSJ_HIP(hipStreamWaitEvent(lanes->stream[i], lanes->fork, 0));
  }
  for (uint32_t first = 0, pass = 0; first < count; first += per, ++pass) {
    const uint32_t n = std::min(per, count - first);
    const uint32_t total = n * channels;
    const size_t cp = padded_coils(total);
    if (lanes) stream = lanes->stream[pass % nlanes];
    char* p = static_cast<char*>(d_ws) + (pass % nlanes) * lane_ws;
    auto* sorted = reinterpret_cast<float2*>(p);
    p += align256(static_cast<size_t>(d.samples) * cp * sizeof(float2));
    auto* grid = reinterpret_cast<float2*>(p);
    p += align256(cp * d.grid_rows * d.grid_cols * sizeof(float2));
    auto* rowout = reinterpret_cast<float2*>(p);
    p += align256(cp * d.grid_rows * d.image_cols * sizeof(float2));
    auto* img = reinterpret_cast<float2*>(p);
    p += align256(static_cast<size_t>(total) * d.image_rows * d.image_cols * sizeof(float2));
    auto* partial = reinterpret_cast<float2*>(p);

    CoilPlanes planes;
    planes.coils = total;
    for (uint32_t i = 0; i < n; ++i) {
      require(scans[first + i].kspace != nullptr, "null k-space");
      for (uint32_t c = 0; c < channels; ++c)
        planes.plane[i * channels + c] = scans[first + i].kspace + static_cast<uint64_t>(c) * d.samples;
    }
    {
      ProfScope ps("grid_prep", stream);
      grid_prep(plan, planes, sorted, stream);
    }
    {
      ProfScope ps("grid_gather", stream);
      grid_gather(plan, sorted, total, grid, partial, stream);
    }
    if (plan.multi_count > 0) {
      ProfScope ps("grid_reduce", stream);
      grid_reduce(plan, total, grid, partial, stream);
    }

    // Coil-last outputs: rows stored coil-fastest, then one column pass that writes
    // every scan's coil-last images and RSS itself (no coil planes, no split pass).
    const FftArgs b2 = col_pass_coil_last_args(plan, total, cp, rowout);
    if (layout == kLayoutCoilLast && fft_columns_supported(b2, channels)) {
      {
        ProfScope ps("fft_rows", stream);
        fft_launch(row_pass_coil_last_args(plan, total, cp, grid, rowout), stream);
      }
      ColumnOutputs co;
      co.channels = channels;
      co.cols = d.image_cols;
      for (uint32_t i = 0; i < n; ++i) {
        co.coil_last[i] = scans[first + i].images;
        co.rss[i] = scans[first + i].rss;
      }
      ProfScope ps("fft_cols_deapod", stream);
      fft_launch_columns(b2, co, stream);
      continue;
    }

    const FftArgs a = row_pass_args(plan, total, cp, grid, rowout);
    {
      ProfScope ps("fft_rows", stream);
      fft_launch(a, stream);
    }

    const bool direct = (count == 1 && layout == kLayoutCoilMajor && scans[0].images != nullptr);
    const FftArgs b = col_pass_args(plan, total, rowout, direct ? scans[0].images : img);
    {
      ProfScope ps("fft_cols_deapod", stream);
      fft_launch(b, stream);
    }

    ProfScope ps("coil_combine", stream);
    CoilGroupOutputs out;
    out.groups = n;
    out.coils = channels;
    bool any = false;
    for (uint32_t i = 0; i < n; ++i) {
      const AdjointScan& sc = scans[first + i];
      out.coil_last[i] = layout == kLayoutCoilLast ? sc.images : nullptr;
      out.rss[i] = sc.rss;
      any = any || out.coil_last[i] || out.rss[i];
      if (layout == kLayoutCoilMajor && sc.images && !direct)
        KSJ_HIP(hipMemcpyAsync(sc.images, img + static_cast<uint64_t>(i) * channels * pixels,
                               static_cast<size_t>(channels) * pixels * sizeof(float2), hipMemcpyDeviceToDevice,
                               stream));
    }
    if (any) coil_major_split(b.out, out, pixels, stream);
  }
  if (lanes) {
    for (uint32_t i = 0; i < nlanes; ++i) {
      KSJ_HIP(hipEventRecord(lanes->join[i], lanes->stream[i]));
      KSJ_HIP(hipStreamWaitEvent(caller, lanes->join[i], 0));
    }
  }
}

namespace {

struct MultiLayout {
  uint32_t lanes = 1;
  std::vector<uint32_t> order;   // items, largest first
  std::vector<size_t> lane_ws;   // workspace bytes of each lane (its largest item)
};

MultiLayout multi_layout(const MultiItem* items, uint32_t count) {
  MultiLayout m;
  m.lanes = static_cast<uint32_t>(std::min<int64_t>(tune_get(kTuneMultiLanes), std::max(count, 1u)));
  m.order.resize(count);
  std::vector<size_t> ws(count);
  for (uint32_t i = 0; i < count; ++i) {
    require(items[i].plan != nullptr && items[i].plan->has_trajectory, "multi item without a ready plan");
    ws[i] = adjoint2_workspace_bytes(*items[i].plan, items[i].channels, 1);
    m.order[i] = i;
  }
  std::stable_sort(m.order.begin(), m.order.end(), [&](uint32_t a, uint32_t b) { return ws[a] > ws[b]; });
  m.lane_ws.assign(m.lanes, 0);
  for (uint32_t j = 0; j < count; ++j) m.lane_ws[j % m.lanes] = std::max(m.lane_ws[j % m.lanes], ws[m.order[j]]);
  return m;
}

}  // namespace

namespace {

// The fused form runs every stage once for all items (one launch per stage and
// shape class, argument tables in the workspace) when every item fits one pass of
// the 2x2-block schedule; otherwise items go round-robin over lanes.
// The fused form takes 2x2-block and matrix plans (all matrix plans of one block
// width) of at most 64 channels.
bool fused_shapes_ok(const MultiItem* items, uint32_t count) {
  if (count > kMaxFusedItems) return false;
  uint32_t matrix_w = 0;
  for (uint32_t i = 0; i < count; ++i) {
    const Plan2& p = *items[i].plan;
    if (p.scheme == kSchemeTiles || items[i].channels > kMaxStackedCoils) return false;
    if (p.scheme == kSchemeMatrix) {
      if (matrix_w != 0 && matrix_w != p.block_w) return false;
      matrix_w = p.block_w;
    }
  }
  return true;
}

bool fused_eligible(const MultiItem* items, uint32_t count) {
  return tune_get(kTuneMultiFused) != 0 && fused_shapes_ok(items, count);
}

size_t fused_tables_bytes(uint32_t count) {
  return align256(std::max(grid_fused_table_bytes(), grid_chain_table_bytes())) +
         2 * align256(fft_fused_table_bytes(count)) + align256(coil_split_fused_table_bytes());
}

size_t fused_total_bytes(const MultiItem* items, uint32_t count) {
  size_t total = fused_tables_bytes(count);
  for (uint32_t i = 0; i < count; ++i)
    total += align256(adjoint2_pass_workspace_bytes(*items[i].plan, items[i].channels, 1));
  return total;
}

// The fused form over lanes: items dealt round-robin (largest first) onto the
// multi_lanes streams, each lane running the fused stages over its own items.
std::vector<std::vector<MultiItem>> fused_lanes(const MultiItem* items, uint32_t count) {
  const MultiLayout m = multi_layout(items, count);
  std::vector<std::vector<MultiItem>> lanes(m.lanes);
  for (uint32_t j = 0; j < count; ++j) lanes[j % m.lanes].push_back(items[m.order[j]]);
  return lanes;
}

size_t fused_lanes_bytes(const MultiItem* items, uint32_t count) {
  size_t total = 0;
  for (const auto& l : fused_lanes(items, count)) total += fused_total_bytes(l.data(), static_cast<uint32_t>(l.size()));
  return total;
}

size_t lanes_total_bytes(const MultiItem* items, uint32_t count) {
  const MultiLayout m = multi_layout(items, count);
  size_t total = 0;
  for (size_t b : m.lane_ws) total += align256(b);
  return total;
}

void execute_fused(const MultiItem* items, uint32_t count, int layout, void* d_ws, hipStream_t stream) {
  char* p = static_cast<char*>(d_ws);
  char* grid_tab = p;
  char* row_tab = grid_tab + align256(std::max(grid_fused_table_bytes(), grid_chain_table_bytes()));
  char* col_tab = row_tab + align256(fft_fused_table_bytes(count));
  char* split_tab = col_tab + align256(fft_fused_table_bytes(count));
  p += fused_tables_bytes(count);
  std::vector<GridFusedItem> gi(count);
  std::vector<FftArgs> rows(count), cols(count);
  std::vector<const float2*> split_in(count);
  std::vector<float2*> split_last(count);
  std::vector<float*> split_rss(count);
  std::vector<uint64_t> pixels(count);
  std::vector<uint32_t> coils(count);
  for (uint32_t i = 0; i < count; ++i) {
    const MultiItem& it = items[i];
    const Plan2Desc& d = it.plan->d;
    require(it.scan.kspace != nullptr, "null k-space");
    const uint32_t c = it.channels;
    const size_t cp = padded_coils(c);
    char* q = p;
    p += align256(adjoint2_pass_workspace_bytes(*it.plan, c, 1));
    auto* sorted = reinterpret_cast<float2*>(q);
    q += align256(static_cast<size_t>(d.samples) * cp * sizeof(float2));
    auto* grid = reinterpret_cast<float2*>(q);
    q += align256(cp * d.grid_rows * d.grid_cols * sizeof(float2));
    auto* rowout = reinterpret_cast<float2*>(q);
    q += align256(cp * d.grid_rows * d.image_cols * sizeof(float2));
    auto* img = reinterpret_cast<float2*>(q);
    q += align256(static_cast<size_t>(c) * d.image_rows * d.image_cols * sizeof(float2));
    gi[i] = GridFusedItem{it.plan, it.scan.kspace, c, sorted, grid, reinterpret_cast<float2*>(q)};
    float2* planes_out = (layout == kLayoutCoilMajor && it.scan.images) ? it.scan.images : img;
    rows[i] = row_pass_args(*it.plan, c, cp, grid, rowout);
    cols[i] = col_pass_args(*it.plan, c, rowout, planes_out);
    split_in[i] = planes_out;
    split_last[i] = layout == kLayoutCoilLast ? it.scan.images : nullptr;
    split_rss[i] = it.scan.rss;
    pixels[i] = static_cast<uint64_t>(d.image_rows) * d.image_cols;
    coils[i] = c;
  }
  if (tune_get(kTuneMultiFused) == 2 && grid_chain_supported(gi.data(), count, stream)) {
    ProfScope ps("multi_grid_chain", stream);
    grid_chain(gi.data(), count, grid_tab, stream);
  } else {
    ProfScope ps("multi_grid", stream);
    grid_fused(gi.data(), count, grid_tab, stream);
  }
  {
    ProfScope ps("multi_fft_rows", stream);
    fft_launch_fused(rows.data(), count, row_tab, stream);
  }
  {
    ProfScope ps("multi_fft_cols_deapod", stream);
    fft_launch_fused(cols.data(), count, col_tab, stream);
  }
  ProfScope ps("multi_coil_combine", stream);
  coil_split_fused(split_in.data(), split_last.data(), split_rss.data(), pixels.data(), coils.data(), count,
                   split_tab, stream);
}

}  // namespace

size_t adjoint2_multi_workspace_bytes(const MultiItem* items, uint32_t count) {
  // Enough for either form, so a tuning change between query and call stays valid.
  size_t total = lanes_total_bytes(items, count);
  if (fused_shapes_ok(items, count)) total = std::max(total, fused_lanes_bytes(items, count));
  return total;
}

void adjoint2_execute_multi(const MultiItem* items, uint32_t count, int layout, void* d_ws, size_t ws_bytes,
                            hipStream_t stream) {
  require(count >= 1, "multi adjoint needs at least one item");
  require(layout == kLayoutCoilLast || layout == kLayoutCoilMajor, "unknown image layout");
  const MultiLayout m = multi_layout(items, count);
  if (fused_eligible(items, count)) {
    require(ws_bytes >= fused_lanes_bytes(items, count), "multi adjoint workspace too small");
    const auto lanes = fused_lanes(items, count);
    if (lanes.size() == 1) return execute_fused(items, count, layout, d_ws, stream);
    Lanes& ls = multi_lanes_for(stream);
    KSJ_HIP(hipEventRecord(ls.fork, stream));
    char* p = static_cast<char*>(d_ws);
    for (uint32_t l = 0; l < lanes.size(); ++l) {
      const uint32_t n = static_cast<uint32_t>(lanes[l].size());
      KSJ_HIP(hipStreamWaitEvent(ls.stream[l], ls.fork, 0));
      execute_fused(lanes[l].data(), n, layout, p, ls.stream[l]);
      p += fused_total_bytes(lanes[l].data(), n);
    }
    for (uint32_t l = 0; l < lanes.size(); ++l) {
      KSJ_HIP(hipEventRecord(ls.join[l], ls.stream[l]));
      KSJ_HIP(hipStreamWaitEvent(stream, ls.join[l], 0));
    }
    return;
  }
  require(ws_bytes >= lanes_total_bytes(items, count), "multi adjoint workspace too small");
  if (m.lanes == 1) {
    for (uint32_t j = 0; j < count; ++j) {
      const MultiItem& it = items[m.order[j]];
      adjoint2_execute(*it.plan, &it.scan, 1, it.channels, layout, d_ws, m.lane_ws[0], stream);
    }
    return;
  }
  Lanes& lanes = multi_lanes_for(stream);
  KSJ_HIP(hipEventRecord(lanes.fork, stream));
  std::vector<char*> base(m.lanes);
  char* p = static_cast<char*>(d_ws);
  for (uint32_t l = 0; l < m.lanes; ++l) {
    KSJ_HIP(hipStreamWaitEvent(lanes.stream[l], lanes.fork, 0));
    base[l] = p;
    p += align256(m.lane_ws[l]);
  }
  for (uint32_t j = 0; j < count; ++j) {
    const uint32_t l = j % m.lanes;
    const MultiItem& it = items[m.order[j]];
    adjoint2_execute(*it.plan, &it.scan, 1, it.channels, layout, base[l], m.lane_ws[l], lanes.stream[l]);
  }
  for (uint32_t l = 0; l < m.lanes; ++l) {
    KSJ_HIP(hipEventRecord(lanes.join[l], lanes.stream[l]));
    KSJ_HIP(hipStreamWaitEvent(stream, lanes.join[l], 0));
  }
}

size_t nudft2_adjoint_workspace_bytes(uint32_t channels, uint32_t rows, uint32_t cols) {
  return align256(static_cast<size_t>(channels) * rows * cols * sizeof(float2));
}

}  // namespace ksj::gpu
