// Pipeline orchestration (host side; every step is enqueued on one stream).
#include "recon.hpp"

#include "combine.hpp"
#include "fft.hpp"
#include "grid.hpp"

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
  a.columns = 0;
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
  b.columns = 1;
  fft_launch(b, stream);
  if (layout == kLayoutCoilLast)
    coil_major_to_last(img, d_images, nullptr, static_cast<uint64_t>(rows) * cols, channels, stream);
}

size_t adjoint2_workspace_bytes(const Plan2& plan, uint32_t channels) {
  const auto& d = plan.d;
  const size_t cp = padded_coils(channels);
  const size_t sorted = static_cast<size_t>(d.samples) * cp * sizeof(float2);
  const size_t grid = static_cast<size_t>(channels) * d.grid_rows * d.grid_cols * sizeof(float2);
  const size_t rowout = static_cast<size_t>(channels) * d.grid_rows * d.image_cols * sizeof(float2);
  const size_t img = static_cast<size_t>(channels) * d.image_rows * d.image_cols * sizeof(float2);
  return align256(sorted) + align256(grid) + align256(rowout) + align256(img);
}

void adjoint2_execute(const Plan2& plan, const float2* d_kspace, uint32_t channels, float2* d_images, int layout,
                      float* d_rss, void* d_ws, size_t ws_bytes, hipStream_t stream) {
  const auto& d = plan.d;
  require(plan.has_trajectory, "plan has no trajectory");
  require(channels >= 1 && channels <= 64, "adjoint supports 1..64 channels");
  require(layout == kLayoutCoilLast || layout == kLayoutCoilMajor, "unknown image layout");
  require(ws_bytes >= adjoint2_workspace_bytes(plan, channels), "adjoint workspace too small");
  char* p = static_cast<char*>(d_ws);
  const size_t cp = padded_coils(channels);
  auto* sorted = reinterpret_cast<float2*>(p);
  p += align256(static_cast<size_t>(d.samples) * cp * sizeof(float2));
  auto* grid = reinterpret_cast<float2*>(p);
  p += align256(static_cast<size_t>(channels) * d.grid_rows * d.grid_cols * sizeof(float2));
  auto* rowout = reinterpret_cast<float2*>(p);
  p += align256(static_cast<size_t>(channels) * d.grid_rows * d.image_cols * sizeof(float2));
  auto* img = reinterpret_cast<float2*>(p);

  grid_prep(plan, d_kspace, channels, sorted, stream);
  grid_gather(plan, sorted, channels, grid, stream);

  FftArgs a;
  a.in = grid;
  a.out = rowout;
  a.src = rows_of(channels, d.grid_rows, d.grid_cols, d.grid_cols);
  a.dst = rows_of(channels, d.grid_rows, d.image_cols, d.image_cols);
  a.n = d.grid_cols;
  a.n_out = d.image_cols;
  a.out_offset = d.centred ? (d.grid_cols - d.image_cols / 2) % d.grid_cols : 0;
  a.inverse = 1;
  a.columns = 0;
  fft_launch(a, stream);

  FftArgs b;
  b.in = rowout;
  b.out = (layout == kLayoutCoilMajor && d_images) ? d_images : img;
  b.src = cols_of(channels, d.grid_rows, d.image_cols);
  b.dst = cols_of(channels, d.image_rows, d.image_cols);
  b.n = d.grid_rows;
  b.n_out = d.image_rows;
  b.out_offset = d.centred ? (d.grid_rows - d.image_rows / 2) % d.grid_rows : 0;
  b.out_weight = plan.deapod_rows;
  b.line_weight = plan.deapod_cols;
  b.inverse = 1;
  b.columns = 1;
  fft_launch(b, stream);

  const uint64_t pixels = static_cast<uint64_t>(d.image_rows) * d.image_cols;
  float2* last = (layout == kLayoutCoilLast) ? d_images : nullptr;
  coil_major_to_last(b.out, last, d_rss, pixels, channels, stream);
}

size_t nudft2_adjoint_workspace_bytes(uint32_t channels, uint32_t rows, uint32_t cols) {
  return align256(static_cast<size_t>(channels) * rows * cols * sizeof(float2));
}

}  // namespace ksj::gpu
