// The operators.  Each process function validates the firing in the order of
// its reference counterpart (every rejection of an input comes before the
// output grant is acquired; a grant acquired and then found invalid is
// released), stages the host-pageable payloads to HBM, runs the HIP path of
// libksj_gpu.so on the operator's stream, and copies the result into the grant.
#include "ops.hpp"

#include <cmath>
#include <limits>
#include <numbers>

namespace ksj::prov {

namespace {

constexpr uint32_t kMaxDim = 512;
constexpr uint32_t kMaxChannels = 64;
constexpr uint32_t kMaxSamples = 65536;             // reference envelope of the non-Cartesian operators
constexpr uint32_t kMaxKbSamples = uint32_t{1} << 22;
constexpr uint64_t kMaxPixels = uint64_t{kMaxDim} * kMaxDim;
constexpr uint64_t kCf32 = 8;
constexpr uint64_t kMaxCoilImageBytes = kMaxPixels * kMaxChannels * kCf32;  // 128 MiB
constexpr uint64_t kMaxImageBytes = kMaxPixels * 4;                         // 1 MiB
constexpr uint64_t kMaxComplexImageBytes = kMaxPixels * kCf32;               // 2 MiB
constexpr uint64_t kMaxDirectWork = uint64_t{1} << 28;
constexpr uint64_t kCaps3 = KSJ_LEASE_CAP_INPUT_BATCHES | KSJ_LEASE_CAP_OUTPUT_GRANTS | KSJ_LEASE_CAP_SCRATCH;
constexpr uint64_t kCaps2 = KSJ_LEASE_CAP_INPUT_BATCHES | KSJ_LEASE_CAP_OUTPUT_GRANTS;

uint64_t pixels(const Shape& s) { return uint64_t{s.rows} * s.cols; }
uint64_t coil_image_bytes(const Shape& s) { return pixels(s) * s.channels * kCf32; }
uint64_t kspace_bytes(const Shape& s) { return uint64_t{s.channels} * s.samples * kCf32; }
uint64_t trajectory_bytes(const Shape& s) { return uint64_t{s.samples} * 2 * 4; }

bool dims_ok(const Shape& s, uint32_t min_dim) {
  return s.channels >= 1 && s.channels <= kMaxChannels && s.rows >= min_dim && s.rows <= kMaxDim &&
         s.cols >= min_dim && s.cols <= kMaxDim;
}

// Shared texts.
constexpr std::string_view kBadAbi = "provider received ABI storage it cannot use";
constexpr std::string_view kInvalidLifecycle = "provider received an invalid lifecycle argument";

// =================================================================== Cartesian

// Reference: providers/kspacejet-cartesian-recon/src/operators/cartesian_ifft2_coil_images.cpp.
namespace cart {
constexpr std::string_view kConfig = "cartesian_ifft2_coil_images needs the canonical config "
                                     "{\"channels\":N,\"cols\":N,\"rows\":N} with channels in [1,64] and rows, cols "
                                     "in [2,512]";
constexpr std::string_view kLease = "cartesian_ifft2_coil_images needs one input frame, one output grant and its "
                                    "declared transform scratch";
constexpr std::string_view kFrame = "cartesian_ifft2_coil_images needs one exact ksj.kspace-frame payload";
constexpr std::string_view kOutput = "cartesian_ifft2_coil_images needs one exact ksj.coil-image-frame output grant";
constexpr std::string_view kScratch = "cartesian_ifft2_coil_images scratch grant is too small or unusable";

bool configure(const ksj_byte_view& c, Shape& s) {
  return parse_config(c, {u32_field("channels", 1, kMaxChannels, &s.channels), u32_field("cols", 2, kMaxDim, &s.cols),
                          u32_field("rows", 2, kMaxDim, &s.rows)});
}
bool valid(const Shape& s) { return dims_ok(s, 2); }
// One plane + one intermediate plane + two line buffers of cf32
// (cartesian_ifft2_coil_images.cpp:125-138): kept as the declared host scratch
// so a host reserves exactly what it reserves for the CPU operator.
uint64_t scratch_bytes(const Shape& s) { return (2 * pixels(s) + 2 * uint64_t{std::max(s.rows, s.cols)}) * kCf32; }

ksj_status process(Firing& f) {
  if (!callbacks_ok(f.cb, kCaps3)) return reject(f.error, KSJ_STATUS_BAD_ABI, kBadAbi);
  const Shape& s = f.op.shape;
  const uint64_t bytes = coil_image_bytes(s);
  ksj_firing_lease_info info;
  if (ksj_status st = lease_info(f, 1, bytes, scratch_bytes(s), true, kLease, info)) return st;
  ksj_input_batch_view batch{};
  batch.abi = make_header(sizeof(batch));
  if (ksj_status st = f.cb->get_input_batch(f.cb->host_context, f.lease, 0, &batch, f.error)) return st;
  if (!single_item_batch(batch, 0)) return reject(f.error, KSJ_STATUS_CONTRACT_VIOLATION, kFrame);
  const ksj_input_item_view& in = batch.items[0];
  if (!input_payload_ok(in, bytes, TypeId::kKspaceFrame)) return reject(f.error, KSJ_STATUS_CONTRACT_VIOLATION, kFrame);

  GrantGuard guard;
  ksj_mutable_payload_view out;
  if (ksj_status st = output_grant(f, bytes, TypeId::kCoilImageFrame, kOutput, guard, out)) return st;
  // The reference takes its scratch after mapping the output (cartesian_ifft2_coil_images.cpp:362-373).
  if (ksj_status st = scratch_ok(f, scratch_bytes(s), 4, kScratch)) return st;

  DeviceSession& d = device(f);
  void* d_in = d.buffer(DeviceSession::kInput, bytes);
  void* d_out = d.buffer(DeviceSession::kOutput, bytes);
  const size_t ws_bytes = ksj_gpu_ifft2_workspace_bytes(s.channels, s.rows, s.cols);
  void* ws = d.buffer(DeviceSession::kWorkspace, ws_bytes);
  d.upload(d_in, in.payload.data, bytes);
  check_gpu(ksj_gpu_ifft2_coil_images(d_in, d_out, s.channels, s.rows, s.cols, 0, KSJ_GPU_LAYOUT_COIL_LAST, ws,
                                      ws_bytes, d.stream()),
            "ifft2");
  d.download(out.data, d_out, bytes);
  // The Cartesian reference seals an empty metadata view (cartesian_ifft2_coil_images.cpp:414-423).
  return seal_and_finish(f, guard, out, bytes, in, false, 1, info.terminal_epoch);
}
}  // namespace cart

// Centred or uncentred inverse FFT fused with the RSS combine (BASELINE Cartesian recon:
// ifft2 per coil then root-sum-of-squares) -> ksj.image-frame.  No reference operator.
namespace cart_rss {
constexpr std::string_view kConfig = "cartesian_ifft2_rss needs the canonical config "
                                     "{\"centred\":false|true,\"channels\":N,\"cols\":N,\"rows\":N} with channels in "
                                     "[1,64] and rows, cols in [2,512]";
constexpr std::string_view kLease = "cartesian_ifft2_rss needs one input frame and one output grant";
constexpr std::string_view kFrame = "cartesian_ifft2_rss needs one exact ksj.kspace-frame payload";
constexpr std::string_view kOutput = "cartesian_ifft2_rss needs one exact ksj.image-frame output grant";

bool configure(const ksj_byte_view& c, Shape& s) {
  return parse_config(c, {literal_field("centred", "false|true", &s.centred),
                          u32_field("channels", 1, kMaxChannels, &s.channels), u32_field("cols", 2, kMaxDim, &s.cols),
                          u32_field("rows", 2, kMaxDim, &s.rows)});
}
bool valid(const Shape& s) { return dims_ok(s, 2) && s.centred <= 1; }

ksj_status process(Firing& f) {
  if (!callbacks_ok(f.cb, kCaps2)) return reject(f.error, KSJ_STATUS_BAD_ABI, kBadAbi);
  const Shape& s = f.op.shape;
  const uint64_t in_bytes = coil_image_bytes(s), out_bytes = pixels(s) * 4;
  ksj_firing_lease_info info;
  if (ksj_status st = lease_info(f, 1, out_bytes, 0, false, kLease, info)) return st;
  ksj_input_batch_view batch{};
  batch.abi = make_header(sizeof(batch));
  if (ksj_status st = f.cb->get_input_batch(f.cb->host_context, f.lease, 0, &batch, f.error)) return st;
  if (!single_item_batch(batch, 0)) return reject(f.error, KSJ_STATUS_CONTRACT_VIOLATION, kFrame);
  const ksj_input_item_view& in = batch.items[0];
  if (!input_payload_ok_device(in, in_bytes, TypeId::kKspaceFrame, f.op.device_ordinal))
    return reject(f.error, KSJ_STATUS_CONTRACT_VIOLATION, kFrame);
  GrantGuard guard;
  ksj_mutable_payload_view out;
  if (ksj_status st = output_grant(f, out_bytes, TypeId::kImageFrame, kOutput, guard, out, true)) return st;

  DeviceSession& d = device(f);
  const void* d_in = device_input(d, DeviceSession::kInput, in.payload, in_bytes);
  void* d_coils = d.buffer(DeviceSession::kInput2, in_bytes);
  void* d_out = device_output(d, out, out_bytes);
  const size_t ws_bytes = ksj_gpu_ifft2_workspace_bytes(s.channels, s.rows, s.cols);
  void* ws = d.buffer(DeviceSession::kWorkspace, ws_bytes);
  check_gpu(ksj_gpu_ifft2_coil_images(d_in, d_coils, s.channels, s.rows, s.cols, static_cast<int>(s.centred),
                                      KSJ_GPU_LAYOUT_COIL_LAST, ws, ws_bytes, d.stream()),
            "ifft2");
  check_gpu(ksj_gpu_rss(d_coils, static_cast<float*>(d_out), pixels(s), s.channels, d.stream()), "rss");
  finish_output(d, out, d_out, out_bytes);
  return seal_and_finish(f, guard, out, out_bytes, in, false, 1, info.terminal_epoch);
}
}  // namespace cart_rss

// =================================================================== coil combine

// Reference: providers/kspacejet-coil-combine/src/operators/coil_combine_rss.cpp:232-330.
namespace rss {
constexpr std::string_view kConfig = "coil_combine_rss needs the canonical config {\"channels\":N,\"cols\":N,\"rows\":N} "
                                     "with channels in [1,64] and rows, cols in [2,512]";
constexpr std::string_view kLease = "coil_combine_rss needs one input frame and one output grant";
constexpr std::string_view kInput = "coil_combine_rss needs one exact ksj.coil-image-frame payload";
constexpr std::string_view kOutput = "coil_combine_rss needs one exact ksj.image-frame output grant";

bool configure(const ksj_byte_view& c, Shape& s) {
  return parse_config(c, {u32_field("channels", 1, kMaxChannels, &s.channels), u32_field("cols", 2, kMaxDim, &s.cols),
                          u32_field("rows", 2, kMaxDim, &s.rows)});
}
bool valid(const Shape& s) { return dims_ok(s, 2); }

ksj_status process(Firing& f) {
  if (!callbacks_ok(f.cb, kCaps2)) return reject(f.error, KSJ_STATUS_BAD_ABI, kBadAbi);
  const Shape& s = f.op.shape;
  const uint64_t in_bytes = coil_image_bytes(s), out_bytes = pixels(s) * 4;
  ksj_firing_lease_info info;
  if (ksj_status st = lease_info(f, 1, out_bytes, 0, false, kLease, info)) return st;
  ksj_input_batch_view batch{};
  batch.abi = make_header(sizeof(batch));
  if (ksj_status st = f.cb->get_input_batch(f.cb->host_context, f.lease, 0, &batch, f.error)) return st;
  if (!single_item_batch(batch, 0)) return reject(f.error, KSJ_STATUS_CONTRACT_VIOLATION, kInput);
  const ksj_input_item_view& in = batch.items[0];
  if (!input_payload_ok(in, in_bytes, TypeId::kCoilImageFrame))
    return reject(f.error, KSJ_STATUS_CONTRACT_VIOLATION, kInput);
  GrantGuard guard;
  ksj_mutable_payload_view out;
  if (ksj_status st = output_grant(f, out_bytes, TypeId::kImageFrame, kOutput, guard, out)) return st;

  DeviceSession& d = device(f);
  void* d_in = d.buffer(DeviceSession::kInput, in_bytes);
  void* d_out = d.buffer(DeviceSession::kOutput, out_bytes);
  d.upload(d_in, in.payload.data, in_bytes);
  check_gpu(ksj_gpu_rss(d_in, static_cast<float*>(d_out), pixels(s), s.channels, d.stream()), "rss");
  d.download(out.data, d_out, out_bytes);
  return seal_and_finish(f, guard, out, out_bytes, in, false, 1, info.terminal_epoch);
}
}  // namespace rss

// Sensitivity-weighted combine, x = sum_c conj(S_c) I_c / sum_c |S_c|^2.  Ports
// follow the planned interface providers/interfaces/org.kspacejet.coil-combine/
// coil_combine_adaptive.json: port 0 coil images, port 1 sensitivity maps
// (ksj.sensitivity-map), output ksj.complex-image-frame.  No reference implementation.
namespace sense {
constexpr std::string_view kConfig = "coil_combine_sensitivity needs the canonical config "
                                     "{\"channels\":N,\"cols\":N,\"rows\":N} with channels in [1,64] and rows, cols "
                                     "in [2,512]";
constexpr std::string_view kLease = "coil_combine_sensitivity needs two input batches and one output grant";
constexpr std::string_view kImages = "coil_combine_sensitivity needs one exact ksj.coil-image-frame on input port 0";
constexpr std::string_view kMaps = "coil_combine_sensitivity needs one exact ksj.sensitivity-map on input port 1";
constexpr std::string_view kOutput = "coil_combine_sensitivity needs one exact ksj.complex-image-frame output grant";

bool configure(const ksj_byte_view& c, Shape& s) { return rss::configure(c, s); }
bool valid(const Shape& s) { return dims_ok(s, 2); }

ksj_status process(Firing& f) {
  if (!callbacks_ok(f.cb, kCaps2)) return reject(f.error, KSJ_STATUS_BAD_ABI, kBadAbi);
  const Shape& s = f.op.shape;
  const uint64_t in_bytes = coil_image_bytes(s), out_bytes = pixels(s) * kCf32;
  ksj_firing_lease_info info;
  if (ksj_status st = lease_info(f, 2, out_bytes, 0, false, kLease, info)) return st;
  ksj_input_batch_view storage[2];
  const ksj_input_batch_view *images = nullptr, *maps = nullptr;
  if (ksj_status st = two_port_batches(f, storage, images, maps, kLease, kImages, kMaps)) return st;
  const ksj_input_item_view& img = images->items[0];
  const ksj_input_item_view& map = maps->items[0];
  if (!input_payload_ok_device(img, in_bytes, TypeId::kCoilImageFrame, f.op.device_ordinal))
    return reject(f.error, KSJ_STATUS_CONTRACT_VIOLATION, kImages);
  if (!input_payload_ok_device(map, in_bytes, TypeId::kSensitivityMap, f.op.device_ordinal))
    return reject(f.error, KSJ_STATUS_CONTRACT_VIOLATION, kMaps);
  GrantGuard guard;
  ksj_mutable_payload_view out;
  if (ksj_status st = output_grant(f, out_bytes, TypeId::kComplexImageFrame, kOutput, guard, out, true)) return st;

  DeviceSession& d = device(f);
  const void* d_img = device_input(d, DeviceSession::kInput, img.payload, in_bytes);
  const void* d_map = device_input(d, DeviceSession::kInput2, map.payload, in_bytes);
  void* d_out = device_output(d, out, out_bytes);
  check_gpu(ksj_gpu_sense_combine(d_img, d_map, d_out, pixels(s), s.channels, d.stream()), "sensitivity combine");
  finish_output(d, out, d_out, out_bytes);
  return seal_and_finish(f, guard, out, out_bytes, img, false, 2, info.terminal_epoch);
}
}  // namespace sense

// The same combine under the planned interface's own id
// (providers/interfaces/org.kspacejet.coil-combine/coil_combine_adaptive.json: coil images
// and the maps of sensitivity_map_estimate in, one complex image out).  No reference
// implementation; the interface's optional "normalization" key is not taken.
namespace adaptive {
constexpr std::string_view kConfig = "coil_combine_adaptive needs the canonical config "
                                     "{\"channels\":N,\"cols\":N,\"rows\":N} with channels in [1,64] and rows, cols "
                                     "in [2,512]";
constexpr std::string_view kLease = "coil_combine_adaptive needs two input batches and one output grant";
constexpr std::string_view kImages = "coil_combine_adaptive needs one exact ksj.coil-image-frame on input port 0";
constexpr std::string_view kMaps = "coil_combine_adaptive needs one exact ksj.sensitivity-map on input port 1";
constexpr std::string_view kOutput = "coil_combine_adaptive needs one exact ksj.complex-image-frame output grant";

ksj_status process(Firing& f) {
  if (!callbacks_ok(f.cb, kCaps2)) return reject(f.error, KSJ_STATUS_BAD_ABI, kBadAbi);
  const Shape& s = f.op.shape;
  const uint64_t in_bytes = coil_image_bytes(s), out_bytes = pixels(s) * kCf32;
  ksj_firing_lease_info info;
  if (ksj_status st = lease_info(f, 2, out_bytes, 0, false, kLease, info)) return st;
  ksj_input_batch_view storage[2];
  const ksj_input_batch_view *images = nullptr, *maps = nullptr;
  if (ksj_status st = two_port_batches(f, storage, images, maps, kLease, kImages, kMaps)) return st;
  const ksj_input_item_view& img = images->items[0];
  const ksj_input_item_view& map = maps->items[0];
  if (!input_payload_ok_device(img, in_bytes, TypeId::kCoilImageFrame, f.op.device_ordinal))
    return reject(f.error, KSJ_STATUS_CONTRACT_VIOLATION, kImages);
  if (!input_payload_ok_device(map, in_bytes, TypeId::kSensitivityMap, f.op.device_ordinal))
    return reject(f.error, KSJ_STATUS_CONTRACT_VIOLATION, kMaps);
  GrantGuard guard;
  ksj_mutable_payload_view out;
  if (ksj_status st = output_grant(f, out_bytes, TypeId::kComplexImageFrame, kOutput, guard, out, true)) return st;

  DeviceSession& d = device(f);
  const void* d_img = device_input(d, DeviceSession::kInput, img.payload, in_bytes);
  const void* d_map = device_input(d, DeviceSession::kInput2, map.payload, in_bytes);
  void* d_out = device_output(d, out, out_bytes);
  check_gpu(ksj_gpu_sense_combine(d_img, d_map, d_out, pixels(s), s.channels, d.stream()), "adaptive combine");
  finish_output(d, out, d_out, out_bytes);
  return seal_and_finish(f, guard, out, out_bytes, img, false, 2, info.terminal_epoch);
}
}  // namespace adaptive

// =================================================================== sensitivity maps

// sensitivity_map_estimate of the planned interface
// providers/interfaces/org.kspacejet.calibration/sensitivity_map_estimate.json: coil images
// in, ksj.sensitivity-map out, method required.  Method "walsh" (the adaptive combine's
// estimate): the dominant eigenvector of the windowed coil covariance at each pixel.  No
// reference implementation; the optional calibration_region and regularization keys are not
// taken.
namespace smap {
constexpr std::string_view kConfig = "sensitivity_map_estimate needs the canonical config "
                                     "{\"channels\":N,\"cols\":N,\"iterations\":N,\"method\":\"walsh\",\"rows\":N,"
                                     "\"window\":N} with channels in [1,64], rows, cols in [2,512], iterations in "
                                     "[1,64] and an odd window in [1,31]";
constexpr std::string_view kLease = "sensitivity_map_estimate needs one input frame and one output grant";
constexpr std::string_view kImages = "sensitivity_map_estimate needs one exact ksj.coil-image-frame payload";
constexpr std::string_view kOutput = "sensitivity_map_estimate needs one exact ksj.sensitivity-map output grant";

bool configure(const ksj_byte_view& c, Shape& s) {
  return parse_config(c, {u32_field("channels", 1, kMaxChannels, &s.channels), u32_field("cols", 2, kMaxDim, &s.cols),
                          u32_field("iterations", 1, 64, &s.iterations), literal_field("method", "\"walsh\"", &s.method),
                          u32_field("rows", 2, kMaxDim, &s.rows), u32_field("window", 1, 31, &s.window)}) &&
         s.window % 2 == 1;  // a centred neighbourhood
}
bool valid(const Shape& s) {
  return dims_ok(s, 2) && s.iterations >= 1 && s.iterations <= 64 && s.window % 2 == 1 && s.window <= 31 &&
         s.method == 0;
}

ksj_status process(Firing& f) {
  if (!callbacks_ok(f.cb, kCaps2)) return reject(f.error, KSJ_STATUS_BAD_ABI, kBadAbi);
  const Shape& s = f.op.shape;
  const uint64_t bytes = coil_image_bytes(s);
  ksj_firing_lease_info info;
  if (ksj_status st = lease_info(f, 1, bytes, 0, false, kLease, info)) return st;
  ksj_input_batch_view batch{};
  batch.abi = make_header(sizeof(batch));
  if (ksj_status st = f.cb->get_input_batch(f.cb->host_context, f.lease, 0, &batch, f.error)) return st;
  if (!single_item_batch(batch, 0)) return reject(f.error, KSJ_STATUS_CONTRACT_VIOLATION, kImages);
  const ksj_input_item_view& in = batch.items[0];
  if (!input_payload_ok_device(in, bytes, TypeId::kCoilImageFrame, f.op.device_ordinal))
    return reject(f.error, KSJ_STATUS_CONTRACT_VIOLATION, kImages);
  GrantGuard guard;
  ksj_mutable_payload_view out;
  if (ksj_status st = output_grant(f, bytes, TypeId::kSensitivityMap, kOutput, guard, out, true)) return st;

  DeviceSession& d = device(f);
  const void* d_in = device_input(d, DeviceSession::kInput, in.payload, bytes);
  void* d_out = device_output(d, out, bytes);
  check_gpu(ksj_gpu_sensitivity_maps_walsh(d_in, s.rows, s.cols, s.channels, s.window, s.iterations, d_out,
                                           d.stream()),
            "sensitivity maps");
  finish_output(d, out, d_out, bytes);
  return seal_and_finish(f, guard, out, bytes, in, false, 1, info.terminal_epoch);
}
}  // namespace smap

// =================================================================== non-Cartesian

// The two-input firing shared by the non-Cartesian operators
// (radial_gridding_reconstruct.cpp:228-330 and noncartesian_adjoint_reconstruct.cpp:202-316):
// k-space on port 0, trajectory on port 1, equal identity keys, then a
// trajectory value check.
struct NoncartesianTexts {
  std::string_view lease, kspace, trajectory, identity, output, scratch;
};

enum class TrajectoryRule { kFinite, kRadiansPerPixel };

bool trajectory_ok(const float* v, uint64_t n, TrajectoryRule rule) {
  constexpr float kPi = std::numbers::pi_v<float>;
  for (uint64_t i = 0; i < n; ++i) {
    if (!std::isfinite(v[i])) return false;
    if (rule == TrajectoryRule::kRadiansPerPixel && (v[i] < -kPi || v[i] > kPi)) return false;
  }
  return true;
}

struct NoncartesianInputs {
  const ksj_input_item_view* kspace = nullptr;
  const ksj_input_item_view* trajectory = nullptr;
  bool resident = false;  // the trajectory's bytes are those of the session's resident plan
  ksj_firing_lease_info info{};
  ksj_input_batch_view storage[2];
};

// `plan_desc`: the gridding plan the firing will run (operators with a resident plan).
// A trajectory whose bytes equal the resident plan's passed the value check when that
// plan was built, so the check is skipped for it.  `allow_device`: the payloads may
// also be device-resident on the operator's GPU (KSJ_PROVIDER_MEMORY_DEVICE); a device
// trajectory is compared with the resident plan's and value-checked on the GPU.
ksj_status noncartesian_inputs(Firing& f, const NoncartesianTexts& t, uint64_t output_bytes, uint64_t scratch_bytes,
                               bool check_scratch, TrajectoryRule rule, NoncartesianInputs& in,
                               const ksj_gpu_plan2_desc* plan_desc = nullptr, bool allow_device = false) {
  const Shape& s = f.op.shape;
  if (ksj_status st = lease_info(f, 2, output_bytes, scratch_bytes, check_scratch, t.lease, in.info)) return st;
  const ksj_input_batch_view *k = nullptr, *tr = nullptr;
  if (ksj_status st = two_port_batches(f, in.storage, k, tr, t.lease, t.kspace, t.trajectory)) return st;
  in.kspace = &k->items[0];
  in.trajectory = &tr->items[0];
  auto payload_ok = [&](const ksj_input_item_view& item, uint64_t bytes, TypeId type) {
    return allow_device ? input_payload_ok_device(item, bytes, type, f.op.device_ordinal)
                        : input_payload_ok(item, bytes, type);
  };
  if (!payload_ok(*in.kspace, kspace_bytes(s), TypeId::kNoncartesianKspaceFrame))
    return reject(f.error, KSJ_STATUS_CONTRACT_VIOLATION, t.kspace);
  if (!payload_ok(*in.trajectory, trajectory_bytes(s), TypeId::kTrajectoryFrame))
    return reject(f.error, KSJ_STATUS_CONTRACT_VIOLATION, t.trajectory);
  if (in.kspace->semantic_key_hash != in.trajectory->semantic_key_hash ||
      in.kspace->order_key != in.trajectory->order_key)
    return reject(f.error, KSJ_STATUS_CONTRACT_VIOLATION, t.identity);
  const auto* traj = static_cast<const float*>(in.trajectory->payload.data);
  if (on_device(in.trajectory->payload)) {
    if (plan_desc == nullptr) return reject(f.error, KSJ_STATUS_CONTRACT_VIOLATION, t.trajectory);
    const float bound = rule == TrajectoryRule::kRadiansPerPixel ? std::numbers::pi_v<float>
                                                                  : std::numeric_limits<float>::max();
    const uint32_t flags = device(f).check_device_trajectory(*plan_desc, traj, bound);
    if (flags & KSJ_GPU_TRAJ_INVALID) return reject(f.error, KSJ_STATUS_CONTRACT_VIOLATION, t.trajectory);
    in.resident = (flags & KSJ_GPU_TRAJ_DIFFERS) == 0;
    return KSJ_STATUS_OK;
  }
  in.resident = plan_desc != nullptr && f.op.device && f.op.device->resident(*plan_desc, traj);
  if (!in.resident && !trajectory_ok(traj, uint64_t{s.samples} * 2, rule))
    return reject(f.error, KSJ_STATUS_CONTRACT_VIOLATION, t.trajectory);
  return KSJ_STATUS_OK;
}

// Reference: providers/kspacejet-noncartesian-recon/src/operators/noncartesian_adjoint_reconstruct.cpp.
namespace direct {
constexpr std::string_view kConfig = "noncartesian_adjoint_reconstruct needs the canonical config "
                                     "{\"channels\":N,\"image_cols\":N,\"image_rows\":N,\"sample_count\":N} within "
                                     "its direct-adjoint bound (channels*rows*cols*samples <= 2^28)";
constexpr NoncartesianTexts kTexts{
    "noncartesian_adjoint_reconstruct needs two input batches, one output grant and one image scratch plane",
    "noncartesian_adjoint_reconstruct needs one exact ksj.noncartesian-kspace-frame on input port 0",
    "noncartesian_adjoint_reconstruct needs one exact finite ksj.trajectory-frame on input port 1",
    "non-Cartesian k-space and trajectory inputs must carry equal semantic_key_hash and order_key",
    "noncartesian_adjoint_reconstruct needs one exact ksj.coil-image-frame output grant",
    "noncartesian_adjoint_reconstruct scratch grant is too small or unusable"};

bool configure(const ksj_byte_view& c, Shape& s) {
  return parse_config(c, {u32_field("channels", 1, kMaxChannels, &s.channels),
                          u32_field("image_cols", 2, kMaxDim, &s.cols), u32_field("image_rows", 2, kMaxDim, &s.rows),
                          u32_field("sample_count", 1, kMaxSamples, &s.samples)}) &&
         pixels(s) * s.channels * s.samples <= kMaxDirectWork;  // an over-budget config is a config rejection
}
bool valid(const Shape& s) {
  return dims_ok(s, 2) && s.samples >= 1 && s.samples <= kMaxSamples &&
         pixels(s) * s.channels * s.samples <= kMaxDirectWork;
}
uint64_t scratch_bytes(const Shape& s) { return pixels(s) * kCf32; }

ksj_status process(Firing& f) {
  if (!callbacks_ok(f.cb, kCaps3)) return reject(f.error, KSJ_STATUS_BAD_ABI, kBadAbi);
  const Shape& s = f.op.shape;
  const uint64_t bytes = coil_image_bytes(s);
  NoncartesianInputs in;
  if (ksj_status st = noncartesian_inputs(f, kTexts, bytes, scratch_bytes(s), true, TrajectoryRule::kFinite, in))
    return st;
  if (ksj_status st = scratch_ok(f, scratch_bytes(s), 64, kTexts.scratch)) return st;
  GrantGuard guard;
  ksj_mutable_payload_view out;
  if (ksj_status st = output_grant(f, bytes, TypeId::kCoilImageFrame, kTexts.output, guard, out)) return st;

  DeviceSession& d = device(f);
  void* d_k = d.buffer(DeviceSession::kInput, kspace_bytes(s));
  void* d_t = d.buffer(DeviceSession::kTrajectory, trajectory_bytes(s));
  void* d_out = d.buffer(DeviceSession::kOutput, bytes);
  const size_t ws_bytes = ksj_gpu_nudft2_adjoint_workspace_bytes(s.channels, s.rows, s.cols);
  void* ws = d.buffer(DeviceSession::kWorkspace, ws_bytes);
  d.upload(d_k, in.kspace->payload.data, kspace_bytes(s));
  d.upload(d_t, in.trajectory->payload.data, trajectory_bytes(s));
  check_gpu(ksj_gpu_nudft2_adjoint(static_cast<const float*>(d_t), d_k, d_out, s.channels, s.samples, s.rows, s.cols,
                                   (s.rows - 1) / 2.0, (s.cols - 1) / 2.0, KSJ_GPU_LAYOUT_COIL_LAST, ws, ws_bytes,
                                   d.stream()),
            "direct adjoint");
  d.download(out.data, d_out, bytes);
  return seal_and_finish(f, guard, out, bytes, *in.kspace, true, 2, in.info.terminal_epoch);
}
}  // namespace direct

// Reference: providers/kspacejet-noncartesian-recon/src/operators/radial_gridding_reconstruct.cpp.
namespace radial {
constexpr std::string_view kConfig =
    "radial_gridding_reconstruct needs the canonical config {\"channels\":N,\"density_compensation\":"
    "\"radial_analytic_ramp\",\"image_cols\":N,\"image_rows\":N,\"sample_count\":N,\"trajectory_units\":"
    "\"radians_per_pixel\"} with power-of-two image dimensions within its radial-gridding envelope";
constexpr NoncartesianTexts kTexts{
    "radial_gridding_reconstruct needs two input batches, one output grant and its declared gridding scratch",
    "radial_gridding_reconstruct needs one exact ksj.noncartesian-kspace-frame on input port 0",
    "radial_gridding_reconstruct needs one exact finite ksj.trajectory-frame on input port 1 in radians_per_pixel "
    "within [-pi, pi]",
    "non-Cartesian radial k-space and trajectory inputs must carry equal semantic_key_hash and order_key",
    "radial_gridding_reconstruct needs one exact ksj.coil-image-frame output grant",
    "radial_gridding_reconstruct scratch grant is too small or unusable"};

bool configure(const ksj_byte_view& c, Shape& s) {
  return parse_config(c, {u32_field("channels", 1, kMaxChannels, &s.channels),
                          literal_field("density_compensation", "\"radial_analytic_ramp\""),
                          u32_field("image_cols", 2, kMaxDim, &s.cols), u32_field("image_rows", 2, kMaxDim, &s.rows),
                          u32_field("sample_count", 1, kMaxSamples, &s.samples),
                          literal_field("trajectory_units", "\"radians_per_pixel\"")}) &&
         is_power_of_two(s.rows) && is_power_of_two(s.cols);
}
bool valid(const Shape& s) {
  return dims_ok(s, 2) && is_power_of_two(s.rows) && is_power_of_two(s.cols) && s.samples >= 1 &&
         s.samples <= kMaxSamples;
}
// Image plane + FFT plane + two line buffers of cf32 + the float DCF
// (radial_gridding_reconstruct.cpp:175-181), reserved as on the CPU operator.
uint64_t scratch_bytes(const Shape& s) {
  return (2 * pixels(s) + 2 * uint64_t{std::max(s.rows, s.cols)}) * kCf32 + uint64_t{s.samples} * 4;
}

ksj_status process(Firing& f) {
  if (!callbacks_ok(f.cb, kCaps3)) return reject(f.error, KSJ_STATUS_BAD_ABI, kBadAbi);
  const Shape& s = f.op.shape;
  const uint64_t bytes = coil_image_bytes(s);
  ksj_gpu_plan2_desc desc{};
  desc.struct_size = sizeof(desc);
  desc.image_rows = desc.grid_rows = s.rows;
  desc.image_cols = desc.grid_cols = s.cols;
  desc.kernel = KSJ_GPU_KERNEL_LINEAR;
  desc.kernel_width = 2;
  desc.density = KSJ_GPU_DENSITY_RAMP;
  desc.centred = 0;
  desc.samples = s.samples;
  desc.row_origin = (s.rows - 1) / 2.0;
  desc.col_origin = (s.cols - 1) / 2.0;
  NoncartesianInputs in;
  if (ksj_status st = noncartesian_inputs(f, kTexts, bytes, scratch_bytes(s), true, TrajectoryRule::kRadiansPerPixel,
                                          in, &desc))
    return st;
  if (ksj_status st = scratch_ok(f, scratch_bytes(s), 64, kTexts.scratch)) return st;
  GrantGuard guard;
  ksj_mutable_payload_view out;
  if (ksj_status st = output_grant(f, bytes, TypeId::kCoilImageFrame, kTexts.output, guard, out)) return st;

  DeviceSession& d = device(f);
  ksj_gpu_plan2* plan =
      d.plan(desc, static_cast<const float*>(in.trajectory->payload.data), false, in.resident);
  void* d_k = d.buffer(DeviceSession::kInput, kspace_bytes(s));
  void* d_out = d.buffer(DeviceSession::kOutput, bytes);
  const size_t ws_bytes = ksj_gpu_adjoint2_workspace_bytes(plan, s.channels);
  void* ws = d.buffer(DeviceSession::kWorkspace, ws_bytes);
  d.upload(d_k, in.kspace->payload.data, kspace_bytes(s));
  check_gpu(ksj_gpu_adjoint2_execute(plan, d_k, s.channels, d_out, KSJ_GPU_LAYOUT_COIL_LAST, nullptr, ws, ws_bytes,
                                     d.stream()),
            "radial gridding");
  d.download(out.data, d_out, bytes);
  return seal_and_finish(f, guard, out, bytes, *in.kspace, true, 2, in.info.terminal_epoch);
}
}  // namespace radial

// Kaiser-Bessel gridding onto the sigma-oversampled grid, pruned inverse FFT,
// deapodize and centre crop: the adjoint NUFFT the reference's radial route
// stands in for.  Same ports and identity rules as radial_gridding_reconstruct;
// no host scratch.  Output: coil images (kb_gridding_reconstruct) or their RSS
// image (kb_gridding_reconstruct_rss).
namespace kb {
constexpr std::string_view kConfig =
    "kb_gridding_reconstruct needs the canonical config {\"channels\":N,\"density_compensation\":"
    "\"radial_analytic_ramp\"|\"none\",\"image_cols\":N,\"image_rows\":N,\"kernel_width\":4|6|8,"
    "\"oversampling\":2,\"sample_count\":N,\"trajectory_units\":\"radians_per_pixel\"} with channels in [1,64], "
    "image dimensions in [8,512] and sample_count in [1,4194304]";
constexpr NoncartesianTexts kTexts{
    "kb_gridding_reconstruct needs two input batches and one output grant",
    "kb_gridding_reconstruct needs one exact ksj.noncartesian-kspace-frame on input port 0",
    "kb_gridding_reconstruct needs one exact finite ksj.trajectory-frame on input port 1 in radians_per_pixel "
    "within [-pi, pi]",
    "kb_gridding_reconstruct k-space and trajectory inputs must carry equal semantic_key_hash and order_key",
    "kb_gridding_reconstruct needs one exact output grant of its output type",
    ""};

// Shape.centred carries the density mode here: 0 = radial_analytic_ramp, 1 = none.
bool configure(const ksj_byte_view& c, Shape& s) {
  return parse_config(c, {u32_field("channels", 1, kMaxChannels, &s.channels),
                          literal_field("density_compensation", "\"radial_analytic_ramp\"|\"none\"", &s.centred),
                          u32_field("image_cols", 8, kMaxDim, &s.cols), u32_field("image_rows", 8, kMaxDim, &s.rows),
                          u32_field("kernel_width", 4, 8, &s.kernel_width),
                          u32_field("oversampling", 2, 2, &s.oversampling),
                          u32_field("sample_count", 1, kMaxKbSamples, &s.samples),
                          literal_field("trajectory_units", "\"radians_per_pixel\"")}) &&
         s.kernel_width % 2 == 0;
}
bool valid(const Shape& s) {
  return dims_ok(s, 8) && s.samples >= 1 && s.samples <= kMaxKbSamples && s.oversampling == 2 &&
         (s.kernel_width == 4 || s.kernel_width == 6 || s.kernel_width == 8) && s.centred <= 1;
}

ksj_status run(Firing& f, bool rss_output) {
  if (!callbacks_ok(f.cb, kCaps2)) return reject(f.error, KSJ_STATUS_BAD_ABI, kBadAbi);
  const Shape& s = f.op.shape;
  const uint64_t bytes = rss_output ? pixels(s) * 4 : coil_image_bytes(s);
  ksj_gpu_plan2_desc desc{};
  desc.struct_size = sizeof(desc);
  desc.image_rows = s.rows;
  desc.image_cols = s.cols;
  desc.grid_rows = s.rows * s.oversampling;
  desc.grid_cols = s.cols * s.oversampling;
  desc.kernel = KSJ_GPU_KERNEL_KAISER_BESSEL;
  desc.kernel_width = s.kernel_width;
  desc.density = s.centred == 0 ? KSJ_GPU_DENSITY_RAMP : KSJ_GPU_DENSITY_NONE;
  desc.centred = 1;
  desc.samples = s.samples;
  desc.row_origin = (s.rows - 1) / 2.0;
  desc.col_origin = (s.cols - 1) / 2.0;
  NoncartesianInputs in;
  if (ksj_status st =
          noncartesian_inputs(f, kTexts, bytes, 0, false, TrajectoryRule::kRadiansPerPixel, in, &desc, true))
    return st;
  GrantGuard guard;
  ksj_mutable_payload_view out;
  if (ksj_status st = output_grant(f, bytes, rss_output ? TypeId::kImageFrame : TypeId::kCoilImageFrame, kTexts.output,
                                   guard, out, true))
    return st;

  DeviceSession& d = device(f);
  ksj_gpu_plan2* plan = d.plan(desc, static_cast<const float*>(in.trajectory->payload.data),
                               on_device(in.trajectory->payload), in.resident);
  // Device-resident payloads are read and written in place; host ones go through the
  // session's staging.
  const void* d_k = device_input(d, DeviceSession::kInput, in.kspace->payload, kspace_bytes(s));
  void* d_out = device_output(d, out, bytes);
  const size_t ws_bytes = ksj_gpu_adjoint2_workspace_bytes(plan, s.channels);
  void* ws = d.buffer(DeviceSession::kWorkspace, ws_bytes);
  check_gpu(ksj_gpu_adjoint2_execute(plan, d_k, s.channels, rss_output ? nullptr : d_out, KSJ_GPU_LAYOUT_COIL_LAST,
                                     rss_output ? static_cast<float*>(d_out) : nullptr, ws, ws_bytes, d.stream()),
            "kaiser-bessel gridding");
  finish_output(d, out, d_out, bytes);
  return seal_and_finish(f, guard, out, bytes, *in.kspace, true, 2, in.info.terminal_epoch);
}
ksj_status process_images(Firing& f) { return run(f, false); }
ksj_status process_rss(Firing& f) { return run(f, true); }
}  // namespace kb

constexpr uint64_t kMaxKbInputBytes = uint64_t{kMaxChannels} * kMaxKbSamples * kCf32;
static_assert(kMaxKbInputBytes == uint64_t{2} << 30, "2 GiB k-space bound of the KB operators");

}  // namespace

const OpSpec kCartesianIfft2CoilImages{"cartesian_ifft2_coil_images", cart::kConfig, 1, kMaxCoilImageBytes,
                                       (2 * kMaxPixels + 2 * kMaxDim) * kCf32, &cart::configure, &cart::valid,
                                       &cart::process};
const OpSpec kCoilCombineRss{"coil_combine_rss", rss::kConfig, 1, kMaxImageBytes, 0, &rss::configure, &rss::valid,
                             &rss::process};
const OpSpec kNoncartesianAdjoint{"noncartesian_adjoint_reconstruct", direct::kConfig, 2, kMaxCoilImageBytes,
                                  kMaxPixels * kCf32, &direct::configure, &direct::valid, &direct::process};
const OpSpec kRadialGridding{"radial_gridding_reconstruct", radial::kConfig, 2, kMaxCoilImageBytes,
                             (2 * kMaxPixels + 2 * kMaxDim) * kCf32 + uint64_t{kMaxSamples} * 4, &radial::configure,
                             &radial::valid, &radial::process};
const OpSpec kKbGridding{"kb_gridding_reconstruct", kb::kConfig, 2, kMaxCoilImageBytes, 0, &kb::configure, &kb::valid,
                         &kb::process_images};
const OpSpec kKbGriddingRss{"kb_gridding_reconstruct_rss", kb::kConfig, 2, kMaxImageBytes, 0, &kb::configure,
                            &kb::valid, &kb::process_rss};
const OpSpec kCartesianIfft2Rss{"cartesian_ifft2_rss", cart_rss::kConfig, 1, kMaxImageBytes, 0, &cart_rss::configure,
                                &cart_rss::valid, &cart_rss::process};
const OpSpec kCoilCombineSensitivity{"coil_combine_sensitivity", sense::kConfig, 2, kMaxComplexImageBytes, 0,
                                     &sense::configure, &sense::valid, &sense::process};
const OpSpec kCoilCombineAdaptive{"coil_combine_adaptive", adaptive::kConfig, 2, kMaxComplexImageBytes, 0,
                                  &sense::configure, &sense::valid, &adaptive::process};
const OpSpec kSensitivityMapEstimate{"sensitivity_map_estimate", smap::kConfig, 1, kMaxCoilImageBytes, 0,
                                     &smap::configure, &smap::valid, &smap::process};

}  // namespace ksj::prov
