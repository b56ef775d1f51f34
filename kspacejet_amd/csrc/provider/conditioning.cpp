// The operators of org.kspacejet.kspace-conditioning on the GPU: noise prewhitening and
// coil compression of Cartesian and non-Cartesian k-space (channel mixing by an explicit
// matrix), per-channel readout phase correction, and readout-oversampling removal.
//
// The firing follows the reference's shared flow
// (providers/kspacejet-kspace-conditioning/src/provider_api.cpp:323-437): lease counts and
// no scratch -> input ports by port number (0 = k-space, 1 = calibration) -> k-space item
// -> calibration item -> calibration identity -> reserved output bytes -> output grant ->
// transform -> seal with the k-space item's keys and empty metadata.  Payloads carry no
// metadata, as there.  The reference-named module keeps them host-pageable; the same
// operators registered in org.kspacejet.gpu-recon also take device-memory payloads and
// output grants (KSJ_PROVIDER_MEMORY_DEVICE) on the operator's GPU.
#include "ops.hpp"

namespace ksj::prov {

namespace {

constexpr uint64_t kCf32 = 8;
constexpr uint32_t kMaxChannels = 64;
constexpr uint32_t kMinDim = 2, kMaxDim = 512;
constexpr uint32_t kMaxSamples = 65536;
constexpr uint64_t kMaxKspaceBytes = uint64_t{kMaxChannels} * kMaxDim * kMaxDim * kCf32;
constexpr uint64_t kCaps = KSJ_LEASE_CAP_INPUT_BATCHES | KSJ_LEASE_CAP_OUTPUT_GRANTS;

constexpr std::string_view kBadAbi = "kspace-conditioning provider received ABI storage it cannot use";
constexpr std::string_view kLease =
    "kspace-conditioning operators need their configured input batches, one output grant and no host scratch";
constexpr std::string_view kPorts = "kspace-conditioning operators need one item on each of their distinct input ports";
constexpr std::string_view kDynamic =
    "kspace-conditioning operators need one exact channel-major k-space item on input port 0";
constexpr std::string_view kCalibration =
    "kspace-conditioning operators need one exact calibration item on input port 1";
constexpr std::string_view kIdentity =
    "a keyed calibration item must carry the k-space item's semantic_key_hash and order_key";
constexpr std::string_view kOutput = "kspace-conditioning operators need one exact k-space output grant";

enum class Kind { kNoise, kPhase, kCoil, kCrop, kNoncartesianNoise, kNoncartesianCoil };

struct Layout {
  uint32_t batches;          // 1 (crop) or 2
  TypeId dynamic, calibration, output;
  uint64_t dynamic_bytes, calibration_bytes, output_bytes;
};

bool noncartesian(Kind k) { return k == Kind::kNoncartesianNoise || k == Kind::kNoncartesianCoil; }

// Payload types and byte counts of a configured operator (the expected_*_byte_count and
// matches_*_type functions of src/operators/*.cpp).
Layout layout(Kind k, const Shape& s) {
  const uint64_t lines = uint64_t{s.channels} * s.rows;
  switch (k) {
    case Kind::kNoise:
      return {2, TypeId::kKspaceFrame, TypeId::kNoiseModel, TypeId::kKspaceFrame, lines * s.cols * kCf32,
              uint64_t{s.channels} * s.channels * kCf32, lines * s.cols * kCf32};
    case Kind::kPhase:
      return {2, TypeId::kKspaceFrame, TypeId::kPhaseModel, TypeId::kKspaceFrame, lines * s.cols * kCf32,
              uint64_t{s.channels} * s.cols * kCf32, lines * s.cols * kCf32};
    case Kind::kCoil:
      return {2, TypeId::kKspaceFrame, TypeId::kCoilCompressionBasis, TypeId::kKspaceFrame, lines * s.cols * kCf32,
              uint64_t{s.out_channels} * s.channels * kCf32, uint64_t{s.out_channels} * s.rows * s.cols * kCf32};
    case Kind::kCrop:
      return {1, TypeId::kKspaceFrame, TypeId::kKspaceFrame, TypeId::kKspaceFrame, lines * s.cols * kCf32, 0,
              lines * s.out_cols * kCf32};
    case Kind::kNoncartesianNoise:
      return {2, TypeId::kNoncartesianKspaceFrame, TypeId::kNoiseModel, TypeId::kNoncartesianKspaceFrame,
              uint64_t{s.channels} * s.samples * kCf32, uint64_t{s.channels} * s.channels * kCf32,
              uint64_t{s.channels} * s.samples * kCf32};
    case Kind::kNoncartesianCoil:
      return {2, TypeId::kNoncartesianKspaceFrame, TypeId::kCoilCompressionBasis, TypeId::kNoncartesianKspaceFrame,
              uint64_t{s.channels} * s.samples * kCf32, uint64_t{s.out_channels} * s.channels * kCf32,
              uint64_t{s.out_channels} * s.samples * kCf32};
  }
  return {};
}

// An input item as the reference accepts it: compatible headers, exact bytes, host-pageable,
// 64-byte aligned, no metadata, the exact registry type (provider_api.cpp:165-174).  With
// `device_edges` (the org.kspacejet.gpu-recon registrations) the payload may instead lie in
// device memory of the operator's GPU.
bool item_ok(const Firing& f, const ksj_input_item_view& item, uint64_t bytes, TypeId type, bool device_edges) {
  const bool payload = device_edges ? input_payload_ok_device(item, bytes, type, f.op.device_ordinal)
                                    : input_payload_ok(item, bytes, type);
  return abi_ok(&item) && payload && item.metadata.size == 0;
}

ksj_status process(Firing& f, Kind kind, bool device_edges) {
  if (!callbacks_ok(f.cb, kCaps)) return reject(f.error, KSJ_STATUS_BAD_ABI, kBadAbi);
  const Shape& s = f.op.shape;
  const Layout lay = layout(kind, s);

  ksj_firing_lease_info info{};
  info.abi = make_header(sizeof(info));
  if (ksj_status st = f.cb->get_info(f.cb->host_context, f.lease, &info, f.error)) return st;
  if (!abi_ok(&info) || info.input_batch_count != lay.batches || info.output_grant_count != 1 ||
      info.reserved_scratch_bytes != 0)
    return reject(f.error, KSJ_STATUS_CONTRACT_VIOLATION, kLease);

  // Ports by number, each at most once (provider_api.cpp:181-219).
  ksj_input_batch_view batches[2];
  const ksj_input_item_view* dynamic = nullptr;
  const ksj_input_item_view* calibration = nullptr;
  for (uint32_t b = 0; b < lay.batches; ++b) {
    batches[b] = ksj_input_batch_view{};
    batches[b].abi = make_header(sizeof(batches[b]));
    if (ksj_status st = f.cb->get_input_batch(f.cb->host_context, f.lease, b, &batches[b], f.error)) return st;
    const ksj_input_batch_view& batch = batches[b];
    if (!abi_ok(&batch) || batch.item_count != 1 || batch.items == nullptr || !abi_ok(&batch.items[0]))
      return reject(f.error, KSJ_STATUS_CONTRACT_VIOLATION, kPorts);
    const ksj_input_item_view** slot = batch.input_port == 0                         ? &dynamic
                                       : lay.batches == 2 && batch.input_port == 1 ? &calibration
                                                                                    : nullptr;
    if (slot == nullptr || *slot != nullptr) return reject(f.error, KSJ_STATUS_CONTRACT_VIOLATION, kPorts);
    *slot = &batch.items[0];
  }
  if (dynamic == nullptr || (lay.batches == 2 && calibration == nullptr))
    return reject(f.error, KSJ_STATUS_CONTRACT_VIOLATION, kPorts);

  if (lay.dynamic_bytes > kMaxKspaceBytes || lay.output_bytes > kMaxKspaceBytes ||
      !item_ok(f, *dynamic, lay.dynamic_bytes, lay.dynamic, device_edges))
    return reject(f.error, KSJ_STATUS_CONTRACT_VIOLATION, kDynamic);
  if (lay.batches == 2) {
    if (!item_ok(f, *calibration, lay.calibration_bytes, lay.calibration, device_edges))
      return reject(f.error, KSJ_STATUS_CONTRACT_VIOLATION, kCalibration);
    // A calibration key of zero is unkeyed and matches any k-space item (support/kspace_frame.cpp).
    if ((calibration->semantic_key_hash != 0 && calibration->semantic_key_hash != dynamic->semantic_key_hash) ||
        (calibration->order_key != 0 && calibration->order_key != dynamic->order_key))
      return reject(f.error, KSJ_STATUS_CONTRACT_VIOLATION, kIdentity);
  }
  if (info.reserved_output_bytes < lay.output_bytes) return reject(f.error, KSJ_STATUS_CONTRACT_VIOLATION, kLease);

  GrantGuard guard;
  ksj_mutable_payload_view out;
  if (ksj_status st = output_grant(f, lay.output_bytes, lay.output, kOutput, guard, out, device_edges)) return st;

  // Device-resident payloads are read and written in place; host ones go through the
  // session's staging.
  DeviceSession& d = device(f);
  const void* d_in = device_input(d, DeviceSession::kInput, dynamic->payload, lay.dynamic_bytes);
  const void* d_cal =
      lay.batches == 2 ? device_input(d, DeviceSession::kInput2, calibration->payload, lay.calibration_bytes) : nullptr;
  void* d_out = device_output(d, out, lay.output_bytes);
  const uint64_t frame = noncartesian(kind) ? uint64_t{s.samples} : uint64_t{s.rows} * s.cols;
  switch (kind) {
    case Kind::kNoise:
    case Kind::kNoncartesianNoise:
      check_gpu(ksj_gpu_channel_mix(d_cal, d_in, d_out, s.channels, s.channels, frame, d.stream()),
                "noise prewhitening");
      break;
    case Kind::kCoil:
    case Kind::kNoncartesianCoil:
      check_gpu(ksj_gpu_channel_mix(d_cal, d_in, d_out, s.out_channels, s.channels, frame, d.stream()),
                "coil compression");
      break;
    case Kind::kPhase:
      check_gpu(ksj_gpu_phase_correct(d_in, d_cal, d_out, s.channels, s.rows, s.cols, d.stream()),
                "phase correction");
      break;
    case Kind::kCrop:
      check_gpu(ksj_gpu_readout_crop(d_in, d_out, uint64_t{s.channels} * s.rows, s.cols, s.out_cols, s.offset,
                                     d.stream()),
                "readout crop");
      break;
  }
  finish_output(d, out, d_out, lay.output_bytes);
  return seal_and_finish(f, guard, out, lay.output_bytes, *dynamic, false, lay.batches, info.terminal_epoch);
}

// ---- configurations (canonical keys in sorted order; the parse_config of each
// src/operators/<operator>.cpp) ----

bool frame_dims_ok(const Shape& s) {
  return s.rows >= kMinDim && s.rows <= kMaxDim && s.cols >= kMinDim && s.cols <= kMaxDim;
}

// noise_prewhiten, phase_correct: {"channel_count":N,"cols":N,"rows":N}
bool configure_frame(const ksj_byte_view& c, Shape& s) {
  return parse_config(c, {u32_field("channel_count", 1, kMaxChannels, &s.channels),
                          u32_field("cols", kMinDim, kMaxDim, &s.cols), u32_field("rows", kMinDim, kMaxDim, &s.rows)});
}
bool valid_frame(const Shape& s) { return s.channels >= 1 && s.channels <= kMaxChannels && frame_dims_ok(s); }

// coil_compress: {"cols":N,"physical_channel_count":N,"rows":N,"virtual_channel_count":N}
bool configure_coil(const ksj_byte_view& c, Shape& s) {
  return parse_config(c, {u32_field("cols", kMinDim, kMaxDim, &s.cols),
                          u32_field("physical_channel_count", 1, kMaxChannels, &s.channels),
                          u32_field("rows", kMinDim, kMaxDim, &s.rows),
                          u32_field("virtual_channel_count", 1, kMaxChannels, &s.out_channels)}) &&
         s.out_channels <= s.channels;
}
bool valid_coil(const Shape& s) { return valid_frame(s) && s.out_channels >= 1 && s.out_channels <= s.channels; }

// readout_oversampling_remove:
// {"channel_count":N,"input_cols":N,"output_cols":N,"readout_offset":N,"rows":N}
bool configure_crop(const ksj_byte_view& c, Shape& s) {
  return parse_config(c, {u32_field("channel_count", 1, kMaxChannels, &s.channels),
                          u32_field("input_cols", kMinDim, kMaxDim, &s.cols),
                          u32_field("output_cols", kMinDim, kMaxDim, &s.out_cols),
                          u32_field("readout_offset", 0, UINT32_MAX, &s.offset),
                          u32_field("rows", kMinDim, kMaxDim, &s.rows)}) &&
         uint64_t{s.offset} + s.out_cols <= s.cols;
}
bool valid_crop(const Shape& s) {
  return valid_frame(s) && s.out_cols >= kMinDim && s.out_cols <= kMaxDim && uint64_t{s.offset} + s.out_cols <= s.cols;
}

// noncartesian_noise_prewhiten: {"channel_count":N,"sample_count":N}
bool configure_nc_noise(const ksj_byte_view& c, Shape& s) {
  return parse_config(c, {u32_field("channel_count", 1, kMaxChannels, &s.channels),
                          u32_field("sample_count", 1, kMaxSamples, &s.samples)});
}
bool valid_nc(const Shape& s) {
  return s.channels >= 1 && s.channels <= kMaxChannels && s.samples >= 1 && s.samples <= kMaxSamples;
}

// noncartesian_coil_compress: {"physical_channel_count":N,"sample_count":N,"virtual_channel_count":N}
bool configure_nc_coil(const ksj_byte_view& c, Shape& s) {
  return parse_config(c, {u32_field("physical_channel_count", 1, kMaxChannels, &s.channels),
                          u32_field("sample_count", 1, kMaxSamples, &s.samples),
                          u32_field("virtual_channel_count", 1, kMaxChannels, &s.out_channels)}) &&
         s.out_channels <= s.channels;
}
bool valid_nc_coil(const Shape& s) { return valid_nc(s) && s.out_channels >= 1 && s.out_channels <= s.channels; }

template <Kind K, bool DeviceEdges> ksj_status process_as(Firing& f) { return process(f, K, DeviceEdges); }

constexpr std::string_view kNoiseConfig =
    "noise_prewhiten needs the canonical config {\"channel_count\":N,\"cols\":N,\"rows\":N} with "
    "1 <= channel_count <= 64 and 2 <= rows, cols <= 512";
constexpr std::string_view kPhaseConfig =
    "phase_correct needs the canonical config {\"channel_count\":N,\"cols\":N,\"rows\":N} with "
    "1 <= channel_count <= 64 and 2 <= rows, cols <= 512";
constexpr std::string_view kCoilConfig =
    "coil_compress needs the canonical config {\"cols\":N,\"physical_channel_count\":N,\"rows\":N,"
    "\"virtual_channel_count\":N} with 1 <= virtual <= physical <= 64 and 2 <= rows, cols <= 512";
constexpr std::string_view kCropConfig =
    "readout_oversampling_remove needs the canonical config {\"channel_count\":N,\"input_cols\":N,"
    "\"output_cols\":N,\"readout_offset\":N,\"rows\":N} with 1 <= channel_count <= 64, 2 <= rows, input_cols, "
    "output_cols <= 512 and readout_offset + output_cols <= input_cols";
constexpr std::string_view kNcNoiseConfig =
    "noncartesian_noise_prewhiten needs the canonical config {\"channel_count\":N,\"sample_count\":N} with "
    "1 <= channel_count <= 64 and 1 <= sample_count <= 65536";
constexpr std::string_view kNcCoilConfig =
    "noncartesian_coil_compress needs the canonical config {\"physical_channel_count\":N,\"sample_count\":N,"
    "\"virtual_channel_count\":N} with 1 <= virtual <= physical <= 64 and 1 <= sample_count <= 65536";

}  // namespace

// org.kspacejet.kspace-conditioning: the reference's host-pageable contract.
const OpSpec kNoisePrewhiten{"noise_prewhiten", kNoiseConfig, 2, kMaxKspaceBytes, 0, &configure_frame, &valid_frame,
                             &process_as<Kind::kNoise, false>};
const OpSpec kPhaseCorrect{"phase_correct", kPhaseConfig, 2, kMaxKspaceBytes, 0, &configure_frame, &valid_frame,
                           &process_as<Kind::kPhase, false>};
const OpSpec kCoilCompress{"coil_compress", kCoilConfig, 2, kMaxKspaceBytes, 0, &configure_coil, &valid_coil,
                           &process_as<Kind::kCoil, false>};
const OpSpec kReadoutOversamplingRemove{"readout_oversampling_remove", kCropConfig, 1, kMaxKspaceBytes, 0,
                                        &configure_crop, &valid_crop, &process_as<Kind::kCrop, false>};
const OpSpec kNoncartesianNoisePrewhiten{"noncartesian_noise_prewhiten", kNcNoiseConfig, 2, kMaxKspaceBytes, 0,
                                         &configure_nc_noise, &valid_nc,
                                         &process_as<Kind::kNoncartesianNoise, false>};
const OpSpec kNoncartesianCoilCompress{"noncartesian_coil_compress", kNcCoilConfig, 2, kMaxKspaceBytes, 0,
                                       &configure_nc_coil, &valid_nc_coil,
                                       &process_as<Kind::kNoncartesianCoil, false>};

// org.kspacejet.gpu-recon: the same operators with device-memory edges as well.
const OpSpec kGpuNoisePrewhiten{"noise_prewhiten", kNoiseConfig, 2, kMaxKspaceBytes, 0, &configure_frame,
                                &valid_frame, &process_as<Kind::kNoise, true>};
const OpSpec kGpuPhaseCorrect{"phase_correct", kPhaseConfig, 2, kMaxKspaceBytes, 0, &configure_frame, &valid_frame,
                              &process_as<Kind::kPhase, true>};
const OpSpec kGpuCoilCompress{"coil_compress", kCoilConfig, 2, kMaxKspaceBytes, 0, &configure_coil, &valid_coil,
                              &process_as<Kind::kCoil, true>};
const OpSpec kGpuReadoutOversamplingRemove{"readout_oversampling_remove", kCropConfig, 1, kMaxKspaceBytes, 0,
                                           &configure_crop, &valid_crop, &process_as<Kind::kCrop, true>};
const OpSpec kGpuNoncartesianNoisePrewhiten{"noncartesian_noise_prewhiten", kNcNoiseConfig, 2, kMaxKspaceBytes, 0,
                                            &configure_nc_noise, &valid_nc,
                                            &process_as<Kind::kNoncartesianNoise, true>};
const OpSpec kGpuNoncartesianCoilCompress{"noncartesian_coil_compress", kNcCoilConfig, 2, kMaxKspaceBytes, 0,
                                          &configure_nc_coil, &valid_nc_coil,
                                          &process_as<Kind::kNoncartesianCoil, true>};

}  // namespace ksj::prov
