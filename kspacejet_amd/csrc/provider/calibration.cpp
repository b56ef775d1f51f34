// The operators of org.kspacejet.calibration on the GPU: noise-model, readout phase-model
// and coil-compression-basis estimation from one calibration payload.
//
// The firing follows the reference's shared flow
// (providers/kspacejet-calibration/src/provider_api.cpp:269-370): usable callbacks -> one
// input batch, one output grant and exactly the operator's configured host scratch
// reservation (validated like the reference, though the GPU path does not use it) -> the
// input item on port 0 (host-pageable, 64-byte aligned, no metadata, registry type, a byte
// count that derives the sample or reference-line count) -> reserved output bytes ->
// output grant -> estimate -> seal with the input's keys and (empty) metadata.  An
// estimate that fails (a non-finite sample, no convergence) is a contract violation with
// the grant released, as there.
#include "ops.hpp"

namespace ksj::prov {

namespace {

constexpr uint64_t kCf32 = 8;
constexpr uint32_t kMaxChannels = 64;         // calibration_payload.hpp:18
constexpr uint32_t kMaxReadout = 4096;        // calibration_payload.hpp:19
constexpr uint64_t kMaxInputBytes = 64ull << 20;  // calibration_payload.hpp:21
constexpr uint64_t kMaxOutputBytes = uint64_t{kMaxChannels} * kMaxReadout * kCf32;  // :22-23 (2 MiB)

constexpr std::string_view kBadAbi = "Calibration Provider received incompatible ABI storage";
constexpr std::string_view kLease =
    "Calibration Provider requires one input, one output grant, and its exact configured host scratch reservation";
constexpr std::string_view kScratch = "Calibration Provider received an incompatible caller-owned scratch view";
constexpr std::string_view kInput = "Calibration Provider received an incompatible calibration payload or byte shape";
constexpr std::string_view kOutput = "Calibration Provider received an incompatible calibration output grant";
constexpr std::string_view kEstimate = "Calibration Provider could not estimate the requested calibration artifact";

enum class Kind { kNoise, kPhase, kCoil };

// Host scratch the reference reserves per firing (noise_model_estimate.cpp:119-139,
// coil_compression_basis_estimate.cpp:145-161; phase estimation none).
uint64_t scratch_bytes(Kind k, const Shape& s) {
  const uint64_t n = s.channels;
  switch (k) {
    case Kind::kNoise: return 3 * n * n * kCf32 + n * (kCf32 + sizeof(float));
    case Kind::kCoil: return 3 * n * n * kCf32 + n * sizeof(float);
    case Kind::kPhase: return 0;
  }
  return 0;
}

TypeId input_type(Kind k) {
  return k == Kind::kNoise ? TypeId::kNoiseCalibrationFrame
                           : k == Kind::kPhase ? TypeId::kPhaseReferenceFrame : TypeId::kKspaceFrame;
}
TypeId output_type(Kind k) {
  return k == Kind::kNoise ? TypeId::kNoiseModel : k == Kind::kPhase ? TypeId::kPhaseModel : TypeId::kCoilCompressionBasis;
}

// Samples per channel (noise, coil) or reference lines (phase) from the payload byte
// count; 0 when the count does not fit (has_compatible_input_byte_count of each operator).
uint64_t derived_count(Kind k, const Shape& s, uint64_t bytes) {
  const uint64_t unit = k == Kind::kPhase ? uint64_t{s.channels} * s.cols * kCf32 : uint64_t{s.channels} * kCf32;
  if (unit == 0 || bytes > kMaxInputBytes || bytes % unit != 0) return 0;
  const uint64_t n = bytes / unit;
  return n >= (k == Kind::kNoise ? 2u : 1u) ? n : 0;
}

uint64_t output_bytes(Kind k, const Shape& s) {
  switch (k) {
    case Kind::kNoise: return uint64_t{s.channels} * s.channels * kCf32;
    case Kind::kPhase: return uint64_t{s.channels} * s.cols * kCf32;
    case Kind::kCoil: return uint64_t{s.out_channels} * s.channels * kCf32;
  }
  return 0;
}

ksj_status process(Firing& f, Kind kind) {
  if (!callbacks_ok(f.cb, KSJ_LEASE_CAP_INPUT_BATCHES | KSJ_LEASE_CAP_OUTPUT_GRANTS))
    return reject(f.error, KSJ_STATUS_BAD_ABI, kBadAbi);
  const Shape& s = f.op.shape;
  const uint64_t scratch = scratch_bytes(kind, s);

  ksj_firing_lease_info info{};
  info.abi = make_header(sizeof(info));
  if (ksj_status st = f.cb->get_info(f.cb->host_context, f.lease, &info, f.error)) return st;
  if (!abi_ok(&info) || info.input_batch_count != 1 || info.output_grant_count != 1 ||
      info.reserved_scratch_bytes != scratch)
    return reject(f.error, KSJ_STATUS_CONTRACT_VIOLATION, kLease);
  if (scratch != 0) {
    if ((f.cb->abi.capability_bits & KSJ_LEASE_CAP_SCRATCH) == 0 || f.cb->get_scratch == nullptr)
      return reject(f.error, KSJ_STATUS_CONTRACT_VIOLATION, kScratch);
    ksj_scratch_view sv{};
    sv.abi = make_header(sizeof(sv));
    if (ksj_status st = f.cb->get_scratch(f.cb->host_context, f.lease, &sv, f.error)) return st;
    if (!abi_ok(&sv) || sv.data == nullptr || sv.byte_count != scratch ||
        sv.memory_domain != KSJ_PROVIDER_MEMORY_HOST_PAGEABLE || !host_memory_ok(sv.data, sv.alignment, 64))
      return reject(f.error, KSJ_STATUS_CONTRACT_VIOLATION, kScratch);
  }

  ksj_input_batch_view batch{};
  batch.abi = make_header(sizeof(batch));
  if (ksj_status st = f.cb->get_input_batch(f.cb->host_context, f.lease, 0, &batch, f.error)) return st;
  if (!single_item_batch(batch, 0)) return reject(f.error, KSJ_STATUS_CONTRACT_VIOLATION, kInput);
  const ksj_input_item_view& item = batch.items[0];
  const uint64_t in_bytes = item.payload.byte_count;
  const uint64_t count = derived_count(kind, s, in_bytes);
  if (count == 0 || !input_payload_ok(item, in_bytes, input_type(kind)) || item.metadata.size != 0)
    return reject(f.error, KSJ_STATUS_CONTRACT_VIOLATION, kInput);

  const uint64_t out_bytes = output_bytes(kind, s);
  if (out_bytes > kMaxOutputBytes || info.reserved_output_bytes < out_bytes)
    return reject(f.error, KSJ_STATUS_CONTRACT_VIOLATION, kLease);
  GrantGuard guard;
  ksj_mutable_payload_view out;
  if (ksj_status st = output_grant(f, out_bytes, output_type(kind), kOutput, guard, out)) return st;

  DeviceSession& d = device(f);
  const void* d_in = device_input(d, DeviceSession::kInput, item.payload, in_bytes);
  void* d_out = device_output(d, out, out_bytes);
  int rc = KSJ_GPU_OK;
  switch (kind) {
    case Kind::kNoise: {
      const size_t ws = ksj_gpu_calibration_workspace_bytes(s.channels, count);
      rc = ksj_gpu_noise_model_estimate(d_in, s.channels, count, 1.0e-6f, d_out,
                                        d.buffer(DeviceSession::kWorkspace, ws), ws, d.stream());
      break;
    }
    case Kind::kCoil: {
      const size_t ws = ksj_gpu_calibration_workspace_bytes(s.channels, count);
      rc = ksj_gpu_coil_compression_basis_estimate(d_in, s.channels, s.out_channels, count, d_out,
                                                   d.buffer(DeviceSession::kWorkspace, ws), ws, d.stream());
      break;
    }
    case Kind::kPhase:
      rc = ksj_gpu_phase_correction_estimate(d_in, s.channels, static_cast<uint32_t>(count), s.cols, d_out,
                                             d.stream());
      break;
  }
  // A sample the estimate cannot use is the reference's "could not estimate"; anything
  // else is a device failure.
  if (rc == KSJ_GPU_INVALID_ARGUMENT) return reject(f.error, KSJ_STATUS_CONTRACT_VIOLATION, kEstimate);
  check_gpu(rc, "calibration estimate");
  finish_output(d, out, d_out, out_bytes);
  return seal_and_finish(f, guard, out, out_bytes, item, true, 1, info.terminal_epoch);
}

// ---- configurations (noise_model_estimate.cpp:47-64, phase_correction_estimate.cpp:40-63,
// coil_compression_basis_estimate.cpp:49-75) ----

// {"channel_count":N}
bool configure_noise(const ksj_byte_view& c, Shape& s) {
  return parse_config(c, {u32_field("channel_count", 1, kMaxChannels, &s.channels)});
}
bool valid_noise(const Shape& s) { return s.channels >= 1 && s.channels <= kMaxChannels; }

// {"channel_count":N,"readout_sample_count":K}
bool configure_phase(const ksj_byte_view& c, Shape& s) {
  return parse_config(c, {u32_field("channel_count", 1, kMaxChannels, &s.channels),
                          u32_field("readout_sample_count", 1, kMaxReadout, &s.cols)});
}
bool valid_phase(const Shape& s) { return valid_noise(s) && s.cols >= 1 && s.cols <= kMaxReadout; }

// {"physical_channel_count":P,"virtual_channel_count":V}
bool configure_coil(const ksj_byte_view& c, Shape& s) {
  return parse_config(c, {u32_field("physical_channel_count", 1, kMaxChannels, &s.channels),
                          u32_field("virtual_channel_count", 1, kMaxChannels, &s.out_channels)}) &&
         s.out_channels <= s.channels;
}
bool valid_coil(const Shape& s) { return valid_noise(s) && s.out_channels >= 1 && s.out_channels <= s.channels; }

template <Kind K> ksj_status process_as(Firing& f) { return process(f, K); }

constexpr std::string_view kNoiseConfig =
    "Calibration Provider noise_model_estimate requires canonical {\"channel_count\":N}, where 1 <= N <= 64";
constexpr std::string_view kPhaseConfig =
    "Calibration Provider phase_correction_estimate requires canonical "
    "{\"channel_count\":N,\"readout_sample_count\":K}, where 1 <= N <= 64 and 1 <= K <= 4096";
constexpr std::string_view kCoilConfig =
    "Calibration Provider coil_compression_basis_estimate requires canonical "
    "{\"physical_channel_count\":P,\"virtual_channel_count\":V}, where 1 <= V <= P <= 64";

// Maximum host scratch of the descriptors (the kMaximumScratchBytes of the operators).
constexpr uint64_t kNoiseScratchMax = 3ull * kMaxChannels * kMaxChannels * kCf32 + kMaxChannels * (kCf32 + 4);
constexpr uint64_t kCoilScratchMax = 3ull * kMaxChannels * kMaxChannels * kCf32 + kMaxChannels * 4;

}  // namespace

const OpSpec kNoiseModelEstimate{"noise_model_estimate", kNoiseConfig, 1, kMaxOutputBytes, kNoiseScratchMax,
                                 &configure_noise, &valid_noise, &process_as<Kind::kNoise>};
const OpSpec kPhaseCorrectionEstimate{"phase_correction_estimate", kPhaseConfig, 1, kMaxOutputBytes, 0,
                                      &configure_phase, &valid_phase, &process_as<Kind::kPhase>};
const OpSpec kCoilCompressionBasisEstimate{"coil_compression_basis_estimate", kCoilConfig, 1, kMaxOutputBytes,
                                           kCoilScratchMax, &configure_coil, &valid_coil, &process_as<Kind::kCoil>};

}  // namespace ksj::prov
