// Operator framework of the MI355X provider modules: the opaque handle types,
// the per-operator specification, the generic lifecycle table, and the firing
// helpers (lease info, port resolution, payload checks, output grants) that
// each operator's process function composes in the reference's validation
// order: lease -> inputs -> [scratch] -> output grant -> compute -> seal.
#pragma once

#include "abi_support.hpp"
#include "device.hpp"

#include <memory>
#include <string_view>

namespace ksj::prov {
struct OpSpec;
}

struct ksj_execution_context {
  ksj_provider_operator* owner = nullptr;
};

struct ksj_key_state {
  ksj_provider_operator* owner = nullptr;
};

namespace ksj::prov {

// Configuration of one operator instance, parsed from its canonical config.
struct Shape {
  uint32_t channels = 0;
  uint32_t rows = 0, cols = 0;  // image (Cartesian: k-space frame) dimensions
  uint32_t samples = 0;
  uint32_t kernel_width = 0;
  uint32_t oversampling = 0;
  uint32_t centred = 0;
  uint32_t out_channels = 0;  // k-space conditioning: virtual channels of a coil compression
  uint32_t out_cols = 0;      // k-space conditioning: readout columns kept by the crop, and
  uint32_t offset = 0;        // the first of them
  uint32_t window = 0;        // sensitivity maps: neighbourhood side (odd)
  uint32_t iterations = 0;    // sensitivity maps: power-iteration steps
  uint32_t method = 0;        // sensitivity maps: index of the method literal
};

}  // namespace ksj::prov

struct ksj_provider_operator {
  const ksj::prov::OpSpec* spec = nullptr;
  ksj::prov::Shape shape;
  ksj_execution_context context;
  ksj_key_state key_state;
  bool context_active = false;
  bool key_state_active = false;
  uint32_t device_ordinal = 0;
  std::unique_ptr<ksj::prov::DeviceSession> device;  // created by the first firing that computes
};

namespace ksj::prov {

struct Firing {
  ksj_provider_operator& op;
  ksj_firing_lease* lease;
  const ksj_firing_lease_callbacks* cb;
  ksj_process_result* result;
  ksj_error_view* error;
};

struct OpSpec {
  std::string_view id;
  std::string_view config_error;  // message of an INVALID_ARGUMENT configuration rejection
  uint32_t max_input_items;
  uint64_t max_output_bytes;
  uint64_t max_scratch_bytes;
  bool (*configure)(const ksj_byte_view& config, Shape& shape);
  bool (*valid)(const Shape& shape);
  ksj_status (*process)(Firing& f);
};

struct ProviderSpec {
  std::string_view provider_id;
  std::string_view bundle_digest_hex;  // SHA-256 of the NUL-delimited bundle identity fields
  const OpSpec* const* ops;
  uint32_t op_count;
  std::string_view bad_abi;           // message texts of the lifecycle rejections
  std::string_view invalid_argument;
  std::string_view unsupported_operator;
  std::string_view internal_error;
};

// Defined once per provider module (mod_*.cpp).
const ProviderSpec& provider_spec();

// ---- firing helpers ----

// get_info + the counts and reservations the operator needs.  Returns the
// host's status unchanged when get_info fails, else CONTRACT_VIOLATION with
// `lease_error` when the lease does not fit.
ksj_status lease_info(Firing& f, uint32_t batches, uint64_t output_bytes, uint64_t scratch_bytes, bool check_scratch,
                      std::string_view lease_error, ksj_firing_lease_info& info);

// Exactly one item on the expected port with compatible headers.
bool single_item_batch(const ksj_input_batch_view& batch, uint32_t port);

// Payload of one input item: exact bytes, host-pageable, 64-byte aligned,
// valid metadata view and the exact registry type.
bool input_payload_ok(const ksj_input_item_view& item, uint64_t bytes, TypeId type);
// As input_payload_ok, also accepting a KSJ_PROVIDER_MEMORY_DEVICE payload that lies in
// device memory of GPU `device_ordinal` (64-byte aligned): the device edge of the
// operators that opt in.
bool input_payload_ok_device(const ksj_input_item_view& item, uint64_t bytes, TypeId type, uint32_t device_ordinal);
bool device_memory_ok(const void* p, uint32_t declared_alignment, uint32_t device_ordinal);

// Two input batches resolved by input_port (0 and 1), not by batch index.
// On success `first`/`second` point into `storage`.
ksj_status two_port_batches(Firing& f, ksj_input_batch_view storage[2], const ksj_input_batch_view*& port0,
                            const ksj_input_batch_view*& port1, std::string_view lease_error,
                            std::string_view port0_error, std::string_view port1_error);

// Lease scratch with at least `bytes` bytes, host-pageable and `alignment`-aligned.
ksj_status scratch_ok(Firing& f, uint64_t bytes, uint32_t alignment, std::string_view scratch_error);

// Acquires and maps the output grant of port 0 and validates the mapping.  On
// failure a mapped-but-invalid grant is released by `guard`.
ksj_status output_grant(Firing& f, uint64_t bytes, TypeId type, std::string_view output_error, GrantGuard& guard,
                        ksj_mutable_payload_view& out, bool allow_device = false);

// Seals the output (one item, exact bytes, identity keys of `key_item`,
// metadata either copied from `key_item` or empty) and writes DONE.
ksj_status seal_and_finish(Firing& f, GrantGuard& guard, const ksj_mutable_payload_view& out, uint64_t bytes,
                           const ksj_input_item_view& key_item, bool copy_metadata, uint64_t consumed,
                           uint64_t terminal_epoch);

// The operator's device session, created on first use on the context's device.
DeviceSession& device(Firing& f);

// ---- device-memory edges (KSJ_PROVIDER_MEMORY_DEVICE) of the operators that accept them ----

inline bool on_device(const ksj_payload_view& p) { return p.memory_domain == KSJ_PROVIDER_MEMORY_DEVICE; }
inline bool on_device(const ksj_mutable_payload_view& p) { return p.memory_domain == KSJ_PROVIDER_MEMORY_DEVICE; }
// Device pointer of an input payload: in place when device-resident, else staged into `slot`.
const void* device_input(DeviceSession& d, DeviceSession::Slot slot, const ksj_payload_view& p, uint64_t bytes);
// Device buffer the operator writes its output to (the grant itself when device-resident),
// and the hand-over once it is computed: a stream synchronisation for a device grant, a
// download into a host grant.
void* device_output(DeviceSession& d, const ksj_mutable_payload_view& out, uint64_t bytes);
void finish_output(DeviceSession& d, const ksj_mutable_payload_view& out, const void* d_out, uint64_t bytes);

}  // namespace ksj::prov
