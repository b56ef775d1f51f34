// Generic lifecycle of the provider modules and the firing helpers.  The
// lifecycle checks, their status codes and their order are those of the
// reference providers (providers/kspacejet-noncartesian-recon/src/provider_api.cpp:146-325),
// so a host sees the same behaviour from this module as from the CPU one.
#include "operator.hpp"

#include <hip/hip_runtime_api.h>

#include <new>
#include <string>
#include <vector>

namespace ksj::prov {

namespace {

thread_local std::string t_message;  // text of a dynamic error, valid until the next call on this thread

const ProviderSpec& spec() { return provider_spec(); }

const OpSpec* find_op(const ksj_utf8_view& id) {
  for (uint32_t i = 0; i < spec().op_count; ++i)
    if (text_is(id, spec().ops[i]->id)) return spec().ops[i];
  return nullptr;
}

bool registered(const OpSpec* op) {
  for (uint32_t i = 0; i < spec().op_count; ++i)
    if (spec().ops[i] == op) return true;
  return false;
}

bool is_operator(const ksj_provider_operator* h) { return h != nullptr && registered(h->spec) && h->spec->valid(h->shape); }

bool is_context(const ksj_provider_operator* h, const ksj_execution_context* c) {
  return is_operator(h) && c == &h->context && c->owner == h && h->context_active;
}

bool is_key_state(const ksj_provider_operator* h, const ksj_key_state* k) {
  return is_operator(h) && k == &h->key_state && k->owner == h && h->key_state_active;
}

ksj_operator_descriptor make_operator_descriptor(const OpSpec& op) {
  ksj_operator_descriptor d{};
  d.abi = make_header(sizeof(d), KSJ_OPERATOR_CAP_CANCEL_NO_ALLOCATION | KSJ_OPERATOR_CAP_CANCEL_NO_THROW);
  d.operator_id = utf8(op.id);
  d.max_in_flight = 1;
  d.thread_safety = KSJ_PROVIDER_SERIAL_INSTANCE;
  d.max_private_threads = 0;
  d.max_input_items_per_firing = op.max_input_items;
  d.max_output_items_per_firing = 1;
  d.max_output_bytes_per_firing = op.max_output_bytes;
  d.max_scratch_bytes_per_firing = op.max_scratch_bytes;
  return d;
}

// Descriptors of every operator of the module, built once; the host reads
// descriptor.operators[0, operator_count).
struct Metadata {
  std::vector<ksj_operator_descriptor> ops;
  ksj_provider_descriptor descriptor{};
  Metadata() : ops(spec().op_count) {
    for (uint32_t i = 0; i < spec().op_count; ++i) ops[i] = make_operator_descriptor(*spec().ops[i]);
    descriptor.abi = make_header(sizeof(descriptor), KSJ_PROVIDER_CAP_SYNC_PROCESS | KSJ_PROVIDER_CAP_NO_PRIVATE_THREADS |
                                                         KSJ_PROVIDER_CAP_NO_DIRECT_FILE_IO |
                                                         KSJ_PROVIDER_CAP_NO_DIRECT_NETWORK_IO);
    descriptor.provider_id = utf8(spec().provider_id);
    descriptor.bundle_digest = digest_from_hex(spec().bundle_digest_hex);
    descriptor.operator_count = static_cast<uint32_t>(ops.size());
    descriptor.operators = ops.data();
  }
};

const Metadata& metadata() {
  static const Metadata m;
  return m;
}

ksj_status device_failure(ksj_error_view* e, const DeviceFailure& failure) {
  t_message = failure.what();
  return reject(e, failure.exhausted ? KSJ_STATUS_RESOURCE_EXHAUSTED : KSJ_STATUS_INTERNAL_ERROR, t_message);
}

// ---- the ten lifecycle entries ----

ksj_status operator_create(const ksj_operator_create_request* request, ksj_provider_operator** out_operator,
                           ksj_error_view* out_error) noexcept {
  try {
    if (!abi_ok(request) || out_operator == nullptr)
      return reject(out_error, KSJ_STATUS_INVALID_ARGUMENT, spec().invalid_argument);
    *out_operator = nullptr;
    const OpSpec* op = find_op(request->operator_id);
    if (op == nullptr) return reject(out_error, KSJ_STATUS_UNSUPPORTED, spec().unsupported_operator);
    if (request->host_services != nullptr && !abi_ok(request->host_services))
      return reject(out_error, KSJ_STATUS_BAD_ABI, spec().bad_abi);
    Shape shape;
    if (!op->configure(request->canonical_config, shape))
      return reject(out_error, KSJ_STATUS_INVALID_ARGUMENT, op->config_error);
    if (!op->valid(shape)) return reject(out_error, KSJ_STATUS_INTERNAL_ERROR, spec().internal_error);
    auto* h = new (std::nothrow) ksj_provider_operator{};
    if (h == nullptr) return reject(out_error, KSJ_STATUS_RESOURCE_EXHAUSTED, spec().internal_error);
    h->spec = op;
    h->shape = shape;
    *out_operator = h;
    return KSJ_STATUS_OK;
  } catch (...) {
    return reject(out_error, KSJ_STATUS_INTERNAL_ERROR, spec().internal_error);
  }
}

ksj_status execution_context_create(ksj_provider_operator* h, const ksj_execution_context_descriptor* d,
                                    ksj_execution_context** out_context, ksj_error_view* out_error) noexcept {
  if (!is_operator(h) || !abi_ok(d) || out_context == nullptr)
    return reject(out_error, KSJ_STATUS_INVALID_ARGUMENT, spec().invalid_argument);
  if (d->host_services != nullptr && !abi_ok(d->host_services))
    return reject(out_error, KSJ_STATUS_BAD_ABI, spec().bad_abi);
  if (h->context_active) return reject(out_error, KSJ_STATUS_FAILED_PRECONDITION, spec().invalid_argument);
  h->context.owner = h;
  h->context_active = true;
  h->device_ordinal = d->device_ordinal;
  *out_context = &h->context;
  return KSJ_STATUS_OK;
}

ksj_status key_state_init(ksj_provider_operator* h, ksj_execution_context* c, const ksj_key_state_descriptor* d,
                          ksj_key_state** out_key_state, ksj_error_view* out_error) noexcept {
  if (!is_context(h, c) || !abi_ok(d) || out_key_state == nullptr || !bytes_ok(d->semantic_key))
    return reject(out_error, KSJ_STATUS_INVALID_ARGUMENT, spec().invalid_argument);
  if (h->key_state_active) return reject(out_error, KSJ_STATUS_FAILED_PRECONDITION, spec().invalid_argument);
  h->key_state.owner = h;
  h->key_state_active = true;
  *out_key_state = &h->key_state;
  return KSJ_STATUS_OK;
}

ksj_status operator_on_start(ksj_provider_operator* h, ksj_execution_context* c, ksj_key_state* k,
                             const ksj_scan_start_descriptor* d, ksj_error_view* out_error) noexcept {
  if (!is_context(h, c) || !is_key_state(h, k) || !abi_ok(d))
    return reject(out_error, KSJ_STATUS_INVALID_ARGUMENT, spec().invalid_argument);
  return KSJ_STATUS_OK;
}

ksj_status operator_process_batch(ksj_provider_operator* h, ksj_execution_context* c, ksj_key_state* k,
                                  ksj_firing_lease* lease, const ksj_firing_lease_callbacks* cb,
                                  ksj_process_result* out_result, ksj_error_view* out_error) noexcept {
  if (!is_context(h, c) || !is_key_state(h, k) || lease == nullptr || !abi_ok(out_result))
    return reject(out_error, KSJ_STATUS_INVALID_ARGUMENT, spec().invalid_argument);
  try {
    Firing f{*h, lease, cb, out_result, out_error};
    return h->spec->process(f);
  } catch (const DeviceFailure& e) {
    return device_failure(out_error, e);
  } catch (const std::bad_alloc&) {
    return reject(out_error, KSJ_STATUS_RESOURCE_EXHAUSTED, spec().internal_error);
  } catch (...) {
    return reject(out_error, KSJ_STATUS_INTERNAL_ERROR, spec().internal_error);
  }
}

ksj_status operator_on_scan_end(ksj_provider_operator* h, ksj_execution_context* c, ksj_key_state* k,
                                const ksj_scan_end_descriptor* d, ksj_firing_lease* terminal_lease,
                                const ksj_firing_lease_callbacks* cb, ksj_process_result* out_result,
                                ksj_error_view* out_error) noexcept {
  if (!is_context(h, c) || !is_key_state(h, k) || !abi_ok(d) || d->kind != KSJ_PROVIDER_SCAN_END_NORMAL ||
      d->reserved0 != 0 || terminal_lease == nullptr || !abi_ok(out_result) || !abi_ok(cb))
    return reject(out_error, KSJ_STATUS_INVALID_ARGUMENT, spec().invalid_argument);
  write_done(out_result, 0, 0, d->terminal_epoch);
  return KSJ_STATUS_OK;
}

ksj_status operator_on_cancel(ksj_provider_operator* h, ksj_execution_context* c, ksj_key_state* k,
                              const ksj_cancel_context* d, ksj_error_view* out_error) noexcept {
  if (!is_context(h, c) || !is_key_state(h, k) || !abi_ok(d) || d->reserved0 != 0 ||
      (d->kind != KSJ_PROVIDER_SCAN_END_CANCELLED && d->kind != KSJ_PROVIDER_SCAN_END_FAILED))
    return reject(out_error, KSJ_STATUS_INVALID_ARGUMENT, spec().invalid_argument);
  return KSJ_STATUS_OK;  // every firing completes synchronously: no device work is left to stop
}

void key_state_reset(ksj_provider_operator* h, ksj_execution_context* c, ksj_key_state* k) noexcept {
  if (!is_context(h, c) || !is_key_state(h, k)) return;
  h->key_state.owner = nullptr;
  h->key_state_active = false;
}

void execution_context_destroy(ksj_provider_operator* h, ksj_execution_context* c) noexcept {
  if (!is_context(h, c)) return;
  h->device.reset();  // device buffers and the resident plan belong to the context's device
  h->context.owner = nullptr;
  h->context_active = false;
}

void operator_destroy(ksj_provider_operator* h) noexcept { delete h; }

}  // namespace

// ---- firing helpers ----

ksj_status lease_info(Firing& f, uint32_t batches, uint64_t output_bytes, uint64_t scratch_bytes, bool check_scratch,
                      std::string_view lease_error, ksj_firing_lease_info& info) {
  info = ksj_firing_lease_info{};
  info.abi = make_header(sizeof(info));
  const ksj_status st = f.cb->get_info(f.cb->host_context, f.lease, &info, f.error);
  if (st != KSJ_STATUS_OK) return st;
  if (!abi_ok(&info) || info.input_batch_count != batches || info.output_grant_count != 1 ||
      info.reserved_output_bytes < output_bytes || (check_scratch && info.reserved_scratch_bytes < scratch_bytes))
    return reject(f.error, KSJ_STATUS_CONTRACT_VIOLATION, lease_error);
  return KSJ_STATUS_OK;
}

bool single_item_batch(const ksj_input_batch_view& batch, uint32_t port) {
  return abi_ok(&batch) && batch.input_port == port && batch.item_count == 1 && batch.items != nullptr &&
         abi_ok(&batch.items[0]);
}

bool input_payload_ok(const ksj_input_item_view& item, uint64_t bytes, TypeId type) {
  const ksj_payload_view& p = item.payload;
  return abi_ok(&p) && p.data != nullptr && p.byte_count == bytes && p.memory_domain == KSJ_PROVIDER_MEMORY_HOST_PAGEABLE &&
         host_memory_ok(p.data, p.alignment, 64) && bytes_ok(item.metadata) && type_matches(p.type, type);
}

bool device_memory_ok(const void* p, uint32_t declared_alignment, uint32_t device_ordinal) {
  if (!host_memory_ok(p, declared_alignment, 64)) return false;  // the same pointer and alignment rules
  hipPointerAttribute_t attr{};
  if (hipPointerGetAttributes(&attr, p) != hipSuccess) {
    (void)hipGetLastError();  // an unknown pointer is a rejection, not a sticky error
    return false;
  }
  return attr.type == hipMemoryTypeDevice && attr.device == static_cast<int>(device_ordinal);
}

bool input_payload_ok_device(const ksj_input_item_view& item, uint64_t bytes, TypeId type, uint32_t device_ordinal) {
  const ksj_payload_view& p = item.payload;
  if (!abi_ok(&p) || p.memory_domain != KSJ_PROVIDER_MEMORY_DEVICE) return input_payload_ok(item, bytes, type);
  return p.data != nullptr && p.byte_count == bytes && device_memory_ok(p.data, p.alignment, device_ordinal) &&
         bytes_ok(item.metadata) && type_matches(p.type, type);
}

ksj_status two_port_batches(Firing& f, ksj_input_batch_view storage[2], const ksj_input_batch_view*& port0,
                            const ksj_input_batch_view*& port1, std::string_view lease_error,
                            std::string_view port0_error, std::string_view port1_error) {
  for (uint32_t i = 0; i < 2; ++i) {
    storage[i] = ksj_input_batch_view{};
    storage[i].abi = make_header(sizeof(storage[i]));
    const ksj_status st = f.cb->get_input_batch(f.cb->host_context, f.lease, i, &storage[i], f.error);
    if (st != KSJ_STATUS_OK) return st;
  }
  if (!abi_ok(&storage[0]) || !abi_ok(&storage[1])) return reject(f.error, KSJ_STATUS_CONTRACT_VIOLATION, lease_error);
  port0 = port1 = nullptr;
  for (uint32_t i = 0; i < 2; ++i) {
    if (storage[i].input_port == 0) {
      if (port0) return reject(f.error, KSJ_STATUS_CONTRACT_VIOLATION, port0_error);
      port0 = &storage[i];
    } else if (storage[i].input_port == 1) {
      if (port1) return reject(f.error, KSJ_STATUS_CONTRACT_VIOLATION, port1_error);
      port1 = &storage[i];
    } else {
      return reject(f.error, KSJ_STATUS_CONTRACT_VIOLATION, lease_error);
    }
  }
  if (port0 == nullptr || !single_item_batch(*port0, 0))
    return reject(f.error, KSJ_STATUS_CONTRACT_VIOLATION, port0_error);
  if (port1 == nullptr || !single_item_batch(*port1, 1))
    return reject(f.error, KSJ_STATUS_CONTRACT_VIOLATION, port1_error);
  return KSJ_STATUS_OK;
}

ksj_status scratch_ok(Firing& f, uint64_t bytes, uint32_t alignment, std::string_view scratch_error) {
  ksj_scratch_view s{};
  s.abi = make_header(sizeof(s));
  const ksj_status st = f.cb->get_scratch(f.cb->host_context, f.lease, &s, f.error);
  if (st != KSJ_STATUS_OK) return st;
  if (!abi_ok(&s) || s.data == nullptr || s.byte_count < bytes || s.memory_domain != KSJ_PROVIDER_MEMORY_HOST_PAGEABLE ||
      !host_memory_ok(s.data, s.alignment, alignment))
    return reject(f.error, KSJ_STATUS_CONTRACT_VIOLATION, scratch_error);
  return KSJ_STATUS_OK;
}

ksj_status output_grant(Firing& f, uint64_t bytes, TypeId type, std::string_view output_error, GrantGuard& guard,
                        ksj_mutable_payload_view& out, bool allow_device) {
  ksj_output_grant* grant = nullptr;
  const ksj_status st = f.cb->acquire_output_grant(f.cb->host_context, f.lease, 0, &grant, f.error);
  if (st != KSJ_STATUS_OK) return st;
  if (grant == nullptr) return reject(f.error, KSJ_STATUS_CONTRACT_VIOLATION, output_error);
  guard.callbacks = f.cb->output_grants;
  guard.grant = grant;
  out = ksj_mutable_payload_view{};
  out.abi = make_header(sizeof(out));
  const ksj_status ms = f.cb->output_grants->map_mutable_payload(f.cb->output_grants->host_context, grant, &out, f.error);
  if (ms != KSJ_STATUS_OK) return ms;
  if (!abi_ok(&out) || out.data == nullptr || out.capacity_bytes < bytes || out.committed_bytes != 0 ||
      !type_matches(out.type, type))
    return reject(f.error, KSJ_STATUS_CONTRACT_VIOLATION, output_error);
  const bool host = out.memory_domain == KSJ_PROVIDER_MEMORY_HOST_PAGEABLE && host_memory_ok(out.data, out.alignment, 64);
  const bool dev = allow_device && out.memory_domain == KSJ_PROVIDER_MEMORY_DEVICE &&
                   device_memory_ok(out.data, out.alignment, f.op.device_ordinal);
  if (!host && !dev) return reject(f.error, KSJ_STATUS_CONTRACT_VIOLATION, output_error);
  return KSJ_STATUS_OK;
}

ksj_status seal_and_finish(Firing& f, GrantGuard& guard, const ksj_mutable_payload_view& out, uint64_t bytes,
                           const ksj_input_item_view& key_item, bool copy_metadata, uint64_t consumed,
                           uint64_t terminal_epoch) {
  ksj_output_seal_descriptor seal{};
  seal.abi = make_header(sizeof(seal));
  seal.output_port = 0;
  seal.produced_item_count = 1;
  seal.produced_byte_count = bytes;
  seal.semantic_key_hash = key_item.semantic_key_hash;
  seal.order_key = key_item.order_key;
  seal.type = out.type;
  if (copy_metadata) {
    seal.metadata = key_item.metadata;
  } else {
    seal.metadata.abi = make_header(sizeof(seal.metadata));
  }
  const ksj_status st = f.cb->output_grants->seal(f.cb->output_grants->host_context, guard.grant, &seal, f.error);
  if (st != KSJ_STATUS_OK) return st;
  guard.settled = true;
  write_done(f.result, 1, consumed, terminal_epoch);
  return KSJ_STATUS_OK;
}

const void* device_input(DeviceSession& d, DeviceSession::Slot slot, const ksj_payload_view& p, uint64_t bytes) {
  if (on_device(p)) return p.data;
  void* staged = d.buffer(slot, bytes);
  d.upload(staged, p.data, bytes);
  return staged;
}

void* device_output(DeviceSession& d, const ksj_mutable_payload_view& out, uint64_t bytes) {
  return on_device(out) ? out.data : d.buffer(DeviceSession::kOutput, bytes);
}

void finish_output(DeviceSession& d, const ksj_mutable_payload_view& out, const void* d_out, uint64_t bytes) {
  if (on_device(out))
    d.synchronize();
  else
    d.download(out.data, d_out, bytes);
}

DeviceSession& device(Firing& f) {
  if (!f.op.device) f.op.device = std::make_unique<DeviceSession>(f.op.device_ordinal);
  f.op.device->activate();
  return *f.op.device;
}

}  // namespace ksj::prov

// ---- the one exported symbol ----

using namespace ksj::prov;

KSJ_PROVIDER_ENTRY ksj_status KSJ_PROVIDER_CALL ksj_provider_query(const ksj_provider_query_request* request,
                                                                   ksj_provider_descriptor* out_descriptor,
                                                                   ksj_provider_api* out_api,
                                                                   ksj_error_view* out_error) {
  try {
    if (!abi_ok(request) || !abi_ok(out_descriptor) || !abi_ok(out_api))
      return reject(out_error, KSJ_STATUS_BAD_ABI, spec().bad_abi);
    if (!abi_ok(&request->host_build_id) || (request->host_build_id.size != 0 && request->host_build_id.data == nullptr))
      return reject(out_error, KSJ_STATUS_BAD_ABI, spec().bad_abi);
    *out_descriptor = metadata().descriptor;
    *out_api = ksj_provider_api{};
    out_api->abi = make_header(sizeof(*out_api));
    out_api->operator_create = &operator_create;
    out_api->execution_context_create = &execution_context_create;
    out_api->key_state_init = &key_state_init;
    out_api->operator_on_start = &operator_on_start;
    out_api->operator_process_batch = &operator_process_batch;
    out_api->operator_on_scan_end = &operator_on_scan_end;
    out_api->operator_on_cancel = &operator_on_cancel;
    out_api->key_state_reset = &key_state_reset;
    out_api->execution_context_destroy = &execution_context_destroy;
    out_api->operator_destroy = &operator_destroy;
    return KSJ_STATUS_OK;
  } catch (...) {
    return reject(out_error, KSJ_STATUS_INTERNAL_ERROR, spec().internal_error);
  }
}
