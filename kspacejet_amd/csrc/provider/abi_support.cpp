#include "abi_support.hpp"

#include <cstring>

namespace ksj::prov {

namespace {

struct TypeEntry {
  const char* type_ref;
  uint8_t digest[32];
  uint32_t kind;
  uint32_t element;
  uint32_t rank;
  const char* dims[8];
  uint64_t layout;
  uint32_t domains;
  uint32_t alignment;
  uint32_t mutability;
};

// Generated from types/registry.json by gen_types.py, in TypeId order.
const TypeEntry kTypes[] = {
#include "types.inc"
};
constexpr int kTypeCount = static_cast<int>(sizeof(kTypes) / sizeof(kTypes[0]));
static_assert(kTypeCount == static_cast<int>(TypeId::kPhaseReferenceFrame) + 1, "one entry per TypeId");

// Dimension-name views of each type, built once (static storage: descriptors
// handed out by type_descriptor() borrow them for the life of the module).
struct DimViews {
  ksj_utf8_view v[kTypeCount][8];
  DimViews() {
    for (int t = 0; t < kTypeCount; ++t)
      for (uint32_t d = 0; d < kTypes[t].rank; ++d) v[t][d] = utf8(kTypes[t].dims[d]);
  }
};
const DimViews& dim_views() {
  static const DimViews views;
  return views;
}

int hex_value(char c) {
  if (c >= '0' && c <= '9') return c - '0';
  if (c >= 'a' && c <= 'f') return c - 'a' + 10;
  if (c >= 'A' && c <= 'F') return c - 'A' + 10;
  return -1;
}

bool same_text(const char* data, uint64_t size, const char* expected) {
  const size_t n = std::strlen(expected);
  return data != nullptr && size == n && std::memcmp(data, expected, n) == 0;
}

}  // namespace

ksj_utf8_view utf8(std::string_view text) {
  ksj_utf8_view v{};
  v.abi = make_header(sizeof(v));
  v.data = text.data();
  v.size = text.size();
  return v;
}

bool text_is(const ksj_utf8_view& view, std::string_view expected) {
  return abi_ok(&view) && view.data != nullptr && view.size == expected.size() &&
         (expected.empty() || std::memcmp(view.data, expected.data(), expected.size()) == 0);
}

bool bytes_ok(const ksj_byte_view& bytes) { return abi_ok(&bytes) && (bytes.size == 0 || bytes.data != nullptr); }

ksj_digest256 digest_from_hex(std::string_view hex) {
  ksj_digest256 d{};
  if (hex.size() != 2 * KSJ_PROVIDER_DIGEST256_SIZE) return d;
  for (uint32_t i = 0; i < KSJ_PROVIDER_DIGEST256_SIZE; ++i) {
    const int hi = hex_value(hex[2 * i]), lo = hex_value(hex[2 * i + 1]);
    if (hi < 0 || lo < 0) return ksj_digest256{};
    d.bytes[i] = static_cast<uint8_t>(hi << 4 | lo);
  }
  d.abi = make_header(sizeof(d));
  return d;
}

ksj_status reject(ksj_error_view* out_error, ksj_status status, std::string_view message) {
  if (abi_ok(out_error)) {
    *out_error = ksj_error_view{};
    out_error->abi = make_header(sizeof(*out_error));
    out_error->status = status;
    out_error->message = utf8(message);
  }
  return status;
}

void write_done(ksj_process_result* out, uint32_t sealed, uint64_t consumed, uint64_t terminal_epoch) {
  if (!abi_ok(out)) return;
  *out = ksj_process_result{};
  out->abi = make_header(sizeof(*out));
  out->outcome = KSJ_PROVIDER_PROCESS_DONE;
  out->sealed_output_count = sealed;
  out->consumed_input_item_count = consumed;
  out->terminal_epoch = terminal_epoch;
  out->async_token = nullptr;
}

bool is_power_of_two(uint64_t v) { return v != 0 && (v & (v - 1)) == 0; }

bool host_memory_ok(const void* p, uint32_t declared_alignment, uint32_t required_alignment) {
  return is_power_of_two(required_alignment) && declared_alignment >= required_alignment &&
         is_power_of_two(declared_alignment) && p != nullptr &&
         (reinterpret_cast<uintptr_t>(p) & (uintptr_t{declared_alignment} - 1)) == 0;
}

bool callbacks_ok(const ksj_firing_lease_callbacks* cb, uint64_t required_caps) {
  if (!abi_ok(cb) || cb->get_info == nullptr || (cb->abi.capability_bits & required_caps) != required_caps)
    return false;
  const uint64_t caps = cb->abi.capability_bits;
  if ((caps & KSJ_LEASE_CAP_INPUT_BATCHES) && cb->get_input_batch == nullptr) return false;
  if (caps & KSJ_LEASE_CAP_OUTPUT_GRANTS) {
    const ksj_output_grant_callbacks* g = cb->output_grants;
    if (cb->acquire_output_grant == nullptr || !abi_ok(g) || g->map_mutable_payload == nullptr ||
        g->seal == nullptr || g->release == nullptr)
      return false;
  }
  if ((caps & KSJ_LEASE_CAP_SCRATCH) && cb->get_scratch == nullptr) return false;
  if ((caps & KSJ_LEASE_CAP_KEY_STATE) && cb->get_key_state == nullptr) return false;
  if ((caps & KSJ_LEASE_CAP_RETENTION) && (cb->retain_input == nullptr || cb->release_retention == nullptr))
    return false;
  if ((caps & KSJ_LEASE_CAP_ASYNC) &&
      (cb->register_async == nullptr || cb->complete_async == nullptr || cb->release_async == nullptr))
    return false;
  return !(caps & KSJ_LEASE_CAP_CANCELLATION) || cb->get_cancellation != nullptr;
}

ksj_type_descriptor_view type_descriptor(TypeId id) {
  const int t = static_cast<int>(id);
  const TypeEntry& e = kTypes[t];
  ksj_type_descriptor_view d{};
  d.abi = make_header(sizeof(d));
  d.type_ref = utf8(e.type_ref);
  d.type_identity_digest.abi = make_header(sizeof(d.type_identity_digest));
  std::memcpy(d.type_identity_digest.bytes, e.digest, sizeof(e.digest));
  d.payload_kind = e.kind;
  d.element_type = e.element;
  d.rank = e.rank;
  d.dimension_names = dim_views().v[t];
  d.layout_flags = e.layout;
  d.allowed_memory_domains = e.domains;
  d.minimum_alignment = e.alignment;
  d.mutability = e.mutability;
  return d;
}

bool type_matches(const ksj_type_descriptor_view& t, TypeId id) {
  // Structural validity first (provider_support.hpp:158-182) ...
  if (!abi_ok(&t) || !abi_ok(&t.type_ref) || t.type_ref.data == nullptr || t.type_ref.size == 0 ||
      !abi_ok(&t.type_identity_digest) || t.rank > 8 || t.minimum_alignment == 0 || t.reserved0 != 0 ||
      (t.rank != 0 && t.dimension_names == nullptr) || !is_power_of_two(t.minimum_alignment))
    return false;
  for (uint32_t i = 0; i < t.rank; ++i) {
    const ksj_utf8_view& n = t.dimension_names[i];
    if (!abi_ok(&n) || n.data == nullptr || n.size == 0) return false;
  }
  // ... then the exact registry identity (type_registry.h matchers).
  const TypeEntry& e = kTypes[static_cast<int>(id)];
  if (!same_text(t.type_ref.data, t.type_ref.size, e.type_ref) || t.payload_kind != e.kind ||
      t.element_type != e.element || t.rank != e.rank || t.layout_flags != e.layout ||
      t.allowed_memory_domains != e.domains || t.minimum_alignment != e.alignment || t.mutability != e.mutability ||
      std::memcmp(t.type_identity_digest.bytes, e.digest, 32) != 0)
    return false;
  for (uint32_t i = 0; i < e.rank; ++i)
    if (!same_text(t.dimension_names[i].data, t.dimension_names[i].size, e.dims[i])) return false;
  for (uint64_t s : t.stride_bytes)
    if (s != 0) return false;
  return true;
}

ConfigField u32_field(std::string_view key, uint32_t min, uint32_t max, uint32_t* out) {
  ConfigField f;
  f.key = key;
  f.min = min;
  f.max = max;
  f.out = out;
  return f;
}

ConfigField literal_field(std::string_view key, std::string_view literal, uint32_t* index_out) {
  ConfigField f;
  f.key = key;
  f.literal = literal;
  f.out = index_out;
  f.is_literal = true;
  return f;
}

bool parse_config(const ksj_byte_view& config, std::initializer_list<ConfigField> fields) {
  if (!abi_ok(&config) || config.data == nullptr || config.size == 0) return false;
  const std::string_view text(static_cast<const char*>(config.data), static_cast<size_t>(config.size));
  size_t pos = 0;
  auto expect = [&](std::string_view token) {
    if (text.compare(pos, token.size(), token) != 0) return false;
    pos += token.size();
    return true;
  };
  if (!expect("{")) return false;
  bool first = true;
  for (const ConfigField& f : fields) {
    if ((!first && !expect(",")) || !expect("\"") || !expect(f.key) || !expect("\":")) return false;
    first = false;
    if (f.is_literal) {
      // Alternatives separated by '|'; the longest match at `pos` that is followed
      // by ',' or '}' wins, so "a|ab" cannot accept a prefix of a longer value.
      std::string_view rest = f.literal;
      uint32_t index = 0;
      bool matched = false;
      while (!matched) {
        const size_t bar = rest.find('|');
        const std::string_view alt = rest.substr(0, bar);
        if (text.compare(pos, alt.size(), alt) == 0 && pos + alt.size() < text.size() &&
            (text[pos + alt.size()] == ',' || text[pos + alt.size()] == '}')) {
          pos += alt.size();
          matched = true;
          if (f.out) *f.out = index;
        }
        if (bar == std::string_view::npos) break;
        rest.remove_prefix(bar + 1);
        ++index;
      }
      if (!matched) return false;
      continue;
    }
    // Canonical integer: digits only, no leading zero, fits in uint32; a lone "0"
    // only where the field's range starts at zero.
    size_t end = pos;
    while (end < text.size() && text[end] >= '0' && text[end] <= '9') ++end;
    if (end == pos || end - pos > 10) return false;
    if (text[pos] == '0' && (f.min != 0 || end - pos != 1)) return false;
    uint64_t value = 0;
    for (size_t i = pos; i < end; ++i) value = value * 10 + static_cast<uint64_t>(text[i] - '0');
    if (value > UINT32_MAX || value < f.min || value > f.max) return false;
    *f.out = static_cast<uint32_t>(value);
    pos = end;
  }
  return expect("}") && pos == text.size();
}

GrantGuard::~GrantGuard() {
  if (callbacks == nullptr || grant == nullptr || settled || callbacks->release == nullptr) return;
  ksj_error_view error{};
  error.abi = make_header(sizeof(error));
  error.message.abi = make_header(sizeof(error.message));
  (void)callbacks->release(callbacks->host_context, grant, &error);
}

}  // namespace ksj::prov
