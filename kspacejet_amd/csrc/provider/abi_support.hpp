// Defensive helpers at the Provider ABI boundary: header checks, error views,
// the registry type matchers, canonical-config parsing and output-grant release.
// The checks mirror what the reference providers apply at the same boundary
// (libs/recon/kspacejet-provider-sdk/include/kspacejet/provider/detail/provider_support.hpp:29-285).
#pragma once

#include "ksj_provider_abi.h"

#include <cstdint>
#include <initializer_list>
#include <string_view>

namespace ksj::prov {

inline ksj_provider_abi_header make_header(uint32_t struct_size, uint64_t capability_bits = 0) {
  return ksj_provider_abi_header_make(struct_size, capability_bits);
}

// A struct is usable when it is non-null, at least as large as our definition
// and every reserved header field is zero.
template <typename T> bool abi_ok(const T* v) {
  return v != nullptr && v->abi.struct_size >= sizeof(T) && v->abi.reserved0 == 0 && v->abi.reserved[0] == 0 &&
         v->abi.reserved[1] == 0;
}

ksj_utf8_view utf8(std::string_view text);
bool text_is(const ksj_utf8_view& view, std::string_view expected);
// A borrowed byte range: compatible header, and data present whenever size > 0.
bool bytes_ok(const ksj_byte_view& bytes);
// 64 hex digits -> digest with a valid header (zeroed digest on malformed input).
ksj_digest256 digest_from_hex(std::string_view hex);

// Fills *out_error (when it has a compatible header) and returns `status`.  The
// message must outlive the callback: string literals or thread-local storage.
ksj_status reject(ksj_error_view* out_error, ksj_status status, std::string_view message);
// DONE result with the given counts (provider_support.hpp:255-268).
void write_done(ksj_process_result* out, uint32_t sealed, uint64_t consumed, uint64_t terminal_epoch);

bool is_power_of_two(uint64_t v);
// declared_alignment >= required (a power of two) and the pointer honours it.
bool host_memory_ok(const void* p, uint32_t declared_alignment, uint32_t required_alignment);
// The lease table offers every callback that `required_caps` and its own capability bits name.
bool callbacks_ok(const ksj_firing_lease_callbacks* cb, uint64_t required_caps);

// Payload types of types/registry.json used by these providers.
enum class TypeId : int {
  kKspaceFrame = 0,
  kNoncartesianKspaceFrame,
  kTrajectoryFrame,
  kCoilImageFrame,
  kImageFrame,
  kComplexImageFrame,
  kSensitivityMap,
  kNoiseModel,
  kPhaseModel,
  kCoilCompressionBasis,
  kNoiseCalibrationFrame,
  kPhaseReferenceFrame,
};
// The full registry descriptor of `id` (what a host places in a payload view).
ksj_type_descriptor_view type_descriptor(TypeId id);
// Structural validity of the descriptor and an exact match of every identity field:
// type_ref, identity digest, kind, element type, rank, dimension names, layout
// flags, strides, memory domains, alignment and mutability.
bool type_matches(const ksj_type_descriptor_view& t, TypeId id);

// One field of a canonical config object.  Canonical means keys in a fixed
// (sorted) order, no whitespace, integers without sign or leading zeros (zero
// itself, written "0", only in a field whose minimum is 0), and literal values
// compared byte for byte (string literals include their quotes).
struct ConfigField {
  std::string_view key;
  uint32_t min = 0, max = 0;       // integer range (ignored for literals)
  uint32_t* out = nullptr;         // integer destination; null for a literal field
  std::string_view literal = {};   // accepted literal(s), '|' separated; out receives the index when given
  bool is_literal = false;
};
ConfigField u32_field(std::string_view key, uint32_t min, uint32_t max, uint32_t* out);
ConfigField literal_field(std::string_view key, std::string_view literal, uint32_t* index_out = nullptr);
bool parse_config(const ksj_byte_view& config, std::initializer_list<ConfigField> fields);

// Releases an acquired output grant unless it was sealed (provider.h:508-520).
struct GrantGuard {
  const ksj_output_grant_callbacks* callbacks = nullptr;
  ksj_output_grant* grant = nullptr;
  bool settled = false;
  ~GrantGuard();
};

}  // namespace ksj::prov
