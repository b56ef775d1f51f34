// Provider module org.kspacejet.noncartesian-recon (libksj_provider_noncartesian_recon.so):
// the reference's two operators in the reference's descriptor order
// (providers/kspacejet-noncartesian-recon/src/provider_api.cpp:36-43).
#include "ops.hpp"

namespace ksj::prov {

namespace {
const OpSpec* const kOps[] = {&kNoncartesianAdjoint, &kRadialGridding};
}  // namespace

const ProviderSpec& provider_spec() {
  // bundle_digest = SHA-256 of the NUL-terminated fields "kspacejet.provider-bundle",
  // the provider id and each operator id in descriptor order.
  static const ProviderSpec spec{"org.kspacejet.noncartesian-recon",
                                 "4297da20fe070aae1988f456967aeed82d5bae62f8aad41585e25fe767000ef6",
                                 kOps,
                                 2,
                                 "non-Cartesian reconstruction provider received ABI storage it cannot use",
                                 "non-Cartesian reconstruction provider received an invalid lifecycle argument",
                                 "non-Cartesian reconstruction provider does not expose the requested operator",
                                 "non-Cartesian reconstruction provider trapped an unexpected internal error"};
  return spec;
}

}  // namespace ksj::prov
