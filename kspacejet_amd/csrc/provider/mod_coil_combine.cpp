// Provider module org.kspacejet.coil-combine (libksj_provider_coil_combine.so).
#include "ops.hpp"

namespace ksj::prov {

namespace {
const OpSpec* const kOps[] = {&kCoilCombineRss};
}  // namespace

const ProviderSpec& provider_spec() {
  static const ProviderSpec spec{"org.kspacejet.coil-combine",
                                 "ba4f619047391419c8ffc42220ace7d888f047c9e2fd64aaeb78d9093191aed5",
                                 kOps,
                                 1,
                                 "coil-combine provider received ABI storage it cannot use",
                                 "coil-combine provider received an invalid lifecycle argument",
                                 "coil-combine provider does not expose the requested operator",
                                 "coil-combine provider trapped an unexpected internal error"};
  return spec;
}

}  // namespace ksj::prov
