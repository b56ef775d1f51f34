// Provider module org.kspacejet.cartesian-recon (libksj_provider_cartesian_recon.so).
#include "ops.hpp"

namespace ksj::prov {

namespace {
const OpSpec* const kOps[] = {&kCartesianIfft2CoilImages};
}  // namespace

const ProviderSpec& provider_spec() {
  static const ProviderSpec spec{"org.kspacejet.cartesian-recon",
                                 "b49ba77e1655bdbbc4839ab692eb8fb7d9551bcb7ba138c6274de836d4ab5006",
                                 kOps,
                                 1,
                                 "Cartesian reconstruction provider received ABI storage it cannot use",
                                 "Cartesian reconstruction provider received an invalid lifecycle argument",
                                 "Cartesian reconstruction provider does not expose the requested operator",
                                 "Cartesian reconstruction provider trapped an unexpected internal error"};
  return spec;
}

}  // namespace ksj::prov
