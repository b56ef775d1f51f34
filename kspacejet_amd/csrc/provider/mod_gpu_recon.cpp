// Provider module org.kspacejet.gpu-recon (libksj_provider_gpu_recon.so): the
// operators with no reference implementation -- Kaiser-Bessel gridding on the
// oversampled grid (coil images or fused RSS), the fused Cartesian ifft2 + RSS,
// and the sensitivity-weighted combine of the planned coil_combine_adaptive
// interface.  The bundle digest follows the reference's development rule:
// SHA-256 of NUL-terminated "kspacejet.provider-bundle", provider id, operator ids.
#include "ops.hpp"

namespace ksj::prov {

namespace {
const OpSpec* const kOps[] = {&kKbGridding, &kKbGriddingRss, &kCartesianIfft2Rss, &kCoilCombineSensitivity};
}  // namespace

const ProviderSpec& provider_spec() {
  static const ProviderSpec spec{"org.kspacejet.gpu-recon",
                                 "a0a5cc5bbf29d1a6746f6f81d5748441b50f04dbf32c10650374f61ee91980f3",
                                 kOps,
                                 4,
                                 "GPU reconstruction provider received ABI storage it cannot use",
                                 "GPU reconstruction provider received an invalid lifecycle argument",
                                 "GPU reconstruction provider does not expose the requested operator",
                                 "GPU reconstruction provider trapped an unexpected internal error"};
  return spec;
}

}  // namespace ksj::prov
