// Provider module org.kspacejet.gpu-recon (libksj_provider_gpu_recon.so): the
// operators with no reference implementation -- Kaiser-Bessel gridding on the
// oversampled grid (coil images or fused RSS), the fused Cartesian ifft2 + RSS,
// the sensitivity-weighted combine of the planned coil_combine_adaptive interface,
// and the k-space conditioning operators with device-memory edges.  The bundle
// digest follows the reference's development rule: SHA-256 of NUL-terminated
// "kspacejet.provider-bundle", provider id, operator ids.
#include "ops.hpp"

namespace ksj::prov {

namespace {
const OpSpec* const kOps[] = {&kKbGridding,
                              &kKbGriddingRss,
                              &kCartesianIfft2Rss,
                              &kCoilCombineSensitivity,
                              &kGpuNoisePrewhiten,
                              &kGpuPhaseCorrect,
                              &kGpuCoilCompress,
                              &kGpuReadoutOversamplingRemove,
                              &kGpuNoncartesianNoisePrewhiten,
                              &kGpuNoncartesianCoilCompress};
}  // namespace

const ProviderSpec& provider_spec() {
  static const ProviderSpec spec{"org.kspacejet.gpu-recon",
                                 "caf57397a99179fa253a77ba02b6d3f24784e83ab62093830a8a927cb65ff277",
                                 kOps,
                                 10,
                                 "GPU reconstruction provider received ABI storage it cannot use",
                                 "GPU reconstruction provider received an invalid lifecycle argument",
                                 "GPU reconstruction provider does not expose the requested operator",
                                 "GPU reconstruction provider trapped an unexpected internal error"};
  return spec;
}

}  // namespace ksj::prov
