// Provider module org.kspacejet.gpu-recon (libksj_provider_gpu_recon.so): the
// operators with no reference implementation -- Kaiser-Bessel gridding on the
// oversampled grid (coil images or fused RSS), the fused Cartesian ifft2 + RSS,
// the sensitivity-weighted combine (as coil_combine_sensitivity and under the planned
// coil_combine_adaptive id), the planned sensitivity_map_estimate (Walsh's method), and
// the k-space conditioning operators with device-memory edges.  The bundle
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
                              &kGpuNoncartesianCoilCompress,
                              &kSensitivityMapEstimate,
                              &kCoilCombineAdaptive};
}  // namespace

const ProviderSpec& provider_spec() {
  static const ProviderSpec spec{"org.kspacejet.gpu-recon",
                                 "7325318366a47a792100c624accea7c0bfc109ca062ce21a22c9342df718052a",
                                 kOps,
                                 12,
                                 "GPU reconstruction provider received ABI storage it cannot use",
                                 "GPU reconstruction provider received an invalid lifecycle argument",
                                 "GPU reconstruction provider does not expose the requested operator",
                                 "GPU reconstruction provider trapped an unexpected internal error"};
  return spec;
}

}  // namespace ksj::prov
