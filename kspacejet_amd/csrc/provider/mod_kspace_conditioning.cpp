// Provider module org.kspacejet.kspace-conditioning (libksj_provider_kspace_conditioning.so).
#include "ops.hpp"

namespace ksj::prov {

namespace {
// Descriptor order of the reference module (provider_api.cpp:228-236).
const OpSpec* const kOps[] = {&kNoisePrewhiten,           &kPhaseCorrect,
                              &kCoilCompress,             &kReadoutOversamplingRemove,
                              &kNoncartesianNoisePrewhiten, &kNoncartesianCoilCompress};
}  // namespace

const ProviderSpec& provider_spec() {
  static const ProviderSpec spec{"org.kspacejet.kspace-conditioning",
                                 "6ff48e549c4ff3ad41cfebb766e6d298360bf686dc7882f5df9e6401b7dc7325",
                                 kOps,
                                 6,
                                 "kspace-conditioning provider received ABI storage it cannot use",
                                 "kspace-conditioning provider received an invalid lifecycle argument",
                                 "kspace-conditioning provider does not expose the requested operator",
                                 "kspace-conditioning provider trapped an unexpected internal error"};
  return spec;
}

}  // namespace ksj::prov
