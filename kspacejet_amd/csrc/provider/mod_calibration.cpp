// Provider module org.kspacejet.calibration (libksj_provider_calibration.so).
#include "ops.hpp"

namespace ksj::prov {

namespace {
// Descriptor order of the reference module (provider_api.cpp:142-149).
const OpSpec* const kOps[] = {&kNoiseModelEstimate, &kPhaseCorrectionEstimate, &kCoilCompressionBasisEstimate};
}  // namespace

const ProviderSpec& provider_spec() {
  static const ProviderSpec spec{"org.kspacejet.calibration",
                                 "abd611a1428074691393e7057e21266246647acd9c8ec7d1d4ad65a971d40f25",
                                 kOps,
                                 3,
                                 "Calibration Provider received incompatible ABI storage",
                                 "Calibration Provider received an invalid lifecycle argument",
                                 "Calibration Provider does not expose the requested operator",
                                 "Calibration Provider trapped an unexpected internal exception"};
  return spec;
}

}  // namespace ksj::prov
