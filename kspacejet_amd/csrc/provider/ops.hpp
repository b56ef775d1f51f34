// Operator specifications of the provider modules.
#pragma once

#include "operator.hpp"

namespace ksj::prov {

// org.kspacejet.cartesian-recon
extern const OpSpec kCartesianIfft2CoilImages;  // cartesian_ifft2_coil_images
// org.kspacejet.coil-combine
extern const OpSpec kCoilCombineRss;            // coil_combine_rss
// org.kspacejet.noncartesian-recon
extern const OpSpec kNoncartesianAdjoint;       // noncartesian_adjoint_reconstruct
extern const OpSpec kRadialGridding;            // radial_gridding_reconstruct
// org.kspacejet.kspace-conditioning
extern const OpSpec kNoisePrewhiten;             // noise_prewhiten
extern const OpSpec kPhaseCorrect;               // phase_correct
extern const OpSpec kCoilCompress;               // coil_compress
extern const OpSpec kReadoutOversamplingRemove;  // readout_oversampling_remove
extern const OpSpec kNoncartesianNoisePrewhiten; // noncartesian_noise_prewhiten
extern const OpSpec kNoncartesianCoilCompress;   // noncartesian_coil_compress
// org.kspacejet.calibration
extern const OpSpec kNoiseModelEstimate;            // noise_model_estimate
extern const OpSpec kPhaseCorrectionEstimate;       // phase_correction_estimate
extern const OpSpec kCoilCompressionBasisEstimate;  // coil_compression_basis_estimate
// org.kspacejet.gpu-recon: the conditioning operators with device-memory edges
extern const OpSpec kGpuNoisePrewhiten;
extern const OpSpec kGpuPhaseCorrect;
extern const OpSpec kGpuCoilCompress;
extern const OpSpec kGpuReadoutOversamplingRemove;
extern const OpSpec kGpuNoncartesianNoisePrewhiten;
extern const OpSpec kGpuNoncartesianCoilCompress;
// org.kspacejet.gpu-recon (no reference implementation)
extern const OpSpec kKbGridding;                // kb_gridding_reconstruct
extern const OpSpec kKbGriddingRss;             // kb_gridding_reconstruct_rss
extern const OpSpec kCartesianIfft2Rss;         // cartesian_ifft2_rss
extern const OpSpec kCoilCombineSensitivity;    // coil_combine_sensitivity
extern const OpSpec kCoilCombineAdaptive;       // coil_combine_adaptive
extern const OpSpec kSensitivityMapEstimate;    // sensitivity_map_estimate

}  // namespace ksj::prov
