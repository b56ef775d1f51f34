"""org.kspacejet.calibration on the CPU side: descriptor bounds (the reference's output cap
and exact host-scratch maxima), its canonical configs and the malformed ones it rejects
(calibration_provider_tests.cpp:446-462, the parse_config of each operator), and the
firing's rejections in the reference's order (provider_api.cpp:269-350), none of which
reaches a GPU: a lease without the exact host scratch reservation, a scratch view of the
wrong size, the wrong input type, a byte count that derives no sample count, payload
metadata, a short output reservation, and a wrong output type (acquired, then released
without a seal).  A valid firing without a GPU fails loudly."""
import numpy as np
import pytest

from kspacejet_amd import provider_host as H

MODULE = "calibration"
CB = 8
CONFIGS = {
    "noise_model_estimate": '{"channel_count":2}',
    "phase_correction_estimate": '{"channel_count":2,"readout_sample_count":2}',
    "coil_compression_basis_estimate": '{"physical_channel_count":2,"virtual_channel_count":1}',
}


@pytest.fixture(scope="module")
def module():
    return H.ProviderModule(MODULE)


def opened(module, op_id, config=None):
    st, op, msg = module.create(op_id, config or CONFIGS[op_id])
    assert st == H.OK, msg
    assert op.open() == H.OK
    return op


def c64(n):
    return H.aligned_copy(np.arange(1, n + 1, dtype=np.float32).astype(np.complex64))


def noise_scratch(n):
    return 24 * n * n + 12 * n


def coil_scratch(n):
    return 24 * n * n + 4 * n


def test_operator_bounds(module):
    got = {H.view_text(o.operator_id): (o.max_input_items_per_firing, o.max_output_bytes_per_firing,
                                        o.max_scratch_bytes_per_firing) for o in module.operators}
    cap = 64 * 4096 * 8  # calibration_payload.hpp:22-23
    assert got == {"noise_model_estimate": (1, cap, noise_scratch(64)),
                   "phase_correction_estimate": (1, cap, 0),
                   "coil_compression_basis_estimate": (1, cap, coil_scratch(64))}


@pytest.mark.parametrize("op_id,config", list(CONFIGS.items()) + [
    ("noise_model_estimate", '{"channel_count":64}'),
    ("phase_correction_estimate", '{"channel_count":64,"readout_sample_count":4096}'),
    ("coil_compression_basis_estimate", '{"physical_channel_count":64,"virtual_channel_count":64}'),
])
def test_canonical_configs_create_operators(module, op_id, config):
    st, op, msg = module.create(op_id, config)
    assert st == H.OK, msg
    op.close()


@pytest.mark.parametrize("op_id,config", [
    ("noise_model_estimate", '{"channel_count":0}'),  # calibration_provider_tests.cpp:452
    ("noise_model_estimate", '{"channel_count":65}'),
    ("noise_model_estimate", '{"channel_count":02}'),
    ("noise_model_estimate", '{"channel_count": 2}'),
    ("phase_correction_estimate", '{"channel_count":2,"readout_sample_count":4097}'),
    ("phase_correction_estimate", '{"readout_sample_count":2,"channel_count":2}'),
    ("coil_compression_basis_estimate", '{"physical_channel_count":2,"virtual_channel_count":3}'),
    ("coil_compression_basis_estimate", '{"physical_channel_count":65,"virtual_channel_count":1}'),
    ("coil_compression_basis_estimate", '{"physical_channel_count":2,"virtual_channel_count":0}'),
])
def test_malformed_configs_are_invalid_arguments(module, op_id, config):
    st, op, _ = module.create(op_id, config)
    assert st == H.INVALID_ARGUMENT and op is None


def lease(data, in_type, out_bytes, out_type, scratch, metadata=b""):
    return H.Lease([H.Batch(0, [H.Item(data, in_type, metadata=metadata)])], H.OutputSpec(out_bytes, out_type),
                   scratch_bytes=scratch)


def rejected_before_output(op, ls):
    st, r, msg = op.process(ls)
    assert st == H.CONTRACT_VIOLATION, msg
    assert ls.acquire_calls == 0 and ls.seal_calls == 0 and r.sealed_output_count == 0
    return msg


def test_scratch_reservation_must_be_exact(module):
    noise = opened(module, "noise_model_estimate")
    for scratch in (0, noise_scratch(2) - 8, noise_scratch(2) + 8):
        assert "scratch" in rejected_before_output(
            noise, lease(c64(8), "ksj.noise-calibration-frame", 4 * CB, "ksj.noise-model", scratch))
    phase = opened(module, "phase_correction_estimate")
    assert "scratch" in rejected_before_output(
        phase, lease(c64(8), "ksj.phase-reference-frame", 4 * CB, "ksj.phase-model", 64))
    # the scratch view itself must have the reserved size
    coil = opened(module, "coil_compression_basis_estimate")
    ls = lease(c64(8), "ksj.kspace-frame", 2 * CB, "ksj.coil-compression-basis", coil_scratch(2))
    ls.scratch_bytes = coil_scratch(2) + 64  # view larger than the reservation
    ls.reserved_scratch_bytes = coil_scratch(2)
    assert "scratch" in rejected_before_output(coil, ls)


def test_input_type_byte_count_and_metadata(module):
    noise = opened(module, "noise_model_estimate")
    ok = noise_scratch(2)
    assert "payload" in rejected_before_output(
        noise, lease(c64(8), "ksj.kspace-frame", 4 * CB, "ksj.noise-model", ok))
    # one sample per channel is too few for a covariance (noise_model_estimate.cpp:94-105)
    assert "payload" in rejected_before_output(
        noise, lease(c64(2), "ksj.noise-calibration-frame", 4 * CB, "ksj.noise-model", ok))
    # not a whole number of channel-sample pairs
    assert "payload" in rejected_before_output(
        noise, lease(c64(7), "ksj.noise-calibration-frame", 4 * CB, "ksj.noise-model", ok))
    assert "payload" in rejected_before_output(
        noise, lease(c64(8), "ksj.noise-calibration-frame", 4 * CB, "ksj.noise-model", ok, metadata=b"m"))
    phase = opened(module, "phase_correction_estimate")
    assert "payload" in rejected_before_output(  # not whole [channel][line][readout] lines
        phase, lease(c64(6), "ksj.phase-reference-frame", 4 * CB, "ksj.phase-model", 0))


def test_output_reservation_and_type(module):
    coil = opened(module, "coil_compression_basis_estimate")
    ls = lease(c64(8), "ksj.kspace-frame", 2 * CB, "ksj.coil-compression-basis", coil_scratch(2))
    ls.reserved_output_bytes = 2 * CB - 1
    assert "one input" in rejected_before_output(coil, ls)
    ls = lease(c64(8), "ksj.kspace-frame", 2 * CB, "ksj.noise-model", coil_scratch(2))
    st, r, msg = coil.process(ls)
    assert st == H.CONTRACT_VIOLATION and ls.acquire_calls == 1 and ls.release_calls == 1 and ls.seal_calls == 0


def test_valid_firing_without_a_device_fails_loudly(module):
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    op = opened(module, "noise_model_estimate")
    ls = lease(c64(8), "ksj.noise-calibration-frame", 4 * CB, "ksj.noise-model", noise_scratch(2))
    st, r, msg = op.process(ls)
    assert st in (H.INTERNAL_ERROR, H.RESOURCE_EXHAUSTED), msg
    assert ls.acquire_calls == 1 and ls.seal_calls == 0 and ls.release_calls == 1 and r.sealed_output_count == 0
    assert not np.any(ls.output)
    op.close()
