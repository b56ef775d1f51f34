"""org.kspacejet.kspace-conditioning on the CPU side: descriptor bounds, the canonical
configs of the reference's tests, the malformed configs it rejects
(kspace_conditioning_provider_tests.cpp:506-537), and the firing's rejections in the
reference's order (provider_api.cpp:323-437; tests.cpp:731-825), none of which reaches a
GPU: wrong types and byte counts, duplicate ports, keyed calibration items of another
scan, payload metadata, host scratch, short output reservations, and a wrong output type
(acquired, then released without a seal)."""
import numpy as np
import pytest

from kspacejet_amd import provider_host as H

MODULE = "kspace_conditioning"
CONFIGS = {
    "noise_prewhiten": '{"channel_count":2,"cols":2,"rows":2}',
    "phase_correct": '{"channel_count":2,"cols":2,"rows":2}',
    "coil_compress": '{"cols":2,"physical_channel_count":2,"rows":2,"virtual_channel_count":1}',
    "readout_oversampling_remove": '{"channel_count":2,"input_cols":4,"output_cols":2,"readout_offset":1,"rows":2}',
    "noncartesian_noise_prewhiten": '{"channel_count":2,"sample_count":3}',
    "noncartesian_coil_compress": '{"physical_channel_count":2,"sample_count":3,"virtual_channel_count":1}',
}
CB = 8  # bytes per complex float


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


def test_operator_bounds(module):
    for o in module.operators:
        op_id = H.view_text(o.operator_id)
        assert o.max_input_items_per_firing == (1 if op_id == "readout_oversampling_remove" else 2)
        assert o.max_output_bytes_per_firing == 64 * 512 * 512 * 8
        assert o.max_scratch_bytes_per_firing == 0


@pytest.mark.parametrize("op_id,config", list(CONFIGS.items()) + [
    ("noise_prewhiten", '{"channel_count":64,"cols":512,"rows":512}'),
    ("coil_compress", '{"cols":512,"physical_channel_count":64,"rows":512,"virtual_channel_count":64}'),
    ("readout_oversampling_remove", '{"channel_count":1,"input_cols":512,"output_cols":256,"readout_offset":0,'
                                    '"rows":2}'),
    ("readout_oversampling_remove", '{"channel_count":1,"input_cols":512,"output_cols":2,"readout_offset":510,'
                                    '"rows":2}'),
    ("noncartesian_noise_prewhiten", '{"channel_count":64,"sample_count":65536}'),
    ("noncartesian_coil_compress", '{"physical_channel_count":64,"sample_count":1,"virtual_channel_count":1}'),
])
def test_canonical_configs_create_operators(module, op_id, config):
    st, op, msg = module.create(op_id, config)
    assert st == H.OK, msg
    op.close()


@pytest.mark.parametrize("op_id,config", [
    # kspace_conditioning_provider_tests.cpp:517-524
    ("noise_prewhiten", '{"channel_count":0,"cols":2,"rows":2}'),
    ("phase_correct", '{"channel_count":2,"rows":2,"cols":2}'),
    ("coil_compress", '{"cols":2,"physical_channel_count":1,"rows":2,"virtual_channel_count":2}'),
    ("readout_oversampling_remove", '{"channel_count":2,"input_cols":4,"output_cols":3,"readout_offset":2,"rows":2}'),
    ("noncartesian_noise_prewhiten", '{"channel_count":2,"sample_count":0}'),
    ("noncartesian_coil_compress", '{"physical_channel_count":1,"sample_count":3,"virtual_channel_count":2}'),
    # the envelopes and the integer grammar
    ("noise_prewhiten", '{"channel_count":65,"cols":2,"rows":2}'),
    ("noise_prewhiten", '{"channel_count":2,"cols":1,"rows":2}'),
    ("phase_correct", '{"channel_count":2,"cols":2,"rows":513}'),
    ("readout_oversampling_remove", '{"channel_count":2,"input_cols":4,"output_cols":2,"readout_offset":00,"rows":2}'),
    ("readout_oversampling_remove", '{"channel_count":2,"input_cols":4,"output_cols":2,"readout_offset":01,"rows":2}'),
    ("readout_oversampling_remove", '{"channel_count":2,"input_cols":4,"output_cols":1,"readout_offset":0,"rows":2}'),
    ("noncartesian_noise_prewhiten", '{"channel_count":2,"sample_count":65537}'),
    ("noncartesian_noise_prewhiten", '{"channel_count":0,"sample_count":3}'),
    ("noncartesian_coil_compress", '{"physical_channel_count":65,"sample_count":3,"virtual_channel_count":1}'),
])
def test_malformed_configs_are_invalid_arguments(module, op_id, config):
    st, op, _ = module.create(op_id, config)
    assert st == H.INVALID_ARGUMENT and op is None


def lease(dyn, dyn_type, cal=None, cal_type=None, out_bytes=None, out_type=None, cal_keys=(71, 73)):
    batches = [H.Batch(0, [H.Item(dyn, dyn_type)])]
    if cal is not None:
        batches.append(H.Batch(1, [H.Item(cal, cal_type, *cal_keys)]))
    return H.Lease(batches, H.OutputSpec(out_bytes, out_type or dyn_type), scratch_bytes=0)


def rejected_before_output(op, ls):
    st, r, msg = op.process(ls)
    assert st == H.CONTRACT_VIOLATION, msg
    assert ls.acquire_calls == 0 and ls.seal_calls == 0 and r.sealed_output_count == 0
    return msg


def test_wrong_types_and_byte_counts(module):
    """tests.cpp:731-784."""
    noise = opened(module, "noise_prewhiten")
    assert "k-space" in rejected_before_output(
        noise, lease(c64(8), "ksj.image-frame", c64(4), "ksj.noise-model", 8 * CB, "ksj.kspace-frame"))
    phase = opened(module, "phase_correct")
    rejected_before_output(phase, lease(c64(8)[:7], "ksj.kspace-frame", c64(4), "ksj.phase-model", 8 * CB))
    nc = opened(module, "noncartesian_noise_prewhiten")
    rejected_before_output(nc, lease(c64(6), "ksj.kspace-frame", c64(4), "ksj.noise-model", 6 * CB,
                                     "ksj.noncartesian-kspace-frame"))
    ncc = opened(module, "noncartesian_coil_compress")
    assert "calibration" in rejected_before_output(
        ncc, lease(c64(6), "ksj.noncartesian-kspace-frame", c64(1), "ksj.coil-compression-basis", 3 * CB))
    # a calibration item of another type
    rejected_before_output(noise, lease(c64(8), "ksj.kspace-frame", c64(4), "ksj.phase-model", 8 * CB))
    # a wrong output type is found only after the grant was acquired: released, not sealed
    out = opened(module, "noncartesian_noise_prewhiten")
    ls = lease(c64(6), "ksj.noncartesian-kspace-frame", c64(4), "ksj.noise-model", 6 * CB, "ksj.kspace-frame")
    st, r, _ = out.process(ls)
    assert st == H.CONTRACT_VIOLATION and ls.acquire_calls == 1 and ls.release_calls == 1 and ls.seal_calls == 0


def test_duplicate_port_and_mismatched_identity(module):
    """tests.cpp:786-825, plus an unkeyed (zero-key) calibration item getting past the identity
    check to the next one."""
    coil = opened(module, "coil_compress")
    ls = lease(c64(8), "ksj.kspace-frame", c64(2), "ksj.coil-compression-basis", 4 * CB)
    ls.batches[1].port = 0
    ls._build()
    assert "port" in rejected_before_output(coil, ls)
    assert "key" in rejected_before_output(
        coil, lease(c64(8), "ksj.kspace-frame", c64(2), "ksj.coil-compression-basis", 4 * CB, cal_keys=(12, 73)))
    nc = opened(module, "noncartesian_noise_prewhiten")
    assert "key" in rejected_before_output(
        nc, lease(c64(6), "ksj.noncartesian-kspace-frame", c64(4), "ksj.noise-model", 6 * CB, cal_keys=(71, 18)))
    # zero keys are unkeyed: accepted by the identity rule, then the short reservation is the rejection
    ls = lease(c64(6), "ksj.noncartesian-kspace-frame", c64(4), "ksj.noise-model", 6 * CB, cal_keys=(0, 0))
    ls.reserved_output_bytes = 6 * CB - 1
    assert "no host scratch" in rejected_before_output(nc, ls)


def test_lease_shape_metadata_and_scratch(module):
    crop = opened(module, "readout_oversampling_remove")
    # one input batch only: a second batch is a lease mismatch
    ls = lease(c64(16), "ksj.kspace-frame", c64(4), "ksj.noise-model", 8 * CB)
    rejected_before_output(crop, ls)
    # host scratch is not allowed
    ls = H.Lease([H.Batch(0, [H.Item(c64(16), "ksj.kspace-frame")])], H.OutputSpec(8 * CB, "ksj.kspace-frame"),
                 scratch_bytes=64)
    rejected_before_output(crop, ls)
    # payload metadata must be empty
    ls = H.Lease([H.Batch(0, [H.Item(c64(16), "ksj.kspace-frame", metadata=b"x")])],
                 H.OutputSpec(8 * CB, "ksj.kspace-frame"), scratch_bytes=0)
    rejected_before_output(crop, ls)
    # the port of a one-batch operator is 0
    ls = H.Lease([H.Batch(1, [H.Item(c64(16), "ksj.kspace-frame")])], H.OutputSpec(8 * CB, "ksj.kspace-frame"),
                 scratch_bytes=0)
    rejected_before_output(crop, ls)
    # a device-labelled payload is a contract violation here, as in the reference
    ls = H.Lease([H.Batch(0, [H.Item(c64(16), "ksj.kspace-frame", memory_domain=H.MEMORY_DEVICE)])],
                 H.OutputSpec(8 * CB, "ksj.kspace-frame"), scratch_bytes=0)
    rejected_before_output(crop, ls)


def test_valid_firing_without_a_device_fails_loudly(module):
    """No CPU fallback: on a host with no GPU a valid firing fails after the grant is
    acquired, the grant is released and the output stays untouched."""
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    op = opened(module, "noncartesian_noise_prewhiten")
    ls = lease(c64(6), "ksj.noncartesian-kspace-frame", c64(4), "ksj.noise-model", 6 * CB)
    st, r, msg = op.process(ls)
    assert st in (H.INTERNAL_ERROR, H.RESOURCE_EXHAUSTED), msg
    assert ls.acquire_calls == 1 and ls.seal_calls == 0 and ls.release_calls == 1 and r.sealed_output_count == 0
    assert not np.any(ls.output)
    op.close()
