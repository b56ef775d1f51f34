"""Config parsing of every Provider operator under mutation (CPU only).  The canonical
configs of all modules are truncated at every byte, have every byte replaced by each
of a set of grammar-significant characters, and get digits, signs, exponents and huge
numbers spliced into their values; every result must be OK (the mutation left a
canonical config in the envelope) or INVALID_ARGUMENT, never another status or a crash.
Run under the host sanitizers by test_provider_sanitized.py."""
import pytest

from kspacejet_amd import provider_host as H

CANONICAL = {
    "noncartesian_recon": {
        "noncartesian_adjoint_reconstruct": '{"channels":2,"image_cols":8,"image_rows":8,"sample_count":16}',
        "radial_gridding_reconstruct": '{"channels":2,"density_compensation":"radial_analytic_ramp","image_cols":8,'
                                       '"image_rows":8,"sample_count":16,"trajectory_units":"radians_per_pixel"}',
    },
    "cartesian_recon": {"cartesian_ifft2_coil_images": '{"channels":8,"cols":128,"rows":128}'},
    "coil_combine": {"coil_combine_rss": '{"channels":8,"cols":128,"rows":128}'},
    "gpu_recon": {
        "kb_gridding_reconstruct": '{"channels":16,"density_compensation":"radial_analytic_ramp","image_cols":256,'
                                   '"image_rows":256,"kernel_width":4,"oversampling":2,"sample_count":205824,'
                                   '"trajectory_units":"radians_per_pixel"}',
        "cartesian_ifft2_rss": '{"centred":true,"channels":8,"cols":256,"rows":256}',
        "coil_combine_sensitivity": '{"channels":8,"cols":256,"rows":256}',
    },
    "kspace_conditioning": {
        "noise_prewhiten": '{"channel_count":2,"cols":2,"rows":2}',
        "coil_compress": '{"cols":2,"physical_channel_count":2,"rows":2,"virtual_channel_count":1}',
        "readout_oversampling_remove": '{"channel_count":2,"input_cols":4,"output_cols":2,"readout_offset":1,'
                                       '"rows":2}',
    },
    "calibration": {
        "noise_model_estimate": '{"channel_count":2}',
        "coil_compression_basis_estimate": '{"physical_channel_count":2,"virtual_channel_count":1}',
        "phase_correction_estimate": '{"channel_count":2,"readout_sample_count":2}',
    },
}
CHARS = ['"', ':', ',', '{', '}', ' ', '0', '9', '-', '+', '.', 'e', 'x', '\\', '\x00', '\xff']
SPLICES = ["0", "00", "-1", "+1", "1e3", "1.5", "4294967296", "18446744073709551616", "9" * 40, '"', "null", "true"]


def mutations(cfg: str):
    b = cfg.encode()
    for i in range(len(b)):
        yield b[:i]
    for i in range(len(b)):
        for ch in CHARS:
            yield b[:i] + ch.encode("latin-1") + b[i + 1:]
    for i, c in enumerate(b):
        if chr(c).isdigit():
            for s in SPLICES:
                yield b[:i] + s.encode() + b[i + 1:]
    yield b + b
    yield b"{" + b


@pytest.fixture(scope="module")
def modules():
    return {name: H.ProviderModule(name) for name in CANONICAL}


@pytest.mark.parametrize("module,op_id", [(m, o) for m, ops in CANONICAL.items() for o in ops])
def test_mutated_configs_are_accepted_or_invalid(modules, module, op_id):
    mod = modules[module]
    st, op, msg = mod.create(op_id, CANONICAL[module][op_id])
    assert st == H.OK, msg
    op.close()
    seen = {}
    for cfg in mutations(CANONICAL[module][op_id]):
        st, op, _ = mod.create(op_id, cfg)
        assert st in (H.OK, H.INVALID_ARGUMENT), (cfg, st)
        seen[st] = seen.get(st, 0) + 1
        if op is not None:
            op.close()
    assert seen.get(H.INVALID_ARGUMENT, 0) > 0
