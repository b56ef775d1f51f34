"""Provider ABI checks that need no GPU: struct layout against the reference
SDK, type identity digests, provider descriptors and bundle digests, the
canonical config grammar, the lifecycle, and every firing rejection the
reference's provider tests exercise (all happen before any device work).

Cases mirror tests/unit/providers/kspacejet-{noncartesian-recon,cartesian-recon,
coil-combine}/*_provider_tests.cpp of the reference.
"""
import ctypes
import hashlib
import json
import math
import os
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
sys.path.insert(0, GOLDEN)

from kspacejet_amd import provider_host as H  # noqa: E402

PROVIDERS = {
    # module: (provider id, operator ids in descriptor order, reference bundle digest literal)
    "noncartesian_recon": ("org.kspacejet.noncartesian-recon",
                           ["noncartesian_adjoint_reconstruct", "radial_gridding_reconstruct"],
                           "4297da20fe070aae1988f456967aeed82d5bae62f8aad41585e25fe767000ef6"),
    "cartesian_recon": ("org.kspacejet.cartesian-recon", ["cartesian_ifft2_coil_images"],
                        "b49ba77e1655bdbbc4839ab692eb8fb7d9551bcb7ba138c6274de836d4ab5006"),
    "coil_combine": ("org.kspacejet.coil-combine", ["coil_combine_rss"],
                     "ba4f619047391419c8ffc42220ace7d888f047c9e2fd64aaeb78d9093191aed5"),
    "kspace_conditioning": ("org.kspacejet.kspace-conditioning",
                            ["noise_prewhiten", "phase_correct", "coil_compress", "readout_oversampling_remove",
                             "noncartesian_noise_prewhiten", "noncartesian_coil_compress"],
                            "6ff48e549c4ff3ad41cfebb766e6d298360bf686dc7882f5df9e6401b7dc7325"),
    "calibration": ("org.kspacejet.calibration",
                    ["noise_model_estimate", "phase_correction_estimate", "coil_compression_basis_estimate"],
                    "abd611a1428074691393e7057e21266246647acd9c8ec7d1d4ad65a971d40f25"),
    "gpu_recon": ("org.kspacejet.gpu-recon",
                  ["kb_gridding_reconstruct", "kb_gridding_reconstruct_rss", "cartesian_ifft2_rss",
                   "coil_combine_sensitivity", "noise_prewhiten", "phase_correct", "coil_compress",
                   "readout_oversampling_remove", "noncartesian_noise_prewhiten", "noncartesian_coil_compress",
                   "sensitivity_map_estimate", "coil_combine_adaptive"], None),
}

DIRECT_CFG = '{"channels":2,"image_cols":2,"image_rows":2,"sample_count":2}'
RADIAL_CFG = ('{"channels":1,"density_compensation":"radial_analytic_ramp","image_cols":2,"image_rows":2,'
              '"sample_count":2,"trajectory_units":"radians_per_pixel"}')
TINY_RADIAL_SCRATCH = (2 * 4 + 2 * 2) * 8 + 2 * 4  # radial_gridding_reconstruct scratch of a 2x2 image


def bundle_digest(provider_id, op_ids):
    return hashlib.sha256(b"".join(s.encode() + b"\0" for s in ["kspacejet.provider-bundle", provider_id, *op_ids]))


@pytest.fixture(scope="module")
def modules():
    return {name: H.ProviderModule(name) for name in PROVIDERS}


# ---------------------------------------------------------------- ABI layout and types

def test_provider_header_layout_matches_the_reference_sdk():
    import abi_layout_probe as probe
    ref = json.load(open(os.path.join(GOLDEN, "abi_layout_reference.json")))
    ours = probe.probe(os.path.join(ROOT, "include", "ksj_provider_abi.h"))
    assert len(ref) == 33
    assert ours == ref


def test_python_host_structs_match_the_reference_layout():
    ref = json.load(open(os.path.join(GOLDEN, "abi_layout_reference.json")))
    for name, cls in H.STRUCTS.items():
        assert ctypes.sizeof(cls) == ref[name]["size"], name
        for field, offset in ref[name]["fields"].items():
            assert getattr(cls, field).offset == offset, (name, field)


def _canonical(v):
    return json.dumps(v, ensure_ascii=False, sort_keys=True, separators=(",", ":"), allow_nan=False)


def test_type_identity_digests_reproduce_the_reference_header():
    """Each digest recomputed from the registry identity equals the byte literal of
    the reference's generated type_registry.h and the table the modules use."""
    fixture = json.load(open(os.path.join(GOLDEN, "type_registry_reference.json")))
    table = {t["type_ref"]: t for t in json.load(open(H.TYPES_JSON))}
    assert len(fixture) == 12
    for entry in fixture:
        ident = entry["identity"]
        p, m = ident["payload"], ident["memory"]
        flat = {"type_ref": entry["type_ref"], "payload_kind": p["kind"], "element_type": p["element_type"],
                "rank": p["rank"], "dimensions": p["dimensions"], "layout": p["layout"], "strides": p["strides"],
                "explicit_byte_strides": p["explicit_byte_strides"], "allowed_memory_domains": m["allowed_domains"],
                "min_alignment_bytes": m["min_alignment_bytes"], "mutability": ident["mutability"]}
        digest = hashlib.sha256(b"kspacejet.type-identity\0" + _canonical(flat).encode()).hexdigest()
        assert digest == entry["header_digest"], entry["type_ref"]
        assert table[entry["type_ref"]]["digest"] == digest
        # A dropped field changes the identity: the rule is not satisfied by accident.
        flat.pop("mutability")
        assert hashlib.sha256(b"kspacejet.type-identity\0" + _canonical(flat).encode()).hexdigest() != digest


def test_type_table_include_matches_the_json_table():
    inc = open(os.path.join(H.PKG, "csrc", "provider", "types.inc")).read()
    for t in json.load(open(H.TYPES_JSON)):
        hexbytes = ", ".join(f"0x{b:02x}" for b in bytes.fromhex(t["digest"]))
        assert f'{{"{t["type_ref"]}", {{{hexbytes}}}' in inc


# ---------------------------------------------------------------- descriptors

def test_each_module_exports_only_the_query_symbol():
    for name in PROVIDERS:
        lib = os.path.join(ROOT, "kspacejet_amd", f"libksj_provider_{name}.so")
        out = subprocess.run(["nm", "-D", "--defined-only", lib], capture_output=True, text=True, check=True).stdout
        assert [l.split()[-1] for l in out.splitlines() if l.strip()] == ["ksj_provider_query"], name


def test_descriptors_publish_ids_bounds_and_bundle_digests(modules):
    for name, (pid, ops, ref_hex) in PROVIDERS.items():
        m = modules[name]
        assert m.provider_id == pid
        assert [H.view_text(o.operator_id) for o in m.operators] == ops
        digest = bundle_digest(pid, ops).digest()
        assert m.bundle_digest == digest
        if ref_hex:
            assert digest.hex() == ref_hex  # the reference's literal
        d = m.descriptor
        assert d.abi.capability_bits == (1 << 0) | (1 << 6) | (1 << 7) | (1 << 8)  # sync, no threads/file/net io
        for o in m.operators:
            assert o.max_in_flight == 1 and o.thread_safety == H.SERIAL_INSTANCE and o.max_private_threads == 0
            assert o.max_output_items_per_firing == 1 and o.max_retained_input_bytes == 0
            assert o.abi.capability_bits == (1 << 4) | (1 << 5)  # cancel: no allocation, no throw


def test_operator_bounds_equal_the_reference(modules):
    """noncartesian_recon_provider_tests.cpp:415-432, cartesian_recon_provider_tests.cpp:324-337,
    coil_combine_provider_tests.cpp:286-297."""
    direct, radial = modules["noncartesian_recon"].operators
    assert (direct.max_input_items_per_firing, direct.max_output_bytes_per_firing,
            direct.max_scratch_bytes_per_firing) == (2, 128 << 20, 2 << 20)
    assert (radial.max_input_items_per_firing, radial.max_output_bytes_per_firing,
            radial.max_scratch_bytes_per_firing) == (2, 128 << 20, 4464640)
    (cart,) = modules["cartesian_recon"].operators
    assert (cart.max_output_bytes_per_firing, cart.max_scratch_bytes_per_firing) == (
        128 << 20, (2 * 512 * 512 + 2 * 512) * 8)
    (rss,) = modules["coil_combine"].operators
    assert (rss.max_output_bytes_per_firing, rss.max_scratch_bytes_per_firing) == (1 << 20, 0)


def test_query_rejects_incompatible_abi_storage(modules):
    m = modules["coil_combine"]
    req = H.QueryRequest(H.header(H.QueryRequest), 0, H.utf8("host"))
    desc, api = H.ProviderDescriptor(), H.ProviderApi()
    desc.abi = H.header(H.ProviderDescriptor)
    api.abi = H.header(H.ProviderApi)
    api.abi.struct_size -= 8  # a host table one pointer short
    e = H.empty_error()
    assert m.lib.ksj_provider_query(ctypes.byref(req), ctypes.byref(desc), ctypes.byref(api),
                                    ctypes.byref(e)) == H.BAD_ABI
    api.abi = H.header(H.ProviderApi)
    req.abi.reserved0 = 1
    assert m.lib.ksj_provider_query(ctypes.byref(req), ctypes.byref(desc), ctypes.byref(api),
                                    ctypes.byref(e)) == H.BAD_ABI


# ---------------------------------------------------------------- config grammar

@pytest.mark.parametrize("module,op,config", [
    ("noncartesian_recon", "noncartesian_adjoint_reconstruct", DIRECT_CFG),
    ("noncartesian_recon", "radial_gridding_reconstruct", RADIAL_CFG),
    ("cartesian_recon", "cartesian_ifft2_coil_images", '{"channels":8,"cols":128,"rows":128}'),
    ("coil_combine", "coil_combine_rss", '{"channels":8,"cols":128,"rows":128}'),
    ("gpu_recon", "kb_gridding_reconstruct",
     '{"channels":16,"density_compensation":"radial_analytic_ramp","image_cols":256,"image_rows":256,'
     '"kernel_width":4,"oversampling":2,"sample_count":205824,"trajectory_units":"radians_per_pixel"}'),
    ("gpu_recon", "kb_gridding_reconstruct_rss",
     '{"channels":64,"density_compensation":"none","image_cols":512,"image_rows":512,'
     '"kernel_width":8,"oversampling":2,"sample_count":4194304,"trajectory_units":"radians_per_pixel"}'),
    ("gpu_recon", "cartesian_ifft2_rss", '{"centred":true,"channels":8,"cols":256,"rows":256}'),
    ("gpu_recon", "coil_combine_sensitivity", '{"channels":8,"cols":256,"rows":256}'),
    ("gpu_recon", "coil_combine_adaptive", '{"channels":8,"cols":256,"rows":256}'),
    ("gpu_recon", "sensitivity_map_estimate",
     '{"channels":16,"cols":256,"iterations":8,"method":"walsh","rows":256,"window":5}'),
    ("gpu_recon", "sensitivity_map_estimate",
     '{"channels":64,"cols":512,"iterations":64,"method":"walsh","rows":512,"window":31}'),
])
def test_canonical_configs_create_operators(modules, module, op, config):
    st, h, msg = modules[module].create(op, config)
    assert st == H.OK, msg
    h.close()


@pytest.mark.parametrize("module,op,config", [
    # noncartesian_recon_provider_tests.cpp:616-635
    ("noncartesian_recon", "noncartesian_adjoint_reconstruct",
     '{"channels":1,"image_rows":2,"image_cols":2,"sample_count":2}'),
    ("noncartesian_recon", "noncartesian_adjoint_reconstruct",
     '{"channels":64,"image_cols":512,"image_rows":512,"sample_count":65536}'),
    # noncartesian_recon_provider_tests.cpp:637-663
    ("noncartesian_recon", "radial_gridding_reconstruct",
     '{"channels":1,"image_cols":2,"image_rows":2,"sample_count":2,"trajectory_units":"radians_per_pixel"}'),
    ("noncartesian_recon", "radial_gridding_reconstruct",
     '{"channels":1,"density_compensation":"radial_analytic_ramp","image_cols":2,"image_rows":2,'
     '"sample_count":2,"trajectory_units":"cycles_per_pixel"}'),
    ("noncartesian_recon", "radial_gridding_reconstruct",
     '{"channels":1,"density_compensation":"radial_analytic_ramp","image_cols":3,"image_rows":2,'
     '"sample_count":2,"trajectory_units":"radians_per_pixel"}'),
    # grammar edges shared by all parsers
    ("cartesian_recon", "cartesian_ifft2_coil_images", '{"channels":08,"cols":128,"rows":128}'),
    ("cartesian_recon", "cartesian_ifft2_coil_images", '{"channels":0,"cols":128,"rows":128}'),
    ("cartesian_recon", "cartesian_ifft2_coil_images", '{"channels":65,"cols":128,"rows":128}'),
    ("cartesian_recon", "cartesian_ifft2_coil_images", '{"channels":8,"cols":1,"rows":128}'),
    ("cartesian_recon", "cartesian_ifft2_coil_images", '{"channels":8,"cols":513,"rows":128}'),
    ("cartesian_recon", "cartesian_ifft2_coil_images", '{"channels":8, "cols":128,"rows":128}'),
    ("cartesian_recon", "cartesian_ifft2_coil_images", '{"channels":8,"cols":128,"rows":128} '),
    ("cartesian_recon", "cartesian_ifft2_coil_images", '{"channels":8,"cols":128,"rows":128,"x":1}'),
    ("cartesian_recon", "cartesian_ifft2_coil_images", '{"channels":4294967304,"cols":128,"rows":128}'),
    ("cartesian_recon", "cartesian_ifft2_coil_images", '{"channels":+8,"cols":128,"rows":128}'),
    ("cartesian_recon", "cartesian_ifft2_coil_images", ''),
    ("coil_combine", "coil_combine_rss", '{"cols":128,"channels":8,"rows":128}'),
    ("gpu_recon", "kb_gridding_reconstruct",
     '{"channels":16,"density_compensation":"radial_analytic_ramp","image_cols":256,"image_rows":256,'
     '"kernel_width":5,"oversampling":2,"sample_count":205824,"trajectory_units":"radians_per_pixel"}'),
    ("gpu_recon", "kb_gridding_reconstruct",
     '{"channels":16,"density_compensation":"radial_analytic_ramp","image_cols":256,"image_rows":256,'
     '"kernel_width":4,"oversampling":3,"sample_count":205824,"trajectory_units":"radians_per_pixel"}'),
    ("gpu_recon", "kb_gridding_reconstruct",
     '{"channels":16,"density_compensation":"radial_analytic_rampx","image_cols":256,"image_rows":256,'
     '"kernel_width":4,"oversampling":2,"sample_count":205824,"trajectory_units":"radians_per_pixel"}'),
    ("gpu_recon", "cartesian_ifft2_rss", '{"centred":1,"channels":8,"cols":256,"rows":256}'),
    ("gpu_recon", "sensitivity_map_estimate",
     '{"channels":16,"cols":256,"iterations":8,"method":"walsh","rows":256,"window":4}'),
    ("gpu_recon", "sensitivity_map_estimate",
     '{"channels":16,"cols":256,"iterations":8,"method":"walsh","rows":256,"window":33}'),
    ("gpu_recon", "sensitivity_map_estimate",
     '{"channels":16,"cols":256,"iterations":0,"method":"walsh","rows":256,"window":5}'),
    ("gpu_recon", "sensitivity_map_estimate",
     '{"channels":16,"cols":256,"iterations":8,"method":"espirit","rows":256,"window":5}'),
    ("gpu_recon", "sensitivity_map_estimate",
     '{"channels":16,"cols":256,"iterations":8,"rows":256,"window":5}'),
    ("gpu_recon", "sensitivity_map_estimate",
     '{"channels":16,"cols":256,"method":"walsh","iterations":8,"rows":256,"window":5}'),
])
def test_noncanonical_or_out_of_envelope_configs_are_rejected(modules, module, op, config):
    st, h, _ = modules[module].create(op, config)
    assert st == H.INVALID_ARGUMENT and h is None


def test_unknown_operator_is_unsupported(modules):
    st, h, _ = modules["noncartesian_recon"].create("cartesian_ifft2_coil_images", '{"channels":1,"cols":2,"rows":2}')
    assert st == H.UNSUPPORTED and h is None


# ---------------------------------------------------------------- lifecycle

def test_lifecycle_admits_one_context_and_key_state(modules):
    st, op, _ = modules["noncartesian_recon"].create("noncartesian_adjoint_reconstruct", DIRECT_CFG)
    assert op.open() == H.OK
    api = modules["noncartesian_recon"].api
    d = H.ExecutionContextDescriptor()
    d.abi = H.header(H.ExecutionContextDescriptor)
    out = ctypes.c_void_p()
    e = H.empty_error()
    assert api.execution_context_create(op.handle, ctypes.byref(d), ctypes.byref(out),
                                        ctypes.byref(e)) == H.FAILED_PRECONDITION
    lease = _direct_lease()
    st, r = op.scan_end(lease, terminal_epoch=99)
    assert st == H.OK and r.outcome == H.PROCESS_DONE and r.terminal_epoch == 99 and r.sealed_output_count == 0
    assert op.scan_end(lease, kind=H.SCAN_END_CANCELLED)[0] == H.INVALID_ARGUMENT
    assert op.cancel(H.SCAN_END_FAILED) == H.OK
    assert op.cancel(H.SCAN_END_NORMAL) == H.INVALID_ARGUMENT
    op.close()


def test_process_without_context_is_an_invalid_argument(modules):
    st, op, _ = modules["coil_combine"].create("coil_combine_rss", '{"channels":2,"cols":2,"rows":2}')
    st, r, _ = op.process(_rss_lease())
    assert st == H.INVALID_ARGUMENT
    op.close()


# ---------------------------------------------------------------- firings rejected before output

def _direct_lease(channels=2, scratch=4 * 8, reverse=False):
    """The reference's HostState (noncartesian_recon_provider_tests.cpp:74-135)."""
    kspace = H.aligned_copy(np.array([1 + 2j, 3 - 1j, 10 - 4j, -2 + 7j], np.complex64)[:channels * 2])
    traj = H.aligned_copy(np.array([0.0, 0.0, math.pi, -0.5 * math.pi], np.float32))
    k = H.Item(kspace, "ksj.noncartesian-kspace-frame", 71, 73, metadata=b"123")
    t = H.Item(traj, "ksj.trajectory-frame", 71, 73)
    batches = [H.Batch(0, [k]), H.Batch(1, [t])]
    if reverse:
        batches.reverse()
    return H.Lease(batches, H.OutputSpec(channels * 4 * 8, "ksj.coil-image-frame"), scratch_bytes=scratch)


def _rss_lease():
    """coil_combine_provider_tests.cpp:62-70 (2x2 pixels, 2 coils)."""
    x = H.aligned_copy(np.array([3 + 4j, 0, 0, 12j, 5, 12j, 1 + 1j, 2j], np.complex64))
    return H.Lease([H.Batch(0, [H.Item(x, "ksj.coil-image-frame", 11, 17)])],
                   H.OutputSpec(16, "ksj.image-frame", np.float32), scratch_bytes=0)


def _cartesian_lease():
    """cartesian_recon_provider_tests.cpp:60-64: impulses 4 (coil 0) and 8i (coil 1) at DC."""
    x = H.aligned_empty(64, np.complex64)
    x[0], x[4] = 4, 8j
    return H.Lease([H.Batch(0, [H.Item(x, "ksj.kspace-frame", 11, 17)])],
                   H.OutputSpec(64, "ksj.coil-image-frame"), scratch_bytes=96)


def _opened(modules, module, op_id, config):
    st, op, msg = modules[module].create(op_id, config)
    assert st == H.OK, msg
    assert op.open() == H.OK
    return op


def _expect_rejected_before_output(op, lease, status=H.CONTRACT_VIOLATION):
    st, r, msg = op.process(lease)
    assert st == status, msg
    assert lease.acquire_calls == 0 and lease.seal_calls == 0 and lease.release_calls == 0
    assert r.sealed_output_count == 0
    return msg


def test_direct_rejects_wrong_trajectory_type(modules):
    op = _opened(modules, "noncartesian_recon", "noncartesian_adjoint_reconstruct", DIRECT_CFG)
    lease = _direct_lease()
    lease.batches[1].items[0].type = H.type_descriptor("ksj.coil-image-frame")
    lease._build()
    _expect_rejected_before_output(op, lease)
    op.close()


def test_direct_rejects_incorrect_trajectory_size(modules):
    op = _opened(modules, "noncartesian_recon", "noncartesian_adjoint_reconstruct", DIRECT_CFG)
    lease = _direct_lease()
    lease.batches[1].items[0].byte_count = 12
    lease._build()
    _expect_rejected_before_output(op, lease)
    op.close()


def test_direct_rejects_mismatched_identity(modules):
    op = _opened(modules, "noncartesian_recon", "noncartesian_adjoint_reconstruct", DIRECT_CFG)
    lease = _direct_lease()
    lease.batches[1].items[0].order_key = 74
    lease._build()
    _expect_rejected_before_output(op, lease)
    op.close()


def test_direct_rejects_missing_or_duplicate_second_input(modules):
    op = _opened(modules, "noncartesian_recon", "noncartesian_adjoint_reconstruct", DIRECT_CFG)
    lease = _direct_lease()
    lease.input_batch_count = 1
    _expect_rejected_before_output(op, lease)
    assert lease.input_batch_requests == 0
    lease = _direct_lease()
    lease.batches[1].port = 0
    lease._build()
    _expect_rejected_before_output(op, lease)
    op.close()


def test_direct_rejects_non_finite_trajectory_and_misaligned_payload(modules):
    op = _opened(modules, "noncartesian_recon", "noncartesian_adjoint_reconstruct", DIRECT_CFG)
    lease = _direct_lease()
    lease.batches[1].items[0].data[1] = np.nan
    _expect_rejected_before_output(op, lease)
    lease = _direct_lease()
    lease.batches[0].items[0].alignment = 32
    lease._build()
    _expect_rejected_before_output(op, lease)
    lease = _direct_lease()
    lease.batches[0].items[0].memory_domain = H.MEMORY_DEVICE
    lease._build()
    _expect_rejected_before_output(op, lease)
    op.close()


def test_direct_rejects_short_lease_reservations_and_missing_caps(modules):
    op = _opened(modules, "noncartesian_recon", "noncartesian_adjoint_reconstruct", DIRECT_CFG)
    lease = _direct_lease()
    lease.reserved_output_bytes = 63
    _expect_rejected_before_output(op, lease)
    lease = _direct_lease(scratch=31)
    _expect_rejected_before_output(op, lease)
    lease = H.Lease(_direct_lease().batches, H.OutputSpec(64, "ksj.coil-image-frame"), 32,
                    caps=H.LEASE_CAP_INPUT_BATCHES | H.LEASE_CAP_OUTPUT_GRANTS)
    _expect_rejected_before_output(op, lease, H.BAD_ABI)
    op.close()


def test_radial_rejects_out_of_range_trajectory_and_small_workspace(modules):
    """noncartesian_recon_provider_tests.cpp:665-687."""
    op = _opened(modules, "noncartesian_recon", "radial_gridding_reconstruct", RADIAL_CFG)
    lease = _direct_lease(channels=1, scratch=TINY_RADIAL_SCRATCH)
    lease.batches[1].items[0].data[0] = np.float32(math.pi) + np.float32(0.01)
    _expect_rejected_before_output(op, lease)
    lease = _direct_lease(channels=1, scratch=TINY_RADIAL_SCRATCH - 1)
    _expect_rejected_before_output(op, lease)
    op.close()


def test_invalid_output_grant_is_released_without_seal(modules):
    """noncartesian_recon_provider_tests.cpp:689-703."""
    op = _opened(modules, "noncartesian_recon", "noncartesian_adjoint_reconstruct", DIRECT_CFG)
    lease = _direct_lease()
    lease.output_spec.type = H.type_descriptor("ksj.noncartesian-kspace-frame")
    lease._build()
    st, r, _ = op.process(lease)
    assert st == H.CONTRACT_VIOLATION
    assert (lease.acquire_calls, lease.map_calls, lease.seal_calls, lease.release_calls) == (1, 1, 0, 1)
    assert r.sealed_output_count == 0
    op.close()


def test_cartesian_rejects_wrong_input_and_output_types(modules):
    """cartesian_recon_provider_tests.cpp:368-399."""
    op = _opened(modules, "cartesian_recon", "cartesian_ifft2_coil_images", '{"channels":2,"cols":2,"rows":2}')
    lease = _cartesian_lease()
    lease.batches[0].items[0].type = H.type_descriptor("ksj.coil-image-frame")
    lease._build()
    _expect_rejected_before_output(op, lease)
    lease = _cartesian_lease()
    lease.output_spec.type = H.type_descriptor("ksj.kspace-frame")
    lease._build()
    st, _, _ = op.process(lease)
    assert st == H.CONTRACT_VIOLATION and lease.seal_calls == 0 and lease.release_calls == 1
    # The Cartesian operator takes its scratch after mapping the output: a short scratch releases the grant.
    lease = _cartesian_lease()
    lease.scratch_bytes = 95
    lease.reserved_scratch_bytes = 96
    st, _, _ = op.process(lease)
    assert st == H.CONTRACT_VIOLATION and lease.acquire_calls == 1 and lease.release_calls == 1
    op.close()


def test_rss_rejects_wrong_input_and_output_types(modules):
    """coil_combine_provider_tests.cpp:322-353."""
    op = _opened(modules, "coil_combine", "coil_combine_rss", '{"channels":2,"cols":2,"rows":2}')
    lease = _rss_lease()
    lease.batches[0].items[0].type = H.type_descriptor("ksj.kspace-frame")
    lease._build()
    _expect_rejected_before_output(op, lease)
    lease = _rss_lease()
    lease.output_spec.type = H.type_descriptor("ksj.complex-image-frame")
    lease._build()
    st, _, _ = op.process(lease)
    assert st == H.CONTRACT_VIOLATION and lease.release_calls == 1 and lease.seal_calls == 0
    op.close()


def test_type_matcher_checks_every_identity_field(modules):
    op = _opened(modules, "coil_combine", "coil_combine_rss", '{"channels":2,"cols":2,"rows":2}')
    mutations = [
        lambda t: setattr(t, "element_type", 10),
        lambda t: setattr(t, "rank", 2),
        lambda t: setattr(t, "layout_flags", t.layout_flags ^ (1 << 1)),
        lambda t: setattr(t, "minimum_alignment", 32),
        lambda t: setattr(t, "mutability", 1),
        lambda t: setattr(t, "allowed_memory_domains", 3),
        lambda t: t.stride_bytes.__setitem__(0, 8),
        lambda t: t.type_identity_digest.bytes.__setitem__(31, t.type_identity_digest.bytes[31] ^ 1),
        lambda t: setattr(t, "type_ref", H.utf8("ksj.coil-image-framf")),
        lambda t: setattr(t, "dimension_names",
                          ctypes.cast((H.Utf8View * 3)(H.utf8("ky"), H.utf8("kx"), H.utf8("coil")),
                                      ctypes.POINTER(H.Utf8View))),
        lambda t: setattr(t, "reserved0", 1),
    ]
    for mutate in mutations:
        lease = _rss_lease()
        t = H.type_descriptor("ksj.coil-image-frame")
        mutate(t)
        lease.batches[0].items[0].type = t
        lease._build()
        _expect_rejected_before_output(op, lease)
    op.close()


def test_kb_operator_needs_no_scratch_and_rejects_out_of_range_trajectory(modules):
    cfg = ('{"channels":1,"density_compensation":"radial_analytic_ramp","image_cols":8,"image_rows":8,'
           '"kernel_width":4,"oversampling":2,"sample_count":2,"trajectory_units":"radians_per_pixel"}')
    op = _opened(modules, "gpu_recon", "kb_gridding_reconstruct", cfg)
    lease = _direct_lease(channels=1, scratch=0)
    lease.output_spec = H.OutputSpec(64 * 8, "ksj.coil-image-frame")
    lease.output = H.aligned_empty(64 * 8, np.complex64)
    lease.caps = H.LEASE_CAP_INPUT_BATCHES | H.LEASE_CAP_OUTPUT_GRANTS
    lease.batches[1].items[0].data[3] = -4.0
    lease._build()
    _expect_rejected_before_output(op, lease)
    op.close()


def test_valid_firing_without_a_device_fails_loudly_and_releases_the_grant(modules):
    """On a host with no GPU the operator must not fall back to CPU work: the
    firing fails after the grant is acquired, and the grant is released."""
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    op = _opened(modules, "coil_combine", "coil_combine_rss", '{"channels":2,"cols":2,"rows":2}')
    lease = _rss_lease()
    st, r, msg = op.process(lease)
    assert st in (H.INTERNAL_ERROR, H.RESOURCE_EXHAUSTED), msg
    assert lease.acquire_calls == 1 and lease.seal_calls == 0 and lease.release_calls == 1
    assert np.all(lease.output == 0)
    op.close()


@pytest.mark.parametrize("module,op_id,config,ports", [
    ("coil_combine", "coil_combine_rss", '{"channels":2,"cols":2,"rows":2}', None),
    ("gpu_recon", "cartesian_ifft2_rss", '{"centred":false,"channels":2,"cols":2,"rows":2}', None)])
def test_device_label_on_host_memory_is_a_contract_violation(modules, module, op_id, config, ports):
    """The reference-named modules take host-pageable payloads only; the gpu-recon
    operators also take device-resident ones, but only memory that really is device
    memory of their GPU.  A host array labelled KSJ_PROVIDER_MEMORY_DEVICE is rejected
    before any device work (so on a host without a GPU too)."""
    op = _opened(modules, module, op_id, config)
    x = H.aligned_copy(np.ones((2, 2, 2), np.complex64))
    type_ref = "ksj.coil-image-frame" if module == "coil_combine" else "ksj.kspace-frame"
    lease = H.Lease([H.Batch(0, [H.Item(x, type_ref, memory_domain=H.MEMORY_DEVICE)])],
                    H.OutputSpec(2 * 2 * 4, "ksj.image-frame", np.float32))
    st, r, msg = op.process(lease)
    assert st == H.CONTRACT_VIOLATION and "payload" in msg, msg
    assert lease.seal_calls == 0
    op.close()
