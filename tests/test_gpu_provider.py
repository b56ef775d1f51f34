"""GPU parity through the Provider ABI: each operator of the provider modules,
driven by the Python host exactly as a KSpaceJet host drives a module, against
the FP64 oracle and the reference's own known answers.

The tiny cases are those of the reference's provider tests
(tests/unit/providers/kspacejet-noncartesian-recon/noncartesian_recon_provider_tests.cpp:434-545,
 kspacejet-cartesian-recon/cartesian_recon_provider_tests.cpp:339-366,
 kspacejet-coil-combine/coil_combine_provider_tests.cpp:299-320), with their 1e-5
absolute tolerance.  The larger cases use the rel-L2 <= 1e-5 bar.
"""
import math

import numpy as np
import pytest

import oracle as O
from kspacejet_amd import provider_host as H
from kspacejet_amd import synth
from tests.helpers import per_coil, rel_l2

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

TOL = 1e-5


@pytest.fixture(scope="module")
def modules():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return {name: H.ProviderModule(name) for name in ("noncartesian_recon", "cartesian_recon", "coil_combine",
                                                      "gpu_recon")}


def opened(modules, module, op_id, config):
    st, op, msg = modules[module].create(op_id, config)
    assert st == H.OK, msg
    assert op.open(0) == H.OK
    return op


def fire(op, lease, consumed):
    st, r, msg = op.process(lease)
    assert st == H.OK, msg
    assert r.outcome == H.PROCESS_DONE and r.sealed_output_count == 1 and r.consumed_input_item_count == consumed
    assert r.terminal_epoch == lease.terminal_epoch
    assert lease.seal_calls == 1 and lease.release_calls == 0
    assert lease.sealed.produced_item_count == 1 and lease.sealed.produced_byte_count == lease.output.nbytes
    return lease.output


def noncartesian_lease(kspace, traj, out_type, out_nbytes, out_dtype=np.complex64, scratch=0, reverse=False,
                       keys=(71, 73), metadata=b"123"):
    k = H.Item(H.aligned_copy(np.asarray(kspace, np.complex64)), "ksj.noncartesian-kspace-frame", *keys,
               metadata=metadata)
    t = H.Item(H.aligned_copy(np.asarray(traj, np.float32)), "ksj.trajectory-frame", *keys)
    batches = [H.Batch(0, [k]), H.Batch(1, [t])]
    if reverse:
        batches.reverse()
    return H.Lease(batches, H.OutputSpec(out_nbytes, out_type, out_dtype), scratch_bytes=scratch)


# The reference's tiny two-sample frame (noncartesian_recon_provider_tests.cpp:81-82).
TINY_K = np.array([1 + 2j, 3 - 1j, 10 - 4j, -2 + 7j], np.complex64)
TINY_T = np.array([0.0, 0.0, math.pi, -0.5 * math.pi], np.float32)


# ---------------------------------------------------------------- reference KATs

def test_direct_adjoint_tiny_channel(modules):
    op = opened(modules, "noncartesian_recon", "noncartesian_adjoint_reconstruct",
                '{"channels":1,"image_cols":2,"image_rows":2,"sample_count":2}')
    lease = noncartesian_lease(TINY_K[:2], TINY_T, "ksj.coil-image-frame", 4 * 8, scratch=4 * 8)
    out = fire(op, lease, 2).reshape(2, 2)
    expected = O.nudft2_adjoint(TINY_K[:2], TINY_T.astype(np.float64), 2, 2, 0.5, 0.5)
    np.testing.assert_allclose(out, expected, atol=1e-5)
    assert (lease.sealed.semantic_key_hash, lease.sealed.order_key) == (71, 73)
    assert lease.sealed_metadata == b"123"  # the k-space item's metadata travels with the image
    op.close()


def test_radial_gridding_tiny_channel(modules):
    op = opened(modules, "noncartesian_recon", "radial_gridding_reconstruct",
                '{"channels":1,"density_compensation":"radial_analytic_ramp","image_cols":2,"image_rows":2,'
                '"sample_count":2,"trajectory_units":"radians_per_pixel"}')
    lease = noncartesian_lease(TINY_K[:2], TINY_T, "ksj.coil-image-frame", 4 * 8, scratch=(2 * 4 + 2 * 2) * 8 + 8)
    out = fire(op, lease, 2).reshape(2, 2)
    t = TINY_T.astype(np.float64)
    expected = O.radial_linear_gridding2(TINY_K[:2], t, O.ramp_dcf2(t), 2, 2, 0.5, 0.5)
    np.testing.assert_allclose(out, expected, atol=1e-5)
    op.close()


def test_direct_adjoint_emits_every_channel_in_coil_last_order(modules):
    op = opened(modules, "noncartesian_recon", "noncartesian_adjoint_reconstruct",
                '{"channels":2,"image_cols":2,"image_rows":2,"sample_count":2}')
    lease = noncartesian_lease(TINY_K, TINY_T, "ksj.coil-image-frame", 2 * 4 * 8, scratch=4 * 8)
    out = fire(op, lease, 2).reshape(2, 2, 2)
    t = TINY_T.astype(np.float64)
    for c in range(2):
        np.testing.assert_allclose(out[..., c], O.nudft2_adjoint(TINY_K[2 * c:2 * c + 2], t, 2, 2, 0.5, 0.5),
                                   atol=1e-5)
    op.close()


def test_inputs_are_resolved_by_port_not_batch_index(modules):
    op = opened(modules, "noncartesian_recon", "noncartesian_adjoint_reconstruct",
                '{"channels":2,"image_cols":2,"image_rows":2,"sample_count":2}')
    a = fire(op, noncartesian_lease(TINY_K, TINY_T, "ksj.coil-image-frame", 64, scratch=32), 2).copy()
    lease = noncartesian_lease(TINY_K, TINY_T, "ksj.coil-image-frame", 64, scratch=32, reverse=True)
    b = fire(op, lease, 2)
    assert np.array_equal(a, b)
    assert (lease.sealed.semantic_key_hash, lease.sealed.order_key) == (71, 73)
    op.close()


def test_cartesian_tiny_impulses(modules):
    """Impulses 4 and 8i at DC of a 2x2, 2-coil frame give 1 and 2i everywhere, coil-last."""
    op = opened(modules, "cartesian_recon", "cartesian_ifft2_coil_images", '{"channels":2,"cols":2,"rows":2}')
    x = H.aligned_empty(64, np.complex64)
    x[0], x[4] = 4, 8j
    lease = H.Lease([H.Batch(0, [H.Item(x, "ksj.kspace-frame", 11, 17, metadata=b"m")])],
                    H.OutputSpec(64, "ksj.coil-image-frame"), scratch_bytes=96)
    out = fire(op, lease, 1).reshape(4, 2)
    np.testing.assert_allclose(out[:, 0], 1.0, atol=1e-5)
    np.testing.assert_allclose(out[:, 1], 2j, atol=1e-5)
    assert (lease.sealed.semantic_key_hash, lease.sealed.order_key) == (11, 17)
    assert lease.sealed_metadata == b""  # the Cartesian operator seals empty metadata
    op.close()


def test_rss_tiny(modules):
    op = opened(modules, "coil_combine", "coil_combine_rss", '{"channels":2,"cols":2,"rows":2}')
    x = H.aligned_copy(np.array([3 + 4j, 0, 0, 12j, 5, 12j, 1 + 1j, 2j], np.complex64))
    lease = H.Lease([H.Batch(0, [H.Item(x, "ksj.coil-image-frame", 11, 17)])],
                    H.OutputSpec(16, "ksj.image-frame", np.float32))
    out = fire(op, lease, 1)
    np.testing.assert_allclose(out, [5, 12, 13, math.sqrt(6)], atol=1e-5)
    op.close()


# ---------------------------------------------------------------- full-size parity

def test_cartesian_config1_shape(modules):
    """256x256, 8 coils: uncentred ifft2 per coil, coil-last, against the FP64 oracle."""
    n, coils = 256, 8
    k = synth.cartesian_kspace2(n, coils, synth.scan_seed(1, 0), noise=True).astype(np.complex64)
    op = opened(modules, "cartesian_recon", "cartesian_ifft2_coil_images",
                f'{{"channels":{coils},"cols":{n},"rows":{n}}}')
    lease = H.Lease([H.Batch(0, [H.Item(H.aligned_copy(k), "ksj.kspace-frame")])],
                    H.OutputSpec(n * n * coils * 8, "ksj.coil-image-frame"),
                    scratch_bytes=(2 * n * n + 2 * n) * 8)
    out = fire(op, lease, 1).reshape(n, n, coils)
    assert rel_l2(out, O.ifft2_coil_images(k)) <= TOL
    op.close()


@pytest.mark.parametrize("centred", [False, True])
def test_cartesian_ifft2_rss(modules, centred):
    n, coils = 128, 8
    k = synth.cartesian_kspace2(n, coils, synth.scan_seed(1, 3), noise=True).astype(np.complex64)
    op = opened(modules, "gpu_recon", "cartesian_ifft2_rss",
                f'{{"centred":{"true" if centred else "false"},"channels":{coils},"cols":{n},"rows":{n}}}')
    lease = H.Lease([H.Batch(0, [H.Item(H.aligned_copy(k), "ksj.kspace-frame")])],
                    H.OutputSpec(n * n * 4, "ksj.image-frame", np.float32))
    out = fire(op, lease, 1).reshape(n, n)
    assert rel_l2(out, O.rss(O.ifft2_coil_images(k, centred))) <= TOL
    op.close()


def golden_scan(n, coils, spokes, readout, seed):
    """Golden-angle radial scan: FP64 trajectory rounded to float, k-space from the exact
    forward model of the seeded phantom x coil maps (direct NUDFT, FP64 on the host)."""
    traj = O.golden_angle2(spokes, readout).astype(np.float32)
    img = synth.coil_images2(n, coils, seed)
    t64 = traj.astype(np.float64)
    k = np.stack(per_coil(lambda c: O.nudft2_forward(img[c], t64, (n - 1) / 2, (n - 1) / 2), coils))
    return traj, synth.add_noise(k, seed).astype(np.complex64)


def test_radial_gridding_at_the_reference_example_shape(modules):
    """radial_gridding_reconstruct at the operator's documented example config
    (128x128, 8 channels, 4096 samples) against the FP64 oracle of the reference route."""
    n, coils, spokes, readout = 128, 8, 16, 256
    traj, k = golden_scan(n, coils, spokes, readout, synth.scan_seed(2, 7))
    S = spokes * readout
    op = opened(modules, "noncartesian_recon", "radial_gridding_reconstruct",
                f'{{"channels":{coils},"density_compensation":"radial_analytic_ramp","image_cols":{n},'
                f'"image_rows":{n},"sample_count":{S},"trajectory_units":"radians_per_pixel"}}')
    lease = noncartesian_lease(k, traj, "ksj.coil-image-frame", n * n * coils * 8,
                               scratch=(2 * n * n + 2 * n) * 8 + S * 4)
    out = fire(op, lease, 2).reshape(n, n, coils)
    t = traj.astype(np.float64)
    dcf = O.ramp_dcf2(t)
    for c in range(coils):
        ref = O.radial_linear_gridding2(k[c], t, dcf, n, n, (n - 1) / 2, (n - 1) / 2)
        assert rel_l2(out[..., c], ref) <= TOL
    op.close()


def test_direct_adjoint_parity(modules):
    rows, cols, coils, S = 32, 24, 4, 1500
    rng = np.random.default_rng(11)
    traj = rng.uniform(-np.pi, np.pi, (S, 2)).astype(np.float32)
    k = (rng.standard_normal((coils, S)) + 1j * rng.standard_normal((coils, S))).astype(np.complex64)
    op = opened(modules, "noncartesian_recon", "noncartesian_adjoint_reconstruct",
                f'{{"channels":{coils},"image_cols":{cols},"image_rows":{rows},"sample_count":{S}}}')
    lease = noncartesian_lease(k, traj, "ksj.coil-image-frame", rows * cols * coils * 8, scratch=rows * cols * 8)
    out = fire(op, lease, 2).reshape(rows, cols, coils)
    t = traj.astype(np.float64)
    for c in range(coils):
        assert rel_l2(out[..., c], O.nudft2_adjoint(k[c], t, rows, cols, (rows - 1) / 2, (cols - 1) / 2)) <= TOL
    op.close()


def kb_config(coils, n, S, width=4, density="radial_analytic_ramp"):
    return (f'{{"channels":{coils},"density_compensation":"{density}","image_cols":{n},"image_rows":{n},'
            f'"kernel_width":{width},"oversampling":2,"sample_count":{S},"trajectory_units":"radians_per_pixel"}}')


def test_kb_gridding_images_and_rss_match_the_oracle(modules):
    n, coils, spokes, readout = 64, 4, 51, 128
    traj, k = golden_scan(n, coils, spokes, readout, synth.scan_seed(2, 9))
    S = spokes * readout
    t = traj.astype(np.float64)
    ref = O.kb_adjoint2_coils(k, t, O.ramp_dcf2(t), n)
    op = opened(modules, "gpu_recon", "kb_gridding_reconstruct", kb_config(coils, n, S))
    out = fire(op, noncartesian_lease(k, traj, "ksj.coil-image-frame", n * n * coils * 8), 2).reshape(n, n, coils)
    assert rel_l2(out, ref) <= TOL
    op_rss = opened(modules, "gpu_recon", "kb_gridding_reconstruct_rss", kb_config(coils, n, S))
    rss = fire(op_rss, noncartesian_lease(k, traj, "ksj.image-frame", n * n * 4, np.float32), 2).reshape(n, n)
    assert rel_l2(rss, O.rss(ref)) <= TOL
    op.close()
    op_rss.close()


def test_kb_resident_plan_follows_the_trajectory(modules):
    """Firings on one operator keep the plan of the last trajectory: the same
    trajectory reproduces the first result bit for bit, and a new trajectory is
    picked up (its result matches the oracle of that trajectory)."""
    n, coils = 48, 3
    traj1, k1 = golden_scan(n, coils, 31, 96, synth.scan_seed(2, 21))
    rng = np.random.default_rng(5)
    traj2 = rng.uniform(-np.pi, np.pi, traj1.shape).astype(np.float32)
    op = opened(modules, "gpu_recon", "kb_gridding_reconstruct", kb_config(coils, n, traj1.shape[0], 6))
    first = fire(op, noncartesian_lease(k1, traj1, "ksj.coil-image-frame", n * n * coils * 8), 2).copy()
    again = fire(op, noncartesian_lease(k1, traj1, "ksj.coil-image-frame", n * n * coils * 8), 2).copy()
    other = fire(op, noncartesian_lease(k1, traj2, "ksj.coil-image-frame", n * n * coils * 8), 2).reshape(n, n, coils)
    assert np.array_equal(first, again)
    t2 = traj2.astype(np.float64)
    assert rel_l2(other, O.kb_adjoint2_coils(k1, t2, O.ramp_dcf2(t2), n, width=6)) <= TOL
    back = fire(op, noncartesian_lease(k1, traj1, "ksj.coil-image-frame", n * n * coils * 8), 2)
    assert np.array_equal(first, back)
    op.close()


def test_kb_without_density_compensation(modules):
    n, coils = 32, 2
    rng = np.random.default_rng(8)
    S = 900
    traj = rng.uniform(-np.pi, np.pi, (S, 2)).astype(np.float32)
    k = (rng.standard_normal((coils, S)) + 1j * rng.standard_normal((coils, S))).astype(np.complex64)
    op = opened(modules, "gpu_recon", "kb_gridding_reconstruct", kb_config(coils, n, S, 8, "none"))
    out = fire(op, noncartesian_lease(k, traj, "ksj.coil-image-frame", n * n * coils * 8), 2).reshape(n, n, coils)
    assert rel_l2(out, O.kb_adjoint2_coils(k, traj.astype(np.float64), None, n, width=8)) <= TOL
    op.close()


def test_sensitivity_combine(modules):
    """With I_c = S_c x the combine returns x; against the FP64 oracle on noisy input."""
    n, coils = 64, 6
    maps = synth.coil_maps2(n, coils, synth.scan_seed(1, 1)).transpose(1, 2, 0)  # coil-last
    x = synth.phantom2(n, synth.scan_seed(1, 1)).astype(np.complex128) * np.exp(0.3j)
    imgs = (maps * x[..., None]).astype(np.complex64)
    op = opened(modules, "gpu_recon", "coil_combine_sensitivity", f'{{"channels":{coils},"cols":{n},"rows":{n}}}')
    lease = H.Lease([H.Batch(0, [H.Item(H.aligned_copy(imgs), "ksj.coil-image-frame", 5, 6)]),
                     H.Batch(1, [H.Item(H.aligned_copy(maps.astype(np.complex64)), "ksj.sensitivity-map", 1, 2)])],
                    H.OutputSpec(n * n * 8, "ksj.complex-image-frame"))
    out = fire(op, lease, 2).reshape(n, n)
    assert rel_l2(out, x) <= TOL
    assert rel_l2(out, O.sense_combine(imgs, maps.astype(np.complex64))) <= TOL
    assert (lease.sealed.semantic_key_hash, lease.sealed.order_key) == (5, 6)
    op.close()
