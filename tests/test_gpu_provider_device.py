"""Device-memory edges (KSJ_PROVIDER_MEMORY_DEVICE) of the org.kspacejet.gpu-recon
operators: k-space, trajectory, images and maps handed over in HBM of the operator's
GPU, and output grants in HBM, give the same bits as the host-pageable firing in every
combination.  The reference-named modules keep the reference's host-only contract, and
a device-labelled payload that is not device memory of the operator's GPU is a
contract violation."""
import numpy as np
import pytest

from kspacejet_amd import provider_host as H
from kspacejet_amd import synth

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def modules():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return {name: H.ProviderModule(name) for name in ("noncartesian_recon", "gpu_recon")}


def opened(modules, module, op_id, config):
    st, op, msg = modules[module].create(op_id, config)
    assert st == H.OK, msg
    assert op.open(0) == H.OK
    return op


def item(data, type_ref, device, keys=(71, 73)):
    if device:
        return H.Item(torch.from_numpy(np.ascontiguousarray(data)).cuda(), type_ref, *keys,
                      memory_domain=H.MEMORY_DEVICE)
    return H.Item(H.aligned_copy(data), type_ref, *keys)


def result(op, lease):
    st, r, msg = op.process(lease)
    assert st == H.OK, msg
    assert r.outcome == H.PROCESS_DONE and lease.seal_calls == 1
    torch.cuda.synchronize()
    out = lease.output
    return out.cpu().numpy().tobytes() if hasattr(out, "data_ptr") else out.tobytes()


def kb_config(coils, n, S):
    return (f'{{"channels":{coils},"density_compensation":"radial_analytic_ramp","image_cols":{n},"image_rows":{n},'
            f'"kernel_width":4,"oversampling":2,"sample_count":{S},"trajectory_units":"radians_per_pixel"}}')


@pytest.mark.parametrize("op_id,out_type,out_bytes", [
    ("kb_gridding_reconstruct", "ksj.coil-image-frame", lambda n, c: n * n * c * 8),
    ("kb_gridding_reconstruct_rss", "ksj.image-frame", lambda n, c: n * n * 4)])
def test_kb_device_edges_give_the_host_bits(modules, op_id, out_type, out_bytes):
    n, coils, spokes, readout = 64, 16, 51, 128
    traj = synth.golden_angle2(spokes, readout).astype(np.float32)
    rng = np.random.default_rng(3)
    k = (rng.standard_normal((coils, spokes * readout)) + 1j * rng.standard_normal((coils, spokes * readout)))
    k = k.astype(np.complex64)
    op = opened(modules, "gpu_recon", op_id, kb_config(coils, n, spokes * readout))
    results = []
    for k_dev, t_dev, o_dev in [(False, False, False), (True, True, True), (True, False, False), (False, True, True),
                                (True, True, False)]:
        lease = H.Lease([H.Batch(0, [item(k, "ksj.noncartesian-kspace-frame", k_dev)]),
                         H.Batch(1, [item(traj, "ksj.trajectory-frame", t_dev)])],
                        H.OutputSpec(out_bytes(n, coils), out_type,
                                     memory_domain=H.MEMORY_DEVICE if o_dev else H.MEMORY_HOST_PAGEABLE))
        results.append(result(op, lease))
    assert all(r == results[0] for r in results[1:])
    op.close()


@pytest.mark.parametrize("centred", ["false", "true"])
def test_cartesian_rss_device_edges_give_the_host_bits(modules, centred):
    n, coils = 96, 8
    rng = np.random.default_rng(4)
    k = (rng.standard_normal((coils, n, n)) + 1j * rng.standard_normal((coils, n, n))).astype(np.complex64)
    op = opened(modules, "gpu_recon", "cartesian_ifft2_rss", f'{{"centred":{centred},"channels":{coils},"cols":{n},'
                                                            f'"rows":{n}}}')
    results = []
    for dev in (False, True):
        lease = H.Lease([H.Batch(0, [item(k, "ksj.kspace-frame", dev)])],
                        H.OutputSpec(n * n * 4, "ksj.image-frame", np.float32,
                                     memory_domain=H.MEMORY_DEVICE if dev else H.MEMORY_HOST_PAGEABLE))
        results.append(result(op, lease))
    assert results[0] == results[1]
    op.close()


def test_sensitivity_combine_device_edges_give_the_host_bits(modules):
    n, coils = 64, 6
    maps = synth.coil_maps2(n, coils, synth.scan_seed(1, 1)).transpose(1, 2, 0).astype(np.complex64)
    x = synth.phantom2(n, synth.scan_seed(1, 1)).astype(np.complex128) * np.exp(0.3j)
    imgs = (maps * x[..., None]).astype(np.complex64)
    op = opened(modules, "gpu_recon", "coil_combine_sensitivity", f'{{"channels":{coils},"cols":{n},"rows":{n}}}')
    results = []
    for dev in (False, True):
        lease = H.Lease([H.Batch(0, [item(imgs, "ksj.coil-image-frame", dev, (5, 6))]),
                         H.Batch(1, [item(maps, "ksj.sensitivity-map", dev, (1, 2))])],
                        H.OutputSpec(n * n * 8, "ksj.complex-image-frame",
                                     memory_domain=H.MEMORY_DEVICE if dev else H.MEMORY_HOST_PAGEABLE))
        results.append(result(op, lease))
    assert results[0] == results[1]
    op.close()


def test_reference_modules_keep_the_host_only_contract(modules):
    n, coils, S = 16, 2, 64
    rng = np.random.default_rng(6)
    traj = rng.uniform(-3, 3, (S, 2)).astype(np.float32)
    k = (rng.standard_normal((coils, S)) + 1j * rng.standard_normal((coils, S))).astype(np.complex64)
    config = (f'{{"channels":{coils},"density_compensation":"radial_analytic_ramp","image_cols":{n},'
              f'"image_rows":{n},"sample_count":{S},"trajectory_units":"radians_per_pixel"}}')
    op = opened(modules, "noncartesian_recon", "radial_gridding_reconstruct", config)
    lease = H.Lease([H.Batch(0, [item(k, "ksj.noncartesian-kspace-frame", True)]),
                     H.Batch(1, [item(traj, "ksj.trajectory-frame", False)])],
                    H.OutputSpec(n * n * coils * 8, "ksj.coil-image-frame"),
                    scratch_bytes=(2 * n * n + 2 * n) * 8 + S * 4)
    st, _, _ = op.process(lease)
    assert st == H.CONTRACT_VIOLATION
    op.close()


def test_device_label_on_host_memory_is_rejected(modules):
    n, coils, spokes, readout = 32, 2, 21, 64
    traj = synth.golden_angle2(spokes, readout).astype(np.float32)
    k = np.ones((coils, spokes * readout), np.complex64)
    op = opened(modules, "gpu_recon", "kb_gridding_reconstruct_rss", kb_config(coils, n, spokes * readout))
    bad = H.Item(H.aligned_copy(k), "ksj.noncartesian-kspace-frame", 71, 73, memory_domain=H.MEMORY_DEVICE)
    lease = H.Lease([H.Batch(0, [bad]), H.Batch(1, [item(traj, "ksj.trajectory-frame", False)])],
                    H.OutputSpec(n * n * 4, "ksj.image-frame", np.float32))
    st, _, _ = op.process(lease)
    assert st == H.CONTRACT_VIOLATION
    op.close()
