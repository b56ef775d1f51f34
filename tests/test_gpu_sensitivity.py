"""Sensitivity maps on the GPU (Walsh's adaptive method, ksj_gpu_sensitivity_maps_walsh)
and the adaptive combine, for the planned interfaces sensitivity_map_estimate and
coil_combine_adaptive (no reference implementation; the oracle is pinned by
numpy.linalg.eigh and closed forms in test_oracle_sensitivity.py):

* maps against the FP64 oracle of the same algorithm (rel-L2 <= 1e-5) at several image
  shapes, coil counts (1 .. 64, so every lane-group width) and windows, including
  non-square images and windows wider than the image;
* a constant complex gain per coil recovered to FP32 rounding, repeatable bit for bit;
* the chain coil images -> sensitivity_map_estimate -> coil_combine_adaptive through the
  Provider ABI against the oracle's maps and combine, over host and device edges."""
import numpy as np
import pytest

import oracle as O
from kspacejet_amd import provider_host as H
from kspacejet_amd import synth
from tests.helpers import rel_l2

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

TOL = 1e-5


@pytest.fixture(scope="module")
def ops():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from kspacejet_amd import ops as K
    K.init(0)
    return K


def coil_images(rows, cols, coils, seed):
    """Coil-last images of the ellipse phantom times Gaussian coil maps on a rows x cols grid."""
    n = max(rows, cols)
    img = synth.coil_images2(n, coils, synth.scan_seed(7, seed)).transpose(1, 2, 0)  # [n, n, C]
    return np.ascontiguousarray(img[:rows, :cols]).astype(np.complex64)


@pytest.mark.parametrize("rows,cols,coils,window,iterations", [
    (48, 40, 5, 5, 8),
    (64, 64, 16, 7, 8),
    (33, 47, 32, 3, 12),
    (24, 24, 64, 5, 6),
    (20, 28, 1, 5, 4),
    (16, 18, 3, 1, 3),
    (6, 5, 8, 31, 8),
])
def test_maps_match_the_oracle(ops, rows, cols, coils, window, iterations):
    img = coil_images(rows, cols, coils, rows * 100 + coils)
    got = ops.sensitivity_maps_walsh(torch.from_numpy(img).cuda(), window, iterations).cpu().numpy()
    ref = O.walsh_maps(img, window, iterations)
    err = rel_l2(got, ref)
    print(f"{rows}x{cols} coils {coils} window {window} iterations {iterations}: rel-L2 {err:.2e}")
    assert err <= TOL


def test_constant_gains_are_recovered_and_repeatable(ops):
    rng = np.random.default_rng(5)
    gains = (rng.standard_normal(12) + 1j * rng.standard_normal(12))
    obj = synth.phantom2(40, synth.scan_seed(7, 1)) + 0.1
    img = (obj[..., None] * gains).astype(np.complex64)
    t = torch.from_numpy(img).cuda()
    a = ops.sensitivity_maps_walsh(t, 5, 4)
    b = ops.sensitivity_maps_walsh(t, 5, 4)
    assert torch.equal(a, b)
    want = gains / np.linalg.norm(gains) * np.conj(gains[0]) / abs(gains[0])
    assert np.max(np.abs(a.cpu().numpy() - want)) < 2e-6


def test_invalid_arguments_are_refused(ops):
    img = torch.zeros((8, 8, 4), dtype=torch.complex64, device="cuda")
    from kspacejet_amd._lib import KsjGpuError
    for window, iterations in ((4, 2), (33, 2), (5, 65)):
        with pytest.raises(KsjGpuError):
            ops.sensitivity_maps_walsh(img, window, iterations)


def smap_config(coils, rows, cols, window, iterations):
    return (f'{{"channels":{coils},"cols":{cols},"iterations":{iterations},"method":"walsh","rows":{rows},'
            f'"window":{window}}}')


def item(data, type_ref, device, keys):
    """A payload on the host or in HBM (complete before the firing: the operator runs on its
    own stream)."""
    if device:
        t = data if isinstance(data, torch.Tensor) else torch.from_numpy(np.ascontiguousarray(data)).cuda()
        torch.cuda.synchronize()
        return H.Item(t, type_ref, *keys, memory_domain=H.MEMORY_DEVICE)
    return H.Item(H.aligned_copy(np.asarray(data)), type_ref, *keys)


def fire(op, lease):
    st, r, msg = op.process(lease)
    assert st == H.OK, msg
    assert r.outcome == H.PROCESS_DONE and r.sealed_output_count == 1 and lease.seal_calls == 1
    torch.cuda.synchronize()
    return lease.output


@pytest.mark.parametrize("device_edges", [False, True])
def test_provider_chain_estimate_then_adaptive_combine(ops, device_edges):
    rows, cols, coils, window, iterations = 64, 56, 12, 5, 8
    img = coil_images(rows, cols, coils, 3)
    mod = H.ProviderModule("gpu_recon")
    st, est, msg = mod.create("sensitivity_map_estimate", smap_config(coils, rows, cols, window, iterations))
    assert st == H.OK, msg
    assert est.open(0) == H.OK
    st, comb, msg = mod.create("coil_combine_adaptive", f'{{"channels":{coils},"cols":{cols},"rows":{rows}}}')
    assert st == H.OK, msg
    assert comb.open(0) == H.OK
    domain = H.MEMORY_DEVICE if device_edges else H.MEMORY_HOST_PAGEABLE
    lease = H.Lease([H.Batch(0, [item(img, "ksj.coil-image-frame", device_edges, (5, 6))])],
                    H.OutputSpec(img.nbytes, "ksj.sensitivity-map", memory_domain=domain))
    maps_out = fire(est, lease)
    assert (lease.sealed.semantic_key_hash, lease.sealed.order_key) == (5, 6)
    if device_edges:
        maps = maps_out.view(torch.complex64).reshape(rows, cols, coils)
        maps_np = maps.cpu().numpy()
    else:
        maps_np = maps_out.view(np.complex64).reshape(rows, cols, coils)
        maps = maps_np
    ref_maps = O.walsh_maps(img, window, iterations)
    assert rel_l2(maps_np, ref_maps) <= TOL
    lease2 = H.Lease([H.Batch(0, [item(img, "ksj.coil-image-frame", device_edges, (5, 6))]),
                      H.Batch(1, [item(maps, "ksj.sensitivity-map", device_edges, (1, 2))])],
                     H.OutputSpec(rows * cols * 8, "ksj.complex-image-frame", memory_domain=domain))
    out = fire(comb, lease2)
    out = (out.view(torch.complex64).cpu().numpy() if device_edges else out.view(np.complex64)).reshape(rows, cols)
    ref = O.sense_combine(img, ref_maps)
    assert rel_l2(out, ref) <= TOL
    # unit-norm maps: the adaptive combine is the projection of the coil vector on the map
    assert rel_l2(out, np.sum(np.conj(ref_maps) * img, axis=-1)) <= TOL
    est.close()
    comb.close()
