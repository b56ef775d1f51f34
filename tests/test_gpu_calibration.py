"""Calibration estimators on the GPU against the FP64 oracle
(providers/kspacejet-calibration/src/operators/*.cpp restated) and through the
org.kspacejet.calibration Provider module:

* the channel covariance (centred with the reference's conjugate on the left, and the
  uncentred second moment) at 1..64 channels, rel <= 1e-6;
* the FP64 Jacobi eigensolver: eigenvalues, residual, unitarity; non-finite input
  reported;
* the whitening matrix rel-L2 <= 1e-5, the compression basis per row up to a unit phase
  (the rotation order fixes the phase) and its projector, the phase model bit for bit;
* the reference's three KATs through the Provider ABI, a sample the estimate cannot use
  rejected with the grant released, and noise estimation -> prewhitening chained through
  the two modules giving whitened noise of unit covariance (for a real channel coupling:
  the reference's covariance convention whitens conj(x) when the coupling is complex,
  see test_oracle_calibration.py)."""
import numpy as np
import pytest

import oracle as O
from kspacejet_amd import provider_host as H
from tests.helpers import rel_l2

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

CB = 8


@pytest.fixture(scope="module")
def ops():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from kspacejet_amd import ops as K
    K.init(0)
    return K


def crandn(rng, *shape):
    return (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(np.complex64)


def correlated(rng, channels, samples, offset=0.3):
    mix = crandn(rng, channels, channels).astype(np.complex128) + 2 * np.eye(channels)
    return (mix @ crandn(rng, channels, samples) + offset * (1 + 1j)).astype(np.complex64)


def dev(x):
    return torch.from_numpy(np.ascontiguousarray(x)).cuda()


@pytest.mark.parametrize("channels,samples", [(1, 7), (3, 1000), (8, 65), (16, 100000), (33, 4097), (64, 20000)])
def test_covariances_match_the_oracle(ops, channels, samples):
    rng = np.random.default_rng(channels * 1000 + samples)
    x = correlated(rng, channels, samples)
    for centred in (True, False):
        got = ops.channel_covariance(dev(x), centred).cpu().numpy()
        ref = O.channel_covariance(x, centred)
        assert rel_l2(got, ref) <= 1e-6, centred


@pytest.mark.parametrize("n", [1, 2, 5, 16, 31, 64])
def test_eigensolver_matches_numpy(ops, n):
    rng = np.random.default_rng(n)
    a = crandn(rng, n, n).astype(np.complex128)
    a = a + np.conj(a).T
    vals, vecs = ops.hermitian_eigen(dev(a))
    vals, vecs = vals.cpu().numpy(), vecs.cpu().numpy()
    ref = np.linalg.eigvalsh(a)
    scale = np.abs(ref).max()
    assert np.all(np.diff(vals) >= 0)
    assert np.allclose(vals, ref, rtol=0, atol=1e-12 * scale)
    assert np.allclose(a @ vecs, vecs * vals, atol=1e-11 * scale)
    assert np.allclose(np.conj(vecs).T @ vecs, np.eye(n), atol=1e-12)
    ovals, _ = O.hermitian_eigen(a)  # the reference's solver order gives the same spectrum
    assert np.allclose(vals, ovals, rtol=0, atol=1e-12 * scale)


def test_eigensolver_reports_non_finite_input(ops):
    a = np.eye(4, dtype=np.complex128)
    a[2, 3] = np.inf
    with pytest.raises(RuntimeError, match="non-finite"):
        ops.hermitian_eigen(dev(a))


@pytest.mark.parametrize("channels,samples", [(2, 50), (8, 5000), (32, 20000), (64, 65536)])
def test_noise_model_matches_the_oracle(ops, channels, samples):
    rng = np.random.default_rng(channels + 17)
    x = correlated(rng, channels, samples)
    w = ops.noise_model_estimate(dev(x)).cpu().numpy()
    ref = O.noise_model_estimate(x)
    err = rel_l2(w, ref)
    print(f"noise model {channels} channels, {samples} samples: rel-L2 {err:.2e}")
    assert err <= 1e-5
    c = O.channel_covariance(x, True)
    assert np.allclose(w.astype(np.complex128) @ c @ np.conj(w.astype(np.complex128)).T, np.eye(channels), atol=1e-4)


@pytest.mark.parametrize("physical,virtual,shape", [(4, 1, (16, 16)), (12, 6, (64, 48)), (32, 16, (128, 128)),
                                                    (64, 64, (96, 64))])
def test_coil_basis_matches_the_oracle_up_to_row_phases(ops, physical, virtual, shape):
    rng = np.random.default_rng(physical * 7 + virtual)
    x = correlated(rng, physical, shape[0] * shape[1]).reshape(physical, *shape)
    b = ops.coil_compression_basis_estimate(dev(x), virtual).cpu().numpy().astype(np.complex128)
    ref = O.coil_compression_basis_estimate(x, virtual)
    for row in range(virtual):
        phase = np.vdot(ref[row], b[row])
        assert abs(abs(phase) - 1.0) <= 1e-5, row
        assert rel_l2(b[row], ref[row] * phase) <= 1e-5, row
    assert rel_l2(np.conj(b).T @ b, np.conj(ref).T @ ref) <= 1e-5


def test_phase_model_is_bit_exact(ops):
    rng = np.random.default_rng(3)
    ref = crandn(rng, 8, 7, 333)
    ref[2, :, 10] = 0
    got = ops.phase_correction_estimate(dev(ref)).cpu().numpy()
    want = O.phase_correction_estimate(ref)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))


# ---------------------------------------------------------------- through the Provider ABI

@pytest.fixture(scope="module")
def module(ops):
    return H.ProviderModule("calibration")


def fire(module, op_id, config, data, in_type, out_bytes, out_type, scratch):
    st, op, msg = module.create(op_id, config)
    assert st == H.OK, msg
    assert op.open() == H.OK
    ls = H.Lease([H.Batch(0, [H.Item(H.aligned_copy(np.ascontiguousarray(data)), in_type)])],
                 H.OutputSpec(out_bytes, out_type), scratch_bytes=scratch)
    st, r, msg = op.process(ls)
    op.close()
    return st, r, msg, ls


def test_provider_noise_model_kat(module):
    """calibration_provider_tests.cpp:464-485."""
    x = np.array([-1, 1, -1, 1, -2, -2, 2, 2], np.complex64)
    st, r, msg, ls = fire(module, "noise_model_estimate", '{"channel_count":2}', x, "ksj.noise-calibration-frame",
                          4 * CB, "ksj.noise-model", 24 * 4 + 24)
    assert st == H.OK, msg
    assert r.sealed_output_count == 1 and ls.seal_calls == 1 and ls.sealed.semantic_key_hash == 71
    w = ls.output.view(np.complex64)
    assert abs(w[0].real - 1) <= 2e-4 and abs(w[3].real - 0.5) <= 2e-4 and abs(w[1]) <= 2e-4 and abs(w[2]) <= 2e-4


def test_provider_phase_model_kat(module):
    """calibration_provider_tests.cpp:487-509."""
    x = np.array([1j, -1, 1j, -1, 1, 0, 1, 0], np.complex64)
    st, r, msg, ls = fire(module, "phase_correction_estimate", '{"channel_count":2,"readout_sample_count":2}', x,
                          "ksj.phase-reference-frame", 4 * CB, "ksj.phase-model", 0)
    assert st == H.OK, msg
    assert np.allclose(ls.output.view(np.complex64), [-1j, -1, 1, 1], atol=1e-5)


def test_provider_coil_basis_kat(module):
    """calibration_provider_tests.cpp:511-530."""
    x = np.array([2, -2, 2, -2, 1, 1, -1, -1], np.complex64)
    st, r, msg, ls = fire(module, "coil_compression_basis_estimate",
                          '{"physical_channel_count":2,"virtual_channel_count":1}', x, "ksj.kspace-frame", 2 * CB,
                          "ksj.coil-compression-basis", 24 * 4 + 8)
    assert st == H.OK, msg
    first, second = ls.output.view(np.complex64)
    assert abs(abs(first) ** 2 + abs(second) ** 2 - 1) <= 1e-4 and abs(first) > 0.999 and abs(second) < 1e-3


def test_provider_rejects_samples_it_cannot_estimate_from(module):
    x = np.array([-1, 1, np.nan, 1, -2, -2, 2, 2], np.complex64)
    st, r, msg, ls = fire(module, "noise_model_estimate", '{"channel_count":2}', x, "ksj.noise-calibration-frame",
                          4 * CB, "ksj.noise-model", 24 * 4 + 24)
    assert st == H.CONTRACT_VIOLATION and "could not estimate" in msg
    assert ls.acquire_calls == 1 and ls.release_calls == 1 and ls.seal_calls == 0


def prewhiten(w, frame):
    channels, rows, cols = frame.shape
    cond = H.ProviderModule("kspace_conditioning")
    st, op, msg = cond.create("noise_prewhiten", f'{{"channel_count":{channels},"cols":{cols},"rows":{rows}}}')
    assert st == H.OK, msg
    assert op.open() == H.OK
    ls = H.Lease([H.Batch(0, [H.Item(H.aligned_copy(frame), "ksj.kspace-frame")]),
                  H.Batch(1, [H.Item(H.aligned_copy(w), "ksj.noise-model")])],
                 H.OutputSpec(frame.nbytes, "ksj.kspace-frame"), scratch_bytes=0)
    st, r, msg = op.process(ls)
    op.close()
    assert st == H.OK, msg
    return ls.output.view(np.complex64).reshape(channels, -1)


def test_estimated_noise_model_whitens_through_the_conditioning_module(module):
    """noise_model_estimate -> noise_prewhiten (org.kspacejet.kspace-conditioning).  The
    reference's noise covariance is sum conj(x_r) x_c, so its W makes conj(x) exactly white;
    x itself, for a real channel coupling, is white up to the O(1/sqrt(S)) imaginary parts of
    the sample covariance."""
    rng = np.random.default_rng(11)
    channels, rows, cols = 8, 64, 64
    mix = rng.standard_normal((channels, channels)) + 2 * np.eye(channels)
    noise = (mix @ crandn(rng, channels, rows * cols)).astype(np.complex64)
    st, _, msg, ls = fire(module, "noise_model_estimate", f'{{"channel_count":{channels}}}', noise,
                          "ksj.noise-calibration-frame", channels * channels * CB, "ksj.noise-model",
                          24 * channels * channels + 12 * channels)
    assert st == H.OK, msg
    w = ls.output.view(np.complex64).reshape(channels, channels).copy()
    exact = prewhiten(w, np.conj(noise).reshape(channels, rows, cols))
    assert np.allclose(O.channel_covariance(exact, True), np.eye(channels), atol=2e-4)
    y = prewhiten(w, noise.reshape(channels, rows, cols))
    cov = O.channel_covariance(y, True)
    assert np.allclose(cov.real, np.eye(channels), atol=5e-3)
    assert np.abs(cov.imag).max() <= 6 / np.sqrt(rows * cols)
