"""Pins of the calibration oracle (providers/kspacejet-calibration/src/operators/*.cpp
restated in FP64): the reference's three KATs
(tests/unit/providers/kspacejet-calibration/calibration_provider_tests.cpp:464-530), the
covariances against numpy's sums (with the reference's conjugate on the left for the
noise covariance), the Jacobi solver against numpy.linalg.eigh (eigenvalues, residual,
unitarity), the whitening matrix against numpy's inverse square root and its defining
identity W C W^H = I, the compression basis against numpy's leading eigenvectors (up to a
unit phase per row, which the rotation order fixes) and its projector, and the phase
model bit for bit against float32 numpy arithmetic."""
import numpy as np
import pytest

import oracle as O


def crandn(rng, *shape):
    return (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(np.complex64)


def correlated(rng, channels, samples):
    mix = crandn(rng, channels, channels).astype(np.complex128) + 2 * np.eye(channels)
    return (mix @ crandn(rng, channels, samples) + 0.3 * (1 + 1j)).astype(np.complex64)


# ---------------------------------------------------------------- reference KATs

def test_noise_model_kat():
    """calibration_provider_tests.cpp:464-485: channels (-1, 1, -1, 1) and (-2, -2, 2, 2)."""
    x = np.array([[-1, 1, -1, 1], [-2, -2, 2, 2]], np.complex64)
    w = O.noise_model_estimate(x)
    assert abs(w[0, 0].real - 1.0) <= 2e-4 and abs(w[1, 1].real - 0.5) <= 2e-4
    assert abs(w[0, 1]) <= 2e-4 and abs(w[1, 0]) <= 2e-4


def test_phase_model_kat():
    """calibration_provider_tests.cpp:487-509."""
    ref = np.array([1j, -1, 1j, -1, 1, 0, 1, 0], np.complex64).reshape(2, 2, 2)
    p = O.phase_correction_estimate(ref)
    assert np.allclose(p.reshape(-1), [-1j, -1, 1, 1], atol=1e-5)


def test_coil_basis_kat():
    """calibration_provider_tests.cpp:511-530: the leading eigenmode of channels (2, -2, 2, -2)
    and (1, 1, -1, -1) is the first channel."""
    x = np.array([[2, -2, 2, -2], [1, 1, -1, -1]], np.complex64)
    b = O.coil_compression_basis_estimate(x.reshape(2, 2, 2), 1)
    assert abs(abs(b[0, 0]) ** 2 + abs(b[0, 1]) ** 2 - 1.0) <= 1e-4
    assert abs(b[0, 0]) > 0.999 and abs(b[0, 1]) < 1e-3


# ---------------------------------------------------------------- covariance

@pytest.mark.parametrize("channels,samples", [(1, 7), (5, 300), (16, 1000)])
def test_covariances_are_numpy_sums(channels, samples):
    rng = np.random.default_rng(channels)
    x = correlated(rng, channels, samples)
    x64 = x.astype(np.complex128)
    centred = x64 - x64.mean(axis=1, keepdims=True)
    noise = np.conj(centred) @ centred.T / samples          # conj on the left, as moments.hpp:129-133
    second = x64 @ np.conj(x64).T / samples                  # x conj(x)^T, as the coil basis
    assert np.allclose(O.channel_covariance(x, True), noise, rtol=1e-12, atol=1e-12)
    assert np.allclose(O.channel_covariance(x, False), second, rtol=1e-12, atol=1e-12)
    if channels > 1:  # the two conventions differ (one is the conjugate of the other's form)
        assert not np.allclose(noise, centred @ np.conj(centred).T / samples)


# ---------------------------------------------------------------- eigensolver

@pytest.mark.parametrize("n", [1, 2, 3, 8, 17, 64])
def test_jacobi_matches_numpy_eigh(n):
    rng = np.random.default_rng(n + 100)
    a = crandn(rng, n, n).astype(np.complex128)
    a = a + np.conj(a).T
    vals, vecs = O.hermitian_eigen(a)
    ref = np.linalg.eigvalsh(a)
    assert np.all(np.diff(vals) >= 0)
    assert np.allclose(vals, ref, rtol=1e-12, atol=1e-12 * np.abs(ref).max())
    assert np.allclose(a @ vecs, vecs * vals, atol=1e-11 * np.abs(ref).max())
    assert np.allclose(np.conj(vecs).T @ vecs, np.eye(n), atol=1e-12)


def test_jacobi_rejects_non_finite_input():
    a = np.eye(3, dtype=np.complex128)
    a[1, 1] = np.nan
    with pytest.raises(ValueError):
        O.hermitian_eigen(a)


# ---------------------------------------------------------------- estimates

@pytest.mark.parametrize("channels,samples", [(2, 50), (8, 2000), (32, 4000)])
def test_whitening_is_the_inverse_square_root(channels, samples):
    rng = np.random.default_rng(channels + 7)
    x = correlated(rng, channels, samples)
    c = O.channel_covariance(x, True)
    w = O.noise_model_estimate(x)
    lam, v = np.linalg.eigh(c)
    ref = (v / np.sqrt(np.maximum(lam, np.float32(1e-6)))) @ np.conj(v).T
    assert np.allclose(w, ref, rtol=1e-10, atol=1e-10 * np.abs(ref).max())
    assert np.allclose(w @ c @ np.conj(w).T, np.eye(channels), atol=1e-9)
    assert np.allclose(w, np.conj(w).T, atol=1e-12)


def test_whitening_floor_limits_rank_deficient_noise():
    x = np.zeros((3, 40), np.complex64)
    x[0] = np.exp(1j * np.arange(40)).astype(np.complex64)
    w = O.noise_model_estimate(x)
    lam = np.linalg.eigvalsh(w @ np.conj(w).T)
    assert np.isclose(lam.max(), 1e6, rtol=1e-6)  # the zero-variance directions scaled by 1/floor


@pytest.mark.parametrize("physical,virtual", [(4, 1), (12, 5), (32, 32)])
def test_basis_rows_are_the_leading_eigenvectors(physical, virtual):
    rng = np.random.default_rng(physical * 10 + virtual)
    x = correlated(rng, physical, 3000)
    b = O.coil_compression_basis_estimate(x, virtual)
    x64 = x.astype(np.complex128)
    r = x64 @ np.conj(x64).T / x.shape[1]
    lam, v = np.linalg.eigh(r)
    lead = np.conj(v[:, ::-1][:, :virtual]).T  # rows: conjugated eigenvectors, largest first
    for row in range(virtual):
        phase = np.vdot(lead[row], b[row])
        assert abs(abs(phase) - 1.0) <= 1e-10
        assert np.allclose(b[row], lead[row] * phase, atol=1e-10)
    assert np.allclose(np.conj(b).T @ b, np.conj(lead).T @ lead, atol=1e-10)
    assert np.allclose(b @ np.conj(b).T, np.eye(virtual), atol=1e-12)


def test_phase_model_is_float_arithmetic():
    rng = np.random.default_rng(5)
    ref = crandn(rng, 3, 5, 17)
    ref[1, :, 4] = 0  # a zero sum emits 1 + 0i
    got = O.phase_correction_estimate(ref)
    s = np.zeros((3, 17), np.complex64)
    for line in range(5):  # float sums in line order
        s = (s.real + ref[:, line].real).astype(np.float32) + 1j * (s.imag + ref[:, line].imag).astype(np.float32)
        s = s.astype(np.complex64)
    mag = np.sqrt(s.real.astype(np.float64) ** 2 + s.imag.astype(np.float64) ** 2).astype(np.float32)
    want = np.where(mag == 0, np.complex64(1), (s.real / np.where(mag == 0, 1, mag)).astype(np.float32)
                    - 1j * (s.imag / np.where(mag == 0, 1, mag)).astype(np.float32)).astype(np.complex64)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
    assert got[1, 4] == 1


def test_reference_convention_whitens_the_conjugate_samples():
    """The noise covariance of the reference is sum conj(x_r) x_c (moments.hpp:129-133), the
    conjugate of E[x x^H].  Its whitening matrix therefore makes conj(x) white, and x
    white only when the channel coupling is real.  Pinned here so the convention cannot
    drift on either side."""
    rng = np.random.default_rng(21)
    n, s = 4, 200000
    mix = crandn(rng, n, n).astype(np.complex128) + 2 * np.eye(n)
    x = (mix @ crandn(rng, n, s)).astype(np.complex64)
    w = O.noise_model_estimate(x)

    def centred(a):
        a = a.astype(np.complex128)
        return a - a.mean(axis=1, keepdims=True)

    def cov_usual(y):
        return y @ np.conj(y).T / s

    assert np.allclose(cov_usual(w @ np.conj(centred(x))), np.eye(n), atol=1e-9)
    assert not np.allclose(cov_usual(w @ centred(x)), np.eye(n), atol=1e-2)
    real_mix = rng.standard_normal((n, n)) + 2 * np.eye(n)
    xr = (real_mix @ crandn(rng, n, s)).astype(np.complex64)
    wr = O.noise_model_estimate(xr)
    assert np.allclose(cov_usual(wr @ centred(xr)), np.eye(n), atol=2e-2)
