"""The oracle's k-space conditioning functions, pinned by the reference's known-answer
tests (tests/unit/providers/kspacejet-kspace-conditioning/kspace_conditioning_provider_tests.cpp:
539-729), by numpy's complex128 matmul and slicing, and, for the phase correction, by
numpy float32 arithmetic with every product and sum rounded on its own (the rounding of
std::complex<float> multiplication of finite values)."""
import numpy as np
import pytest

import oracle as O


def test_reference_kats():
    # noise_prewhiten, 2 channels x 2 x 2, W = diag(2, 3) (tests.cpp:539-569)
    x = np.arange(1, 9, dtype=np.float64).reshape(2, 2, 2).astype(np.complex128)
    y = O.channel_mix(np.diag([2.0, 3.0]), x)
    assert np.array_equal(y.reshape(-1), np.array([2, 4, 6, 8, 15, 18, 21, 24], np.complex128))
    # coil_compress, basis [1, 2] (tests.cpp:608-640)
    y = O.channel_mix(np.array([[1.0, 2.0]]), x)
    assert np.array_equal(y.reshape(-1), np.array([11, 14, 17, 20], np.complex128))
    # non-Cartesian prewhiten and compress, 2 channels x 3 samples (tests.cpp:642-708)
    xn = np.arange(1, 7, dtype=np.float64).reshape(2, 3).astype(np.complex128)
    assert np.array_equal(O.channel_mix(np.diag([2.0, 3.0]), xn).reshape(-1), [2, 4, 6, 12, 15, 18])
    assert np.array_equal(O.channel_mix(np.array([[1.0, 2.0]]), xn).reshape(-1), [9, 12, 15])
    # phase_correct (tests.cpp:571-606)
    k = np.array([1, 2, 3, 4, 1 + 1j, 2 + 1j, 3 + 1j, 4 + 1j], np.complex64).reshape(2, 2, 2)
    p = np.array([1j, -1j, 1, -1], np.complex64).reshape(2, 2)
    want = np.array([1j, -2j, 3j, -4j, 1 + 1j, -2 - 1j, 3 + 1j, -4 - 1j], np.complex64)
    assert np.array_equal(O.phase_correct(k, p).reshape(-1), want)
    # readout_oversampling_remove, 2 x 2 x 4 -> offset 1, 2 columns (tests.cpp:710-729)
    f = np.array([[[100 * c + 10 * r + k for k in range(4)] for r in range(2)] for c in range(2)], np.complex64)
    assert np.array_equal(O.readout_crop(f, 2, 1).reshape(-1).real, [1, 2, 11, 12, 101, 102, 111, 112])


@pytest.mark.parametrize("out_ch,in_ch,samples", [(1, 1, 5), (3, 7, 33), (16, 16, 400), (8, 64, 129), (64, 64, 17)])
def test_channel_mix_is_the_matrix_product(out_ch, in_ch, samples):
    rng = np.random.default_rng(out_ch * 1000 + in_ch)
    w = rng.standard_normal((out_ch, in_ch)) + 1j * rng.standard_normal((out_ch, in_ch))
    x = rng.standard_normal((in_ch, samples)) + 1j * rng.standard_normal((in_ch, samples))
    np.testing.assert_allclose(O.channel_mix(w, x), w @ x, rtol=1e-12, atol=1e-12)
    assert np.array_equal(O.channel_mix(np.eye(in_ch), x), x)


def test_phase_correct_rounds_every_operation_to_float():
    rng = np.random.default_rng(5)
    x = (rng.standard_normal((3, 5, 7)) + 1j * rng.standard_normal((3, 5, 7))).astype(np.complex64)
    p = np.exp(1j * rng.uniform(-np.pi, np.pi, (3, 7))).astype(np.complex64)
    a, b = x.real, x.imag
    c, d = p.real[:, None, :], p.imag[:, None, :]
    want = np.empty_like(x)
    want.real = (a * c) - (b * d)  # float32 arrays: each product and difference rounds on its own
    want.imag = (a * d) + (b * c)
    got = O.phase_correct(x, p)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
    # a conjugated phase or a transposed index would not match
    assert not np.array_equal(O.phase_correct(x, np.conj(p)), got)


def test_readout_crop_is_a_slice():
    rng = np.random.default_rng(6)
    x = (rng.standard_normal((4, 6, 20)) + 1j * rng.standard_normal((4, 6, 20))).astype(np.complex64)
    for out_cols, off in [(20, 0), (10, 5), (1, 19), (3, 0)]:
        assert np.array_equal(O.readout_crop(x, out_cols, off), x[..., off:off + out_cols])
