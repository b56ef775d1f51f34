"""GPU parity: Cartesian multi-coil inverse FFT and coil combination vs the FP64 oracle.

Tolerances: relative L2 <= 1e-5 for complex64 outputs against FP64 (the hot-path
gate); the reference's own known answers are checked at their stated 1e-5
absolute tolerance.
"""
import numpy as np
import pytest

import oracle as O
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


def dev(x):
    return torch.from_numpy(np.ascontiguousarray(x)).cuda()


def test_reference_kat_two_by_two_two_channels(ops):
    # cartesian_recon_provider_tests.cpp:60-63,339-366: DC impulses 4 and 8i -> 1 and 2i everywhere, coil-last.
    k = np.zeros((2, 2, 2), np.complex64)
    k[0, 0, 0] = 4.0
    k[1, 0, 0] = 8.0j
    out = ops.ifft2_coil_images(dev(k)).cpu().numpy()
    assert out.shape == (2, 2, 2)
    np.testing.assert_allclose(out[..., 0], np.ones((2, 2)), atol=1e-5)
    np.testing.assert_allclose(out[..., 1], 2j * np.ones((2, 2)), atol=1e-5)


@pytest.mark.parametrize("shape", [(3, 8, 8), (2, 6, 10), (4, 64, 64), (2, 128, 192), (3, 320, 256), (2, 384, 384),
                                   (1, 512, 640), (2, 100, 36), (1, 2, 512)])
@pytest.mark.parametrize("centred", [False, True])
def test_ifft2_matches_oracle(ops, shape, centred):
    rng = np.random.default_rng(sum(shape) + centred)
    k = (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(np.complex64)
    got = ops.ifft2_coil_images(dev(k), centred=centred).cpu().numpy()
    ref = O.ifft2_coil_images(k.astype(np.complex128), centred=centred)
    assert rel_l2(got, ref) <= TOL
    # numpy is the textbook routine the oracle is pinned to
    npref = np.fft.ifft2(k.astype(np.complex128))
    if centred:
        npref = np.fft.fftshift(np.fft.ifft2(np.fft.ifftshift(k.astype(np.complex128), axes=(-2, -1))), axes=(-2, -1))
    assert rel_l2(got, np.moveaxis(npref, 0, -1)) <= TOL


def test_ifft2_coil_major_layout(ops):
    rng = np.random.default_rng(5)
    k = (rng.standard_normal((3, 64, 32)) + 1j * rng.standard_normal((3, 64, 32))).astype(np.complex64)
    last = ops.ifft2_coil_images(dev(k), layout="coil_last").cpu().numpy()
    major = ops.ifft2_coil_images(dev(k), layout="coil_major").cpu().numpy()
    assert np.array_equal(np.moveaxis(major, 0, -1), last)


def test_cartesian_config1_rss(ops):
    # Cartesian 256x256, 8 coils, centred inverse FFT then RSS (the CPU-runnable route).
    n, coils = 256, 8
    seed = synth.scan_seed(1)
    k = synth.cartesian_kspace2(n, coils, seed, centred=True)
    img = ops.ifft2_coil_images(dev(k), centred=True)
    got_rss = ops.rss(img).cpu().numpy()
    ref_img = O.ifft2_coil_images(k.astype(np.complex128), centred=True)
    ref_rss = O.rss(ref_img)
    assert rel_l2(img.cpu().numpy(), ref_img) <= TOL
    assert rel_l2(got_rss, ref_rss) <= TOL
    # Without noise the centred inverse returns the phantom times the coil RSS.
    clean = synth.cartesian_kspace2(n, coils, seed, centred=True, noise=False)
    truth = synth.phantom2(n, seed) * np.sqrt((np.abs(synth.coil_maps2(n, coils, seed)) ** 2).sum(0))
    assert rel_l2(ops.rss(ops.ifft2_coil_images(dev(clean), centred=True)).cpu().numpy(), truth) <= TOL


def test_rss_reference_kat(ops):
    # coil_combine_provider_tests.cpp:62-70,299-320: 2x2 pixels x 2 coils -> 5, 12, 13, sqrt(6).
    x = np.array([3 + 4j, 0, 0, 12j, 5, 12j, 1 + 1j, 2j], np.complex64).reshape(2, 2, 2)
    out = ops.rss(dev(x)).cpu().numpy()
    np.testing.assert_allclose(out.ravel(), [5, 12, 13, np.sqrt(6)], atol=1e-5)


@pytest.mark.parametrize("coils", [1, 7, 8, 16, 32, 64])
def test_rss_matches_oracle(ops, coils):
    rng = np.random.default_rng(coils)
    x = (rng.standard_normal((33, 17, coils)) + 1j * rng.standard_normal((33, 17, coils))).astype(np.complex64)
    got = ops.rss(dev(x)).cpu().numpy()
    assert rel_l2(got, O.rss(x.astype(np.complex128))) <= TOL


def test_sense_combine_recovers_object(ops):
    # With I_c = S_c x the weighted combine returns x exactly (up to float rounding).
    rng = np.random.default_rng(11)
    n, coils = 48, 6
    x = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    s = np.moveaxis(synth.coil_maps2(n, coils, 3), 0, -1)
    imgs = (s * x[..., None]).astype(np.complex64)
    got = ops.sense_combine(dev(imgs), dev(s.astype(np.complex64))).cpu().numpy()
    assert rel_l2(got, x) <= 1e-5
    assert rel_l2(got, O.sense_combine(imgs.astype(np.complex128), s.astype(np.complex64).astype(np.complex128))) <= TOL
