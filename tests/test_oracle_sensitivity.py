"""Pins of the sensitivity-map oracle (Walsh's adaptive method; the reference only plans
org.kspacejet.calibration/sensitivity_map_estimate, so these pin the mathematics):

* against numpy.linalg.eigh: at every pixel the map is the dominant eigenvector of the
  windowed coil covariance formed here independently (window clipped at the edges), with
  coil 0 rotated real and non-negative;
* a constant complex gain per coil is recovered exactly, whatever the window;
* a 1-pixel window gives I / |I| rotated, the closed form of a rank-one covariance;
* a zero neighbourhood gives a zero map; the maps are unit-norm elsewhere;
* a planted wrap-around window (instead of clipping) fails the eigh comparison at the edges."""
import numpy as np

import oracle as O


def crandn(rng, *shape):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


def rank_one_images(rng, rows, cols, coils, noise=1e-3):
    yy, xx = np.mgrid[0:rows, 0:cols]
    centres = rng.uniform(0, max(rows, cols), (coils, 2))
    sens = np.stack([np.exp(-((yy - cy) ** 2 + (xx - cx) ** 2) / (2 * (0.6 * rows) ** 2)) *
                     np.exp(1j * rng.uniform(0, 2 * np.pi)) for cy, cx in centres], axis=-1)
    obj = 1.0 + 0.5 * np.cos(yy / 3.0) * np.sin(xx / 5.0)
    return sens * obj[..., None] + noise * crandn(rng, rows, cols, coils)


def eigh_maps(img, window, wrap=False):
    rows, cols, coils = img.shape
    r = window // 2
    out = np.zeros_like(img)
    for y in range(rows):
        for x in range(cols):
            R = np.zeros((coils, coils), np.complex128)
            for dy in range(-r, r + 1):
                for dx in range(-r, r + 1):
                    yy, xx = y + dy, x + dx
                    if wrap:
                        yy, xx = yy % rows, xx % cols
                    elif not (0 <= yy < rows and 0 <= xx < cols):
                        continue
                    a = img[yy, xx]
                    R += np.outer(a, a.conj())
            w, v = np.linalg.eigh(R)
            s = v[:, -1]
            out[y, x] = s * np.conj(s[0]) / abs(s[0])
    return out


def test_walsh_maps_are_the_dominant_eigenvectors():
    rng = np.random.default_rng(11)
    img = rank_one_images(rng, 9, 11, 5)
    got = O.walsh_maps(img, 5, 40)
    ref = eigh_maps(img, 5)
    assert np.max(np.abs(got - ref)) < 1e-9
    assert np.allclose(np.linalg.norm(got, axis=-1), 1.0, atol=1e-12)
    assert np.all(np.abs(got[..., 0].imag) < 1e-15) and np.all(got[..., 0].real >= 0)
    wrapped = eigh_maps(img, 5, wrap=True)
    assert np.max(np.abs(got - wrapped)) > 1e-4  # a wrap-around window is told apart


def test_constant_gains_are_recovered_exactly():
    rng = np.random.default_rng(12)
    gains = crandn(rng, 6)
    obj = rng.uniform(0.5, 2.0, (8, 7))
    img = obj[..., None] * gains
    want = gains / np.linalg.norm(gains) * np.conj(gains[0]) / abs(gains[0])
    for window in (1, 3, 7):
        got = O.walsh_maps(img, window, 3)
        assert np.max(np.abs(got - want)) < 1e-13, window


def test_one_pixel_window_is_the_normalised_pixel():
    rng = np.random.default_rng(13)
    img = crandn(rng, 4, 5, 3)
    got = O.walsh_maps(img, 1, 2)
    want = img / np.linalg.norm(img, axis=-1, keepdims=True)
    want = want * np.conj(want[..., :1]) / np.abs(want[..., :1])
    assert np.max(np.abs(got - want)) < 1e-13


def test_zero_neighbourhoods_give_zero_maps():
    img = np.zeros((6, 6, 4), np.complex128)
    img[0, 0] = [1, 2j, -1, 0.5]
    got = O.walsh_maps(img, 3, 4)
    assert np.all(got[3:, 3:] == 0)  # no signal within one pixel
    assert np.allclose(np.linalg.norm(got[:2, :2], axis=-1), 1.0, atol=1e-12)
