"""Route trajectory normalisation (noncartesian_rss_hdf5.cpp:776-834) in the oracle: the
reference's equivalence KAT (radial_rss_hdf5_tests.cpp:353-390: 0.5 cycles/FOV, pi
rad/px and index 1 of a 2x2 encoded matrix give the same trajectory), its rejections
(0.75 cycles/FOV, unspecified units, encoded dimension < 2, non-finite), the inclusive
Nyquist bounds, and float32 arithmetic bit for bit against numpy's float32 products."""
import numpy as np

import oracle as O

PI_F = np.float32(np.pi)


def test_equivalent_units_give_the_same_trajectory():
    outs = []
    for units, v in (("cycles_per_fov", 0.5), ("radians_per_pixel", float(PI_F)), ("encoded_matrix_index", 1.0)):
        raw = np.array([[v, -v], [0.0, v], [-v, 0.0]], np.float32)
        out, rc = O.normalize_trajectory2(raw, units, 2, 2)
        assert rc == 0
        outs.append(out)
    for o in outs[1:]:
        assert np.array_equal(o, outs[0])
    assert np.array_equal(outs[0][0], [-PI_F, PI_F])  # [row = ky, column = kx]


def test_rejections_and_inclusive_bounds():
    assert O.normalize_trajectory2(np.array([[0.75, 0]], np.float32), "cycles_per_fov", 2, 2)[1] == 2
    assert O.normalize_trajectory2(np.array([[0.0, 0]], np.float32), "unspecified", 2, 2)[1] == 3
    assert O.normalize_trajectory2(np.array([[0.0, 0]], np.float32), "encoded_matrix_index", 1, 2)[1] == 4
    assert O.normalize_trajectory2(np.array([[np.nan, 0]], np.float32), "radians_per_pixel", 2, 2)[1] == 1
    assert O.normalize_trajectory2(np.array([[0, np.inf]], np.float32), "cycles_per_fov", 2, 2)[1] == 1
    above_pi = np.nextafter(PI_F, np.float32(4))
    assert O.normalize_trajectory2(np.array([[0, above_pi]], np.float32), "radians_per_pixel", 2, 2)[1] == 2
    ok = np.array([[PI_F, -PI_F]], np.float32)
    assert O.normalize_trajectory2(ok, "radians_per_pixel", 2, 2)[1] == 0
    edge = np.array([[64.0, -32.0]], np.float32)  # encoded 64 rows x 128 columns: kx <= 64, ky <= 32
    assert O.normalize_trajectory2(edge, "encoded_matrix_index", 64, 128)[1] == 0
    assert O.normalize_trajectory2(np.array([[64.0, -32.5]], np.float32), "encoded_matrix_index", 64, 128)[1] == 2


def test_float32_products_bit_for_bit():
    rng = np.random.default_rng(0)
    raw = rng.uniform(-0.5, 0.5, (2000, 2)).astype(np.float32)
    out, rc = O.normalize_trajectory2(raw, "cycles_per_fov", 2, 2)
    assert rc == 0
    two_pi = np.float32(2) * PI_F
    assert np.array_equal(out[:, 0], raw[:, 1] * two_pi) and np.array_equal(out[:, 1], raw[:, 0] * two_pi)
    idx = rng.uniform(-40, 40, (2000, 2)).astype(np.float32)
    out, rc = O.normalize_trajectory2(idx, "encoded_matrix_index", 96, 80)
    assert rc == 0
    sr, sc = np.float32(2) * PI_F / np.float32(96), np.float32(2) * PI_F / np.float32(80)
    assert np.array_equal(out[:, 0], idx[:, 1] * sr) and np.array_equal(out[:, 1], idx[:, 0] * sc)
