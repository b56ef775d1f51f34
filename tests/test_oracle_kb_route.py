"""The FP32 CPU route the benchmark times as its like-for-like baseline
(oracle/ksj_kb_route.c): it reconstructs what the FP64 oracle reconstructs (ramp DCF,
KB W = 4/6, sigma 2, deapodized and cropped, RSS over coils) to FP32 accuracy, and the
threaded timing loop returns the single-scan result bit for bit."""
import numpy as np
import pytest

import oracle as O


def scan(n, coils, spokes, readout, seed):
    rng = np.random.default_rng(seed)
    traj = O.golden_angle2(spokes, readout).astype(np.float32)
    k = (rng.standard_normal((coils, traj.shape[0])) + 1j * rng.standard_normal((coils, traj.shape[0])))
    return traj, k.astype(np.complex64)


@pytest.mark.parametrize("n,width", [(32, 4), (64, 4), (32, 6)])
def test_route_matches_the_fp64_oracle(n, width):
    traj, k = scan(n, 3, 2 * n, 2 * n, n + width)
    got = O.kb_route_rss_f32(k, traj, n, width=width)
    t64 = traj.astype(np.float64)
    ref = O.rss(O.kb_adjoint2_coils(k, t64, O.ramp_dcf2(t64), n, width=width))
    err = np.linalg.norm(got - ref) / np.linalg.norm(ref)
    assert err <= 2e-5, err


def test_threaded_repeats_give_the_single_scan_bits():
    traj, k = scan(32, 2, 40, 64, 9)
    one = O.kb_route_rss_f32(k, traj, 32)
    many = O.kb_route_rss_f32(k, traj, 32, scans=7, threads=3)
    assert np.array_equal(one.view(np.uint32), many.view(np.uint32))


def test_route_rejects_what_it_cannot_grid():
    traj, k = scan(32, 2, 8, 64, 1)
    with pytest.raises(ValueError):
        O.kb_route_rss_f32(k, traj, 24)  # the 48-point grid is not a power of two
    bad = traj.copy()
    bad[3, 0] = np.nan
    with pytest.raises(ValueError):
        O.kb_route_rss_f32(k, bad, 32)


# ---------------------------------------------------------------- Cartesian and 3-D routes
# The FP32 CPU restatements the benchmark times beside its Cartesian (cfg1) and 3-D
# (cfg4) sections, held to the FP64 oracle.

@pytest.mark.parametrize("rows,cols", [(64, 64), (32, 128)])
@pytest.mark.parametrize("centred", [True, False])
def test_cartesian_route_matches_the_oracle(rows, cols, centred):
    rng = np.random.default_rng(rows + cols + centred)
    k = (rng.standard_normal((3, rows, cols)) + 1j * rng.standard_normal((3, rows, cols))).astype(np.complex64)
    got = O.cart_route_rss_f32(k, centred=centred)
    ref = O.rss(O.ifft2_coil_images(k.astype(np.complex128), centred=centred))
    err = np.linalg.norm(got - ref) / np.linalg.norm(ref)
    assert err <= 1e-5, err
    many = O.cart_route_rss_f32(k, centred=centred, scans=5, threads=3)
    assert np.array_equal(got.view(np.uint32), many.view(np.uint32))


def test_cartesian_route_rejects_other_lengths():
    with pytest.raises(ValueError):
        O.cart_route_rss_f32(np.zeros((2, 48, 64), np.complex64))


@pytest.mark.parametrize("width", [4, 6])
def test_3d_route_matches_the_oracle(width):
    n = 16
    traj = O.golden_means3(150, 2 * n).astype(np.float32)
    rng = np.random.default_rng(width)
    k = (rng.standard_normal((3, traj.shape[0])) + 1j * rng.standard_normal((3, traj.shape[0]))).astype(np.complex64)
    got = O.kb_route3_rss_f32(k, traj, n, width=width)
    t64 = traj.astype(np.float64)
    dcf = O.ramp_sq_dcf3(t64)
    coils = [O.kb_adjoint3(k[c], t64, dcf, (n, n, n), width=width) for c in range(k.shape[0])]
    ref = np.sqrt(sum(np.abs(x) ** 2 for x in coils))
    err = np.linalg.norm(got - ref) / np.linalg.norm(ref)
    assert err <= 2e-5, err
    # threads take whole coils: the same volume up to the order of the per-thread sums
    split = O.kb_route3_rss_f32(k, traj, n, width=width, threads=3)
    assert np.allclose(split, got, rtol=1e-5, atol=1e-6 * float(got.max()))


def test_3d_route_rejects_what_it_cannot_grid():
    traj = O.golden_means3(20, 48).astype(np.float32)
    k = np.ones((1, traj.shape[0]), np.complex64)
    with pytest.raises(ValueError):
        O.kb_route3_rss_f32(k, traj, 24)  # the 48-point grid is not a power of two
    bad = O.golden_means3(20, 32).astype(np.float32)
    bad[5, 1] = np.nan
    with pytest.raises(ValueError):
        O.kb_route3_rss_f32(np.ones((1, bad.shape[0]), np.complex64), bad, 16)
