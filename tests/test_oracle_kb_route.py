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
