"""ksj_gpu_trajectory_check, the device-side trajectory check behind device-resident
trajectory payloads: the bitwise comparison with a resident trajectory and the value
rule of the reference's non-Cartesian operators (finite, and within [-pi, pi] in
radians_per_pixel: radial_gridding_reconstruct.cpp:219-221).  Expected flags come from
numpy on the same arrays; planted differences sit at the first and last index and past
the grid-stride boundary of the launch (1024 blocks x 256 threads)."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

PI_F = np.float32(np.pi)
FLT_MAX = float(np.finfo(np.float32).max)


@pytest.fixture(scope="module")
def ops():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from kspacejet_amd import ops as K
    K.init(0)
    return K


def expected(a, b, bound):
    differs = b is None or not np.array_equal(a.view(np.uint32), b.view(np.uint32))
    invalid = not bool(np.all(np.abs(a) <= np.float32(bound)))
    return (1 if differs else 0) | (2 if invalid else 0)


def check(ops, a, b, bound):
    ta = torch.from_numpy(a).cuda()
    tb = torch.from_numpy(b).cuda() if b is not None else None
    got = ops.trajectory_check(ta, tb, bound)
    assert got == expected(a, b, bound), (a.size, bound)
    return got


@pytest.mark.parametrize("size", [1, 2, 255, 257, 1024 * 256 + 3, 4 * 1024 * 256 * 2 + 1])
def test_planted_differences_are_found_everywhere(ops, size):
    rng = np.random.default_rng(size)
    a = rng.uniform(-3, 3, size).astype(np.float32)
    assert check(ops, a, a.copy(), float(PI_F)) == 0
    assert check(ops, a, None, float(PI_F)) == 1
    for at in sorted({0, size // 2, size - 1, min(size - 1, 1024 * 256)}):
        b = a.copy()
        b[at] = np.nextafter(b[at], np.float32(9))
        assert check(ops, a, b, float(PI_F)) == 1
        bad = a.copy()
        bad[at] = np.float32(4.0)
        assert check(ops, bad, a, float(PI_F)) == 3


def test_value_rule_edges(ops):
    base = np.zeros(777, np.float32)
    for v, bound, want in [(PI_F, PI_F, 0), (-PI_F, PI_F, 0), (np.nextafter(PI_F, np.float32(4)), PI_F, 2),
                           (np.nextafter(-PI_F, np.float32(-4)), PI_F, 2), (np.float32(np.nan), PI_F, 2),
                           (np.float32(np.inf), FLT_MAX, 2), (np.float32(-np.inf), FLT_MAX, 2),
                           (np.float32(np.nan), FLT_MAX, 2), (np.float32(3e38), FLT_MAX, 0)]:
        a = base.copy()
        a[401] = v
        assert check(ops, a, a.copy(), float(bound)) == want, (v, bound)


def test_signed_zero_and_nan_payloads_compare_by_bits(ops):
    a = np.zeros(64, np.float32)
    b = a.copy()
    b[7] = np.float32(-0.0)
    assert check(ops, a, b, float(PI_F)) == 1
    n1 = np.full(8, np.nan, np.float32)
    n2 = n1.copy()
    assert check(ops, n1, n2, float(PI_F)) == 2  # same NaN bits: equal, but invalid


def test_empty_trajectory_and_bad_arguments(ops):
    e = torch.empty(0, dtype=torch.float32, device="cuda")
    assert ops.trajectory_check(e, None, float(PI_F)) == 1  # no resident trajectory: differs
    with pytest.raises(RuntimeError, match="bound"):
        ops.trajectory_check(torch.zeros(4, device="cuda"), None, -1.0)
