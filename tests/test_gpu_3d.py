"""3-D KB adjoint (kooshball): the golden-means trajectory, bit-exact binning, and
the volume reconstruction against the FP64 oracle (rel-L2 <= 1e-5) for unequal
axes, every kernel width, multi-pass coil batches and work items cut short so
blocks are summed from several partial slots; coil batching and repeat runs give
bit-identical results.  Every test runs under both gather schedules: 2x2x2 blocks
on the VALU and 4x4x4 blocks on the matrix cores."""
import numpy as np
import pytest

import oracle as O
from tests.helpers import per_coil, rel_l2

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


@pytest.fixture(autouse=True, params=[0, 2], ids=["blocks", "matrix"])
def scheme(ops, request):
    """The 3-D gather schedule of the plans each test builds."""
    saved = ops.tune_get("gather3_scheme")
    ops.tune_set("gather3_scheme", request.param)
    yield request.param
    ops.tune_set("gather3_scheme", saved)


def kspace(coils, samples, seed):
    rng = np.random.default_rng(seed)
    return (rng.standard_normal((coils, samples)) + 1j * rng.standard_normal((coils, samples))).astype(np.complex64)


def oracle_volume(k, traj, shape, width, density):
    t = traj.cpu().numpy().astype(np.float64)
    dcf = O.ramp_sq_dcf3(t) if density == "ramp_squared" else None
    return np.stack(per_coil(lambda c: O.kb_adjoint3(k[c], t, dcf, shape, width=width), k.shape[0]), axis=-1)


def test_golden_means3_matches_the_oracle(ops):
    spokes, readout = 301, 64
    got = ops.golden_means3(spokes, readout).cpu().numpy().astype(np.float64)
    ref = O.golden_means3(spokes, readout)
    assert np.max(np.abs(got - ref)) <= 4e-7  # FP64 on both sides, rounded to float once


@pytest.mark.parametrize("kind", ["kooshball", "random"])
def test_binning_is_bit_exact(ops, kind):
    shape = (12, 10, 14)
    if kind == "kooshball":
        traj = ops.golden_means3(97, 24)
    else:
        rng = np.random.default_rng(3)
        traj = torch.from_numpy(rng.uniform(-np.pi, np.pi, (5000, 3)).astype(np.float32)).cuda()
    plan = ops.Plan3(shape, traj.shape[0]).set_trajectory(traj)
    start, perm = plan.export_bins()
    _, rstart, rperm = O.bin3(traj.cpu().numpy(), tuple(2 * n for n in shape))
    assert np.array_equal(start.cpu().numpy().astype(np.uint32), rstart)
    assert np.array_equal(perm.cpu().numpy().astype(np.uint32), rperm)
    plan.close()


@pytest.mark.parametrize("shape,coils,spokes,readout,width,density,cpp,item_visits", [
    ((16, 16, 16), 3, 120, 32, 4, "ramp_squared", 8, 0),
    ((12, 14, 10), 12, 90, 28, 6, "none", 5, 0),       # unequal axes, passes of 5, 5, 2 coils
    ((10, 10, 10), 20, 80, 20, 8, "ramp_squared", 8, 7),  # short items: most blocks summed from partial slots
    ((14, 14, 14), 9, 100, 28, 2, "ramp_squared", 16, 0),
])
def test_volume_matches_the_oracle(ops, shape, coils, spokes, readout, width, density, cpp, item_visits):
    traj = ops.golden_means3(spokes, readout)
    k = kspace(coils, traj.shape[0], sum(shape) + coils)
    plan = ops.Plan3(shape, traj.shape[0], width=width, density=density, item_visits=item_visits).set_trajectory(traj)
    img, rss = plan.adjoint(torch.from_numpy(k).cuda(), rss=True, coils_per_pass=cpp)
    torch.cuda.synchronize()
    ref = oracle_volume(k, traj, shape, width, density)
    assert rel_l2(img.cpu().numpy(), ref) <= TOL
    assert rel_l2(rss.cpu().numpy(), np.sqrt(np.sum(np.abs(ref) ** 2, axis=-1))) <= TOL
    plan.close()


def test_coil_batching_and_repeats_are_bit_identical(ops):
    shape = (12, 12, 12)
    traj = ops.golden_means3(110, 24)
    k = torch.from_numpy(kspace(13, traj.shape[0], 9)).cuda()
    plan = ops.Plan3(shape, traj.shape[0]).set_trajectory(traj)
    runs = [plan.adjoint(k, rss=True, coils_per_pass=c) for c in (13, 4, 13)]
    torch.cuda.synchronize()
    for img, rss in runs[1:]:
        assert torch.equal(img, runs[0][0]) and torch.equal(rss, runs[0][1])
    plan.close()


def test_random_trajectory_to_the_grid_edge(ops):
    shape = (8, 9, 10)
    rng = np.random.default_rng(11)
    traj = torch.from_numpy(rng.uniform(-np.pi, np.pi, (3000, 3)).astype(np.float32)).cuda()
    k = kspace(4, 3000, 12)
    plan = ops.Plan3(shape, 3000, density="none").set_trajectory(traj)
    img, _ = plan.adjoint(torch.from_numpy(k).cuda())
    torch.cuda.synchronize()
    assert rel_l2(img.cpu().numpy(), oracle_volume(k, traj, shape, 4, "none")) <= TOL
    plan.close()


def test_schedules_agree(ops, scheme):
    """The matrix schedule sums the same products as the 2x2x2 one, in groups of four."""
    if scheme != 2:
        pytest.skip("compared from the matrix run")
    shape = (20, 18, 22)
    traj = ops.golden_means3(151, 40)
    k = torch.from_numpy(kspace(6, traj.shape[0], 21)).cuda()
    results = []
    for s in (2, 0):
        ops.tune_set("gather3_scheme", s)
        plan = ops.Plan3(shape, traj.shape[0]).set_trajectory(traj)
        img, _ = plan.adjoint(k, rss=True)
        torch.cuda.synchronize()
        results.append(img.cpu().numpy())
        plan.close()
    ops.tune_set("gather3_scheme", 2)
    assert rel_l2(results[0], results[1]) <= 2e-6


@pytest.mark.parametrize("width", [4, 6, 8])
def test_shared_tap_loads_give_the_same_bits(ops, scheme, width):
    """The matrix gather's tap-sharing form (each lane loads one or two words of a visit's
    3W taps and takes its own by bpermute: one word at width 4, two at 6 and 8) against
    every lane loading its own taps, bit for bit, with the oracle beside."""
    shape = (10, 13, 11)
    traj = ops.golden_means3(60, 24)
    k = kspace(5, traj.shape[0], 31 + width)
    saved = ops.tune_get("matrix3_tap_share")
    outs = []
    try:
        for share in (1, 0):
            ops.tune_set("matrix3_tap_share", share)
            plan = ops.Plan3(shape, traj.shape[0], width=width).set_trajectory(traj)
            img, _ = plan.adjoint(torch.from_numpy(k).cuda(), coils_per_pass=8)
            outs.append(img.cpu().numpy())
            plan.close()
    finally:
        ops.tune_set("matrix3_tap_share", saved)
    assert np.array_equal(outs[0].view(np.uint32), outs[1].view(np.uint32))
    assert rel_l2(outs[0], oracle_volume(k, traj, shape, width, "ramp_squared")) <= TOL
