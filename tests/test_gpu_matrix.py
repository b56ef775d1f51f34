"""The matrix gather schedule (4x4 or 4x8 cell blocks summed on the matrix cores,
v_mfma_f32_16x16x4_f32): every padded coil count (8..64), kernel width 2..8, grids
that are not a multiple of 4 or smaller than a block window (so window bins wrap
more than once), random trajectories reaching the grid edge, work items short
enough that most blocks are cut into several (including item sizes that are not a
multiple of four), every block grouping and item mapping -- each against the FP64
oracle, deterministic run to run, in agreement with the 2x2-block schedule, and
with each coil's result independent of how many coils are stacked with it."""
import numpy as np
import pytest

import oracle as O
from kspacejet_amd import synth
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


@pytest.fixture(params=[4, 8], ids=["4x4", "4x8"])
def tune(ops, request):
    """Sets tuning keys for one test (matrix block width from the parameter) and
    restores them."""
    saved = {}

    def set_(key, value):
        if key not in saved:
            saved[key] = ops.tune_get(key)
        ops.tune_set(key, value)

    set_("matrix_block_width", request.param)
    yield set_
    for key, value in saved.items():
        ops.tune_set(key, value)


def radial_scan(ops, n, coils, spokes, seed):
    traj = ops.golden_angle2(spokes, 2 * n)
    img = torch.from_numpy(synth.coil_images2(n, coils, seed)).cuda()
    k = ops.nudft2_forward_f64(traj, img, (n - 1) / 2, (n - 1) / 2).cpu().numpy()
    return traj, synth.add_noise(k, seed).astype(np.complex64)


def random_scan(n, coils, samples, seed):
    rng = np.random.default_rng(seed)
    traj = rng.uniform(-3.14, 3.14, size=(samples, 2)).astype(np.float32)
    k = (rng.standard_normal((coils, samples)) + 1j * rng.standard_normal((coils, samples))).astype(np.complex64)
    return torch.from_numpy(traj).cuda(), k


def run(ops, tune, scheme, traj, k, n, width=4, density="ramp", item_visits=0, **plan_kw):
    tune("gather_scheme", scheme)
    plan = ops.Plan2(n, n, traj.shape[0], width=width, density=density, item_visits=item_visits,
                     **plan_kw).set_trajectory(traj)
    img, rss = plan.adjoint(torch.from_numpy(k).cuda(), rss=True)
    torch.cuda.synchronize()
    out = img.cpu().numpy(), rss.cpu().numpy()
    plan.close()
    return out


def oracle_images(traj, k, n, width=4, density="ramp"):
    t = traj.cpu().numpy().astype(np.float64)
    dcf = O.ramp_dcf2(t) if density == "ramp" else None
    return np.stack(per_coil(lambda c: O.kb_adjoint2(k[c], t, dcf, n, n, width=width), k.shape[0]), axis=-1)


@pytest.mark.parametrize("n,coils,spokes,width,item_visits", [
    (64, 3, 45, 4, 0),       # Cp 8
    (48, 12, 37, 4, 0),      # Cp 16
    (41, 20, 29, 6, 0),      # Cp 32, grid 82 (20.5 blocks a side; 10.25 of 8 columns)
    (37, 40, 25, 8, 0),      # Cp 64, grid 74
    (64, 16, 45, 4, 4),      # nearly every block cut into several items
    (64, 16, 45, 4, 10),     # item size rounded up to 12
    (53, 9, 31, 2, 24),      # width 2, grid 106
])
def test_matrix_schedule_matches_the_oracle(ops, tune, n, coils, spokes, width, item_visits):
    traj, k = radial_scan(ops, n, coils, spokes, synth.scan_seed(11, n + coils))
    img, rss = run(ops, tune, 2, traj, k, n, width, item_visits=item_visits)
    ref = oracle_images(traj, k, n, width)
    assert rel_l2(img, ref) <= TOL
    assert rel_l2(rss, O.rss(ref)) <= TOL


@pytest.mark.parametrize("n,coils,samples,width", [(20, 5, 3000, 4), (4, 2, 200, 4), (3, 17, 500, 8), (30, 33, 4000, 6)])
def test_matrix_schedule_on_random_trajectories_and_tiny_grids(ops, tune, n, coils, samples, width):
    tune("matrix_pipeline", 2)  # both pipelines on every coil count (the oracle test above runs 1)
    traj, k = random_scan(n, coils, samples, n * 1000 + coils)
    img, _ = run(ops, tune, 2, traj, k, n, width, density="none", item_visits=8)
    ref = oracle_images(traj, k, n, width, density="none")
    assert rel_l2(img, ref) <= TOL


@pytest.mark.parametrize("block_group,gather_map,batch,run_cost,pipeline,nt_store,slab", [
    (1, 0, 8, 32, 1, 0, 1), (2, 1, 4, 1, 2, 1, 1), (4, 2, 16, 64, 2, 0, 1), (2, 0, 16, 7, 1, 1, 3),
    (2, 0, 8, 13, 2, 1, 16)])
def test_matrix_launch_keys_keep_results(ops, tune, block_group, gather_map, batch, run_cost, pipeline, nt_store,
                                         slab):
    """Block grouping changes only which blocks share a workgroup, the run mapping and run
    cost which wave takes an item, the batch and pipeline depth how many visits are loaded
    together, the store kind how cells reach memory: all bit-exact."""
    traj, k = radial_scan(ops, 56, 24, 41, 3)
    base = run(ops, tune, 2, traj, k, 56)
    tune("block_group", block_group)
    tune("gather_map", gather_map)
    tune("matrix_batch", batch)
    tune("matrix_run_cost", run_cost)
    tune("matrix_pipeline", pipeline)
    tune("matrix_nt_store", nt_store)
    tune("matrix_slab", slab)
    other = run(ops, tune, 2, traj, k, 56)
    assert np.array_equal(base[0], other[0]) and np.array_equal(base[1], other[1])


def test_matrix_schedule_is_deterministic_and_agrees_with_blocks(ops, tune):
    traj, k = radial_scan(ops, 96, 24, 71, 5)
    a = run(ops, tune, 2, traj, k, 96)
    b = run(ops, tune, 2, traj, k, 96)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    blocks = run(ops, tune, 0, traj, k, 96)
    assert rel_l2(a[0], blocks[0]) <= 2e-6


def test_coil_results_do_not_depend_on_the_stack(ops, tune):
    """A coil's products use only its own columns, so a scan gridded alone (Cp 16)
    and inside a 64-coil stack gives bit-identical images."""
    traj, k = radial_scan(ops, 64, 16, 45, 9)
    alone = run(ops, tune, 2, traj, k, 64)[0]
    rng = np.random.default_rng(1)
    other = (rng.standard_normal((48, k.shape[1])) + 1j * rng.standard_normal((48, k.shape[1]))).astype(np.complex64)
    stacked = run(ops, tune, 2, traj, np.concatenate([other[:20], k, other[20:]]), 64)[0]
    assert np.array_equal(stacked[..., 20:36], alone)


def test_linear_gridding_on_the_matrix_schedule(ops, tune):
    """The reference's bilinear route (grid = image, uncentred) on 4x4 blocks agrees
    with the 2x2-block schedule."""
    traj, k = radial_scan(ops, 30, 6, 23, 4)
    kw = dict(kernel="linear", oversampling=1.0, centred=False)
    a = run(ops, tune, 2, traj, k, 30, width=2, **kw)
    b = run(ops, tune, 0, traj, k, 30, width=2, **kw)
    assert rel_l2(a[0], b[0]) <= 2e-6


@pytest.mark.parametrize("n,coils,spokes,width,item_visits", [
    (64, 3, 45, 4, 0), (48, 12, 37, 2, 0), (41, 20, 29, 6, 8), (37, 40, 25, 8, 0), (30, 64, 21, 4, 12)])
def test_weights_from_taps_give_the_table_bits(ops, tune, n, coils, spokes, width, item_visits):
    """Packed visits (sample and window bin, weights formed in the gather from the
    sample's taps, loaded per lane or shared across the visit's lanes) and visits
    carrying the stored cell weights sum the same float products in the same order:
    bit-identical images and RSS at every width, padded coil count and block width,
    with items cut short."""
    traj, k = radial_scan(ops, n, coils, spokes, synth.scan_seed(12, n + coils))
    results = []
    for weights, share in ((1, 0), (1, 1), (0, 0)):
        tune("matrix_weights", weights)
        tune("matrix_tap_share", share)
        results.append(run(ops, tune, 2, traj, k, n, width, item_visits=item_visits))
    for other in results[1:]:
        assert np.array_equal(results[0][0], other[0]) and np.array_equal(results[0][1], other[1])


def test_packed_visit_plans_are_smaller(ops, tune):
    traj, _ = radial_scan(ops, 64, 2, 45, 1)
    sizes = []
    for weights in (1, 0):
        tune("matrix_weights", weights)
        tune("gather_scheme", 2)
        plan = ops.Plan2(64, 64, traj.shape[0]).set_trajectory(traj)
        info = plan.info()
        sizes.append(info["device_bytes"])
        plan.close()
    assert sizes[0] < sizes[1]
