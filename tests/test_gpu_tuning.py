"""Launch parameters the autotuner may set change speed, not results: every
legal setting of the gather mapping, batch, grouping and pipeline keys and of
fft_lines_per_block reproduces the default result bit for bit, under both the
2x2-block and the matrix gather schedule (the work-item size changes the float
summation order and is checked against the oracle instead)."""
import numpy as np
import pytest

import oracle as O
from kspacejet_amd import synth
from tests.helpers import rel_l2

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ops():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from kspacejet_amd import ops as K
    K.init(0)
    return K


@pytest.fixture(params=[0, 2], ids=["blocks", "matrix"])
def tune(ops, request):
    """Sets tuning keys for one test (gather schedule from the parameter) and restores
    the previous values afterwards."""
    saved = {}

    def set_(key, value):
        if key not in saved:
            saved[key] = ops.tune_get(key)
        ops.tune_set(key, value)

    set_("gather_scheme", request.param)
    yield set_
    for key, value in saved.items():
        ops.tune_set(key, value)


def scan(ops, n=96, coils=40, spokes=61, readout=192, seed=3):
    traj = ops.golden_angle2(spokes, readout)
    img = torch.from_numpy(synth.coil_images2(n, coils, seed)).cuda()
    k = ops.nudft2_forward_f64(traj, img, (n - 1) / 2, (n - 1) / 2).cpu().numpy()
    return traj, synth.add_noise(k, seed).astype(np.complex64)


def recon(ops, traj, k, n, item_visits=0):
    plan = ops.Plan2(n, n, traj.shape[0], item_visits=item_visits).set_trajectory(traj)
    img, rss = plan.adjoint(torch.from_numpy(k).cuda(), rss=True)
    torch.cuda.synchronize()
    return img.cpu().numpy(), rss.cpu().numpy()


def test_launch_parameters_do_not_change_results(ops, tune):
    n = 96
    traj, k = scan(ops)
    base_img, base_rss = recon(ops, traj, k, n)
    for key, values in (("gather_map", (0, 1, 2)), ("gather_batch", (4, 8, 16)), ("fft_lines_per_block", (0, 4, 16)),
                        ("gather_coils_per_lane", (1, 2)), ("block_group", (1, 2, 4)), ("matrix_batch", (16, 4, 8)),
                        ("matrix_slab", (16, 1, 5)), ("matrix_pipeline", (1, 2)), ("matrix_nt_store", (1, 0)),
                        ("matrix_run_cost", (12, 1, 64)), ("matrix_tap_share", (1, 0)),
                        ("matrix_persistent", (0, 8, 1)), ("grid_prune", (1, 0)), ("row_order", (0, 1))):
        # (matrix_block_width changes the summation order: covered by test_gpu_matrix.py)
        for v in values:
            tune(key, v)
            img, rss = recon(ops, traj, k, n)
            assert np.array_equal(img, base_img) and np.array_equal(rss, base_rss), (key, v)
        tune(key, values[0])


def test_block_groups_pad_grids_without_changing_results(ops, tune):
    """Block groups pad the block rows and columns (a 100-point grid has 50 blocks a
    side, 52 with groups of 4); the padding blocks write nothing and every cell is
    summed as with row-major blocks."""
    n = 50
    traj, k = scan(ops, n=n, coils=5, spokes=37, readout=100, seed=4)
    results = []
    for g in (1, 2, 4):
        tune("block_group", g)
        results.append(recon(ops, traj, k, n))
    for img, rss in results[1:]:
        assert np.array_equal(img, results[0][0]) and np.array_equal(rss, results[0][1])


def test_pass_lanes_do_not_change_batched_results(ops, tune):
    """Passes of a batched call overlap on 1, 2 or 3 streams with identical results,
    including a batch whose last pass is partial (9 scans of 12 coils: 5 per pass)."""
    n = 64
    traj, k = scan(ops, n=n, coils=12, spokes=41, readout=128, seed=9)
    ks = [torch.from_numpy(np.roll(k, i, axis=1).copy()).cuda() for i in range(9)]
    plan = ops.Plan2(n, n, traj.shape[0]).set_trajectory(traj)
    results = []
    for lanes in (1, 2, 3):
        tune("pass_lanes", lanes)
        imgs, rss = plan.adjoint_batch(ks, images=True, rss=True)
        torch.cuda.synchronize()
        results.append(([i.cpu().numpy() for i in imgs], [r.cpu().numpy() for r in rss]))
    for imgs, rss in results[1:]:
        assert all(np.array_equal(a, b) for a, b in zip(imgs, results[0][0]))
        assert all(np.array_equal(a, b) for a, b in zip(rss, results[0][1]))


@pytest.mark.parametrize("item_visits", [16, 256])
def test_item_size_keeps_oracle_parity(ops, tune, item_visits):
    n = 64
    traj, k = scan(ops, n=n, coils=4, spokes=41, readout=128, seed=5)
    tune("item_visits", item_visits)
    tune("matrix_item_visits", item_visits)
    img, _ = recon(ops, traj, k, n)
    t = traj.cpu().numpy().astype(np.float64)
    ref = O.kb_adjoint2_coils(k, t, O.ramp_dcf2(t), n)
    assert rel_l2(img, ref) <= 1e-5


def test_illegal_values_are_rejected(ops):
    with pytest.raises(RuntimeError):
        ops.tune_set("gather_batch", 5)
    with pytest.raises(RuntimeError):
        ops.tune_set("no_such_key", 1)


@pytest.mark.parametrize("schedule", [0, 1, 2, 3])
def test_fft_schedules_match_the_oracle_at_every_length(ops, tune, schedule):
    """Every radix schedule of every LDS length (and a direct-DFT length) is an exact
    factorisation of the DFT: uncentred and centred inverse FFTs stay at the FP64
    oracle to rounding, for rows and columns of unequal length."""
    tune("fft_schedule", schedule)
    rng = np.random.default_rng(schedule)
    for rows, cols in [(64, 1024), (128, 768), (192, 640), (256, 512), (320, 384), (384, 320), (512, 256),
                       (640, 192), (768, 128), (1024, 64), (100, 64)]:
        k = (rng.standard_normal((2, rows, cols)) + 1j * rng.standard_normal((2, rows, cols))).astype(np.complex64)
        for centred in (False, True):
            got = ops.ifft2_coil_images(torch.from_numpy(k).cuda(), centred=centred).cpu().numpy()
            ref = O.ifft2_coil_images(k, centred=centred)
            assert rel_l2(got, ref) <= 2e-6, (rows, cols, centred)


@pytest.mark.parametrize("schedule", [1, 2])
def test_fft_schedules_keep_kb_parity(ops, tune, schedule):
    tune("fft_schedule", schedule)
    n = 96
    traj, k = scan(ops, n=n, coils=12, spokes=61, readout=192, seed=6)
    img, rss = recon(ops, traj, k, n)
    t = traj.cpu().numpy().astype(np.float64)
    ref = O.kb_adjoint2_coils(k, t, O.ramp_dcf2(t), n)
    assert rel_l2(img, ref) <= 1e-5 and rel_l2(rss, O.rss(ref)) <= 1e-5
