"""The tile gather schedule (64-cell tiles, 16x4 or 8x8, whose window samples are
staged in LDS):
every padded coil count (8..64), kernel width 2..8, grids that are not a multiple of
the tile side or smaller than a tile window (so window bins wrap more than once),
random trajectories reaching the grid edge, and work items short enough that most
tiles are cut into several items -- each against the FP64 oracle, deterministic run
to run, and in agreement with the 2x2-block schedule."""
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


@pytest.fixture(params=[4, 8], ids=["16x4", "8x8"])
def tune(ops, request):
    """Sets tuning keys for one test (tile width from the parameter) and restores them."""
    saved = {}

    def set_(key, value):
        if key not in saved:
            saved[key] = ops.tune_get(key)
        ops.tune_set(key, value)

    set_("tile_width", request.param)
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


def run(ops, tune, scheme, traj, k, n, width=4, density="ramp"):
    tune("gather_scheme", scheme)
    plan = ops.Plan2(n, n, traj.shape[0], width=width, density=density).set_trajectory(traj)
    img, rss = plan.adjoint(torch.from_numpy(k).cuda(), rss=True)
    torch.cuda.synchronize()
    out = img.cpu().numpy(), rss.cpu().numpy()
    plan.close()
    return out


def oracle_images(traj, k, n, width=4, density="ramp"):
    t = traj.cpu().numpy().astype(np.float64)
    dcf = O.ramp_dcf2(t) if density == "ramp" else None
    return np.stack(per_coil(lambda c: O.kb_adjoint2(k[c], t, dcf, n, n, width=width), k.shape[0]), axis=-1)


@pytest.mark.parametrize("n,coils,spokes,width,item_samples", [
    (64, 3, 45, 4, 256),     # Cp 8
    (48, 12, 37, 4, 256),    # Cp 16
    (40, 20, 29, 6, 256),    # Cp 32, grid 80 (10 tile columns of 8, 5 rows of 16)
    (36, 40, 25, 8, 256),    # Cp 64, grid 72 (4.5 tile rows of 16)
    (64, 16, 45, 4, 7),      # nearly every tile cut into several items
    (52, 9, 31, 2, 64),      # width 2, grid 104 (6.5 tile rows of 16)
])
def test_tiles_match_the_oracle(ops, tune, n, coils, spokes, width, item_samples):
    tune("tile_item_samples", item_samples)
    traj, k = radial_scan(ops, n, coils, spokes, synth.scan_seed(7, n + coils))
    img, rss = run(ops, tune, 1, traj, k, n, width)
    ref = oracle_images(traj, k, n, width)
    assert rel_l2(img, ref) <= TOL
    assert rel_l2(rss, O.rss(ref)) <= TOL


@pytest.mark.parametrize("n,coils,samples,width", [(20, 5, 3000, 4), (4, 2, 200, 4), (6, 17, 500, 8), (30, 33, 4000, 6)])
def test_tiles_on_random_trajectories_and_tiny_grids(ops, tune, n, coils, samples, width):
    tune("tile_item_samples", 33)
    traj, k = random_scan(n, coils, samples, n * 1000 + coils)
    img, _ = run(ops, tune, 1, traj, k, n, width, density="none")
    ref = oracle_images(traj, k, n, width, density="none")
    assert rel_l2(img, ref) <= TOL


def test_tiles_are_deterministic_and_agree_with_blocks(ops, tune):
    traj, k = radial_scan(ops, 96, 24, 71, 5)
    a = run(ops, tune, 1, traj, k, 96)
    b = run(ops, tune, 1, traj, k, 96)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    blocks = run(ops, tune, 0, traj, k, 96)
    assert rel_l2(a[0], blocks[0]) <= 2e-6
