"""Grid pruning (tuning key grid_prune): cells of the oversampled grid that no sample
reaches -- the k-space corners outside the sampled disk, a run of blocks around column
G/2 of each row in the grid's uncentred layout -- are neither written by the matrix
gather nor read by the row FFT, which takes them as zero.  Results must be the same bits
with and without it, on every path that reads the grid: the register-staged row pass
(coil-last epilogue), the LDS-staged pass (coil-major outputs, and fft_register = 0),
the direct-DFT lengths, batched scans, the uneven multi-plan call, and a trajectory that
reaches every cell (no span)."""
import numpy as np
import pytest

from kspacejet_amd import synth

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ops():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from kspacejet_amd import ops as K
    K.init(0)
    return K


@pytest.fixture
def keys(ops):
    saved = {k: ops.tune_get(k) for k in ("grid_prune", "fft_register", "gather_scheme", "multi_fused")}
    ops.tune_set("gather_scheme", 2)
    yield
    for k, v in saved.items():
        ops.tune_set(k, v)


def radial(ops, n, coils, spokes, seed):
    traj = ops.golden_angle2(spokes, 2 * n)
    k = (torch.randn(coils, traj.shape[0], dtype=torch.complex64, device="cuda",
                     generator=torch.Generator(device="cuda").manual_seed(seed)))
    return traj, k


def run(ops, traj, k, n, layout, scans=1):
    plan = ops.Plan2(n, n, traj.shape[0]).set_trajectory(traj)
    try:
        if scans == 1:
            img, rss = plan.adjoint(k, rss=True, layout=layout)
            torch.cuda.synchronize()
            return [img.cpu().numpy(), rss.cpu().numpy()]
        outs = plan.adjoint_batch([k] * scans, images=True, rss=True, layout=layout)
        torch.cuda.synchronize()
        return [t.cpu().numpy() for pair in outs for t in pair]
    finally:
        plan.close()


@pytest.mark.parametrize("n,coils,spokes", [(40, 5, 31), (96, 12, 61), (256, 16, 101), (128, 64, 80)])
@pytest.mark.parametrize("layout", ["coil_last", "coil_major"])
@pytest.mark.parametrize("register", [1, 0])
def test_pruning_keeps_every_bit(ops, keys, n, coils, spokes, layout, register):
    traj, k = radial(ops, n, coils, spokes, n + coils)
    ops.tune_set("fft_register", register)
    results = []
    for prune in (1, 0):
        ops.tune_set("grid_prune", prune)
        results.append(run(ops, traj, k, n, layout))
    assert all(np.array_equal(a, b) for a, b in zip(*results))


def test_pruning_keeps_batched_scans(ops, keys):
    n = 256
    traj, k = radial(ops, n, 16, 201, 9)
    results = []
    for prune in (1, 0):
        ops.tune_set("grid_prune", prune)
        results.append(run(ops, traj, k, n, "coil_last", scans=4))
    assert all(np.array_equal(a, b) for a, b in zip(*results))


def test_full_coverage_trajectory_has_nothing_to_prune(ops, keys):
    n, coils, S = 64, 4, 40000
    g = torch.Generator(device="cuda").manual_seed(3)
    traj = (torch.rand(S, 2, device="cuda", generator=g) * 2 - 1) * np.float32(np.pi * 0.999)
    k = torch.randn(coils, S, dtype=torch.complex64, device="cuda", generator=g)
    results = []
    for prune in (1, 0):
        ops.tune_set("grid_prune", prune)
        results.append(run(ops, traj.contiguous(), k, n, "coil_last"))
    assert all(np.array_equal(a, b) for a, b in zip(*results))


@pytest.mark.parametrize("fused", [0, 1, 2])
def test_pruning_keeps_uneven_batches(ops, keys, fused):
    shapes = synth.uneven_items2(8)
    plans, ks = [], []
    for i, (m, c, spokes) in enumerate(shapes):
        traj, k = radial(ops, m, c, spokes, i)
        plans.append(ops.Plan2(m, m, traj.shape[0]).set_trajectory(traj))
        ks.append(k)
    ops.tune_set("multi_fused", fused)
    results = []
    for prune in (1, 0):
        ops.tune_set("grid_prune", prune)
        rss = [torch.zeros((m, m), dtype=torch.float32, device="cuda") for m, _, _ in shapes]
        imgs = [torch.zeros((m, m, c), dtype=torch.complex64, device="cuda") for m, c, _ in shapes]
        ws = torch.empty(ops.adjoint_multi_workspace_bytes(plans, [k.shape[0] for k in ks]), dtype=torch.uint8,
                         device="cuda")
        ops.adjoint_multi(plans, ks, rss=rss, images=imgs, workspace=ws)
        torch.cuda.synchronize()
        results.append([t.cpu().numpy() for t in rss + imgs])
    for p in plans:
        p.close()
    assert all(np.array_equal(a, b) for a, b in zip(*results))
    assert ops.chain_faults() == 0
