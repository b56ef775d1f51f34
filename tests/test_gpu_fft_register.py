"""The register-staged FFT form (fft_register = 1) against the LDS-staged form (0).

The register form runs the first Stockham pass on the loaded registers and the last
pass straight into the stores, with the same butterflies, twiddles and summation order,
so every pass it takes gives the staged form's bits: Cartesian inverse FFTs at every
LDS length (the column pass runs across coils), 3-D volumes (the y and z passes), and
every 512 schedule.  The 2-D radial pipeline with coil-last outputs becomes rows stored
coil-fastest plus one column pass that writes each scan's coil-last images and RSS
itself; its images equal the staged pipeline's bit for bit and its RSS (the same float
sum in coil order) too, at 1..32 channels, several scans per pass, uneven matrices
(grids 384 and 640) and channel counts whose groups do not fit a block (which keep the
staged pipeline), and the headline shape (256, four 16-coil scans stacked in one pass).
Volumes keep every pass output coil-fastest; from grid 256 up the z pass takes the column
epilogue (coil-last images and the RSS carried across coil passes), whose bits equal the
staged pipeline's separate combine kernels."""
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


@pytest.fixture
def tune(ops):
    saved = {}

    def set_(key, value):
        if key not in saved:
            saved[key] = ops.tune_get(key)
        ops.tune_set(key, value)

    yield set_
    for key, value in saved.items():
        ops.tune_set(key, value)


def both_forms(tune, fn):
    out = {}
    for form in (0, 1):
        tune("fft_register", form)
        out[form] = fn()
        torch.cuda.synchronize()
    return out[0], out[1]


@pytest.mark.parametrize("rows,cols", [(64, 64), (128, 192), (256, 256), (320, 384), (512, 512), (640, 768),
                                       (1024, 64), (96, 100)])
@pytest.mark.parametrize("layout", ["coil_last", "coil_major"])
def test_cartesian_ifft_gives_the_staged_bits(ops, tune, rows, cols, layout):
    coils = 6
    rng = np.random.default_rng(rows + cols)
    k = torch.from_numpy((rng.standard_normal((coils, rows, cols)) + 1j * rng.standard_normal((coils, rows, cols)))
                         .astype(np.complex64)).cuda()
    for centred in (False, True):
        staged, reg = both_forms(tune, lambda: ops.ifft2_coil_images(k, centred=centred, layout=layout))
        assert torch.equal(staged, reg), (rows, cols, layout, centred)


@pytest.mark.parametrize("schedule", [0, 1, 2])
def test_every_512_schedule_gives_the_staged_bits(ops, tune, schedule):
    tune("fft_schedule", schedule)
    rng = np.random.default_rng(schedule)
    k = torch.from_numpy((rng.standard_normal((16, 512, 512)) + 1j * rng.standard_normal((16, 512, 512)))
                         .astype(np.complex64)).cuda()
    staged, reg = both_forms(tune, lambda: ops.ifft2_coil_images(k, centred=True))
    assert torch.equal(staged, reg)


def radial(ops, n, coils, spokes, readout, seed):
    traj = ops.golden_angle2(spokes, readout)
    img = torch.from_numpy(synth.coil_images2(n, coils, seed)).cuda()
    k = ops.nudft2_forward_f64(traj, img, (n - 1) / 2, (n - 1) / 2).cpu().numpy()
    return traj, synth.add_noise(k, seed).astype(np.complex64)


@pytest.mark.parametrize("n,coils,scans", [(64, 1, 3), (64, 2, 5), (96, 4, 4), (64, 8, 9), (96, 16, 5), (128, 32, 2),
                                           (192, 16, 2), (320, 8, 2), (64, 12, 3), (50, 16, 2),
                                           (256, 16, 4)])
def test_radial_column_epilogue_matches_the_staged_pipeline(ops, tune, n, coils, scans):
    traj, _ = radial(ops, n, coils, 2 * n // 3 + 1, 2 * n, n + coils)
    ks = [torch.from_numpy(radial(ops, n, coils, 2 * n // 3 + 1, 2 * n, 100 + i)[1]).cuda() for i in range(scans)]
    plan = ops.Plan2(n, n, traj.shape[0]).set_trajectory(traj)
    staged, reg = both_forms(tune, lambda: plan.adjoint_batch(ks, images=True, rss=True))
    for i in range(scans):
        assert torch.equal(staged[0][i], reg[0][i]), ("images", i)
        assert torch.equal(staged[1][i], reg[1][i]), ("rss", i)
    rss_only = plan.adjoint_batch(ks, images=False, rss=True)[1]
    assert all(torch.equal(a, b) for a, b in zip(rss_only, reg[1]))
    one_img, one_rss = plan.adjoint(ks[0], images=True, rss=True)
    assert torch.equal(one_img, reg[0][0]) and torch.equal(one_rss, reg[1][0])
    if scans and n <= 96:
        t = traj.cpu().numpy().astype(np.float64)
        ref = O.kb_adjoint2_coils(ks[0].cpu().numpy(), t, O.ramp_dcf2(t), n)
        assert rel_l2(reg[0][0].cpu().numpy(), ref) <= 1e-5
        assert rel_l2(reg[1][0].cpu().numpy(), O.rss(ref)) <= 1e-5
    plan.close()


@pytest.mark.parametrize("shape,spokes,readout", [
    ((32, 48, 40), 120, 64),      # grids 64 / 96 / 80: the LDS 64 and direct-DFT passes
    ((128, 128, 128), 600, 256),  # grid 256: the z pass writes images and RSS itself
])
def test_volumes_give_the_staged_bits(ops, tune, shape, spokes, readout):
    traj3 = ops.golden_means3(spokes, readout)
    rng = np.random.default_rng(9)
    k3 = torch.from_numpy((rng.standard_normal((5, traj3.shape[0]))
                           + 1j * rng.standard_normal((5, traj3.shape[0]))).astype(np.complex64)).cuda()
    plan = ops.Plan3(shape, traj3.shape[0]).set_trajectory(traj3)
    staged, reg = both_forms(tune, lambda: plan.adjoint(k3, images=True, rss=True, coils_per_pass=4))
    assert torch.equal(staged[0], reg[0]) and torch.equal(staged[1], reg[1])
    plan.close()
