"""FP64 forward NUFFT of a volume (ksj_gpu_nufft3_forward_f64), the model the 3-D workload's
k-space is simulated with: a Kaiser-Bessel kernel of width 4, 6 or 8 on the 2x grid, FP64
throughout.  Pinned against the exact forward model ora_nudft3_forward (itself pinned by a
numpy sum and the adjoint identity, test_oracle3.py): the error falls with the width as the
kernel's approximation error does (W = 4 / 6 / 8 within 2e-3 / 3e-5 / 1e-6 rel-L2), on a
cube and on unequal sides, with the half-pixel origin and another one; a mistake in the
deapodization, the phase or an axis order would leave the error at O(1)."""
import numpy as np
import pytest

import oracle as O
from tests.helpers import rel_l2

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

BOUNDS = {4: 2e-3, 6: 3e-5, 8: 1e-6}


@pytest.fixture(scope="module")
def ops():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from kspacejet_amd import ops as K
    K.init(0)
    return K


def volume(shape, seed):
    rng = np.random.default_rng(seed)
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


@pytest.mark.parametrize("shape", [(16, 16, 16), (8, 16, 32)])
def test_width_ladder_against_the_exact_forward_model(ops, shape):
    x = volume(shape, sum(shape))
    rng = np.random.default_rng(7)
    t = rng.uniform(-np.pi, np.pi, (600, 3)).astype(np.float32)
    ref = O.nudft3_forward(x, t.astype(np.float64))
    errs = []
    for width in (4, 6, 8):
        got = ops.nufft3_forward_f64(torch.from_numpy(x).cuda(), torch.from_numpy(t).cuda(), width=width)
        err = rel_l2(got.cpu().numpy(), ref)
        print(f"{shape} W={width}: rel-L2 {err:.2e}")
        assert err <= BOUNDS[width], width
        errs.append(err)
    assert errs[0] > 10 * errs[1] > 100 * errs[2]


def test_origin_and_kooshball_trajectory(ops):
    shape = (16, 16, 16)
    x = volume(shape, 3)
    traj = ops.golden_means3(40, 32)
    t = traj.cpu().numpy().astype(np.float64)
    for origin in ((7.5, 7.5, 7.5), (3.0, 9.0, 5.5)):
        ref = O.nudft3_forward(x, t, origin=origin)
        got = ops.nufft3_forward_f64(torch.from_numpy(x).cuda(), traj, origin=origin, width=6)
        assert rel_l2(got.cpu().numpy(), ref) <= BOUNDS[6], origin


def test_invalid_shapes_are_refused(ops):
    from kspacejet_amd._lib import KsjGpuError
    x = torch.zeros((12, 16, 16), dtype=torch.complex128, device="cuda")  # grid side 24: not a power of two
    t = torch.zeros((4, 3), dtype=torch.float32, device="cuda")
    with pytest.raises(KsjGpuError):
        ops.nufft3_forward_f64(x, t, width=6)
    with pytest.raises(KsjGpuError):
        ops.nufft3_forward_f64(torch.zeros((16, 16, 16), dtype=torch.complex128, device="cuda"), t, width=5)


def test_simulated_kooshball_kspace_follows_the_exact_model(ops):
    """synth.kooshball_kspace3 (phantom x coil maps through the FP64 W=6 forward NUFFT, here
    without noise) against the exact forward model of the same volumes."""
    from kspacejet_amd import synth
    n, coils, seed = 16, 3, synth.scan_seed(4, 1)
    traj = ops.golden_means3(60, 2 * n)
    k = synth.kooshball_kspace3(ops, torch, n, coils, traj, seed, width=6, noise=0.0)
    t = traj.cpu().numpy().astype(np.float64)
    obj = synth.phantom3(n, seed, torch)
    for c in range(coils):
        img = (obj * synth.coil_map3(n, c, coils, seed, torch)).cpu().numpy()
        ref = O.nudft3_forward(img, t)
        assert rel_l2(k[c].cpu().numpy(), ref) <= 5e-5, c  # W = 6 model error, then rounded to complex64
