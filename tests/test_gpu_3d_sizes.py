"""3-D KB adjoint at sizes whose grid lengths take the LDS Stockham passes (64, 128, 192:
the pruned x / y / z passes with fused deapodization that a 256^3 volume runs at 512),
against the full FP64 oracle; and the 256^3 kooshball volume itself (20 coils, 16384
golden-means spokes x 512 samples, KB width 4 on a 512^3 grid, |k|^2 density):

* two coils against the FP64 oracle at a 6 x 6 x 6 lattice of voxels spread over the
  volume (centre, edges, both halves of every axis), rel-L2 <= 1e-5;
* all 20 coils through properties that hold at any size: the 4x4x4 matrix and 2x2x2
  block schedules (two independent summations) agree, coil passes of 8, 16 and 20 give
  the same bits, a repeat gives the same bits, the operator is linear, and the RSS
  volume is the root sum of squares of the coil volumes.

The k-space is seeded complex Gaussian data made on the device."""
import numpy as np
import pytest

import oracle as O
from tests.helpers import per_coil, rel_l2

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

TOL = 1e-5
N, COILS, SPOKES, READOUT = 256, 20, 16384, 512


@pytest.fixture(scope="module")
def ops():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from kspacejet_amd import ops as K
    K.init(0)
    return K


def with_scheme(ops, scheme, fn):
    saved = ops.tune_get("gather3_scheme")
    ops.tune_set("gather3_scheme", scheme)
    try:
        return fn()
    finally:
        ops.tune_set("gather3_scheme", saved)


@pytest.mark.parametrize("scheme", [2, 0], ids=["matrix", "blocks"])
@pytest.mark.parametrize("shape,spokes", [((32, 32, 32), 700), ((64, 64, 64), 1500), ((32, 64, 96), 1200)])
def test_stockham_sized_volumes_match_the_oracle(ops, shape, spokes, scheme):
    readout = 2 * max(shape)
    coils = 3
    traj = ops.golden_means3(spokes, readout)
    rng = np.random.default_rng(sum(shape) + spokes)
    k = (rng.standard_normal((coils, traj.shape[0])) + 1j * rng.standard_normal((coils, traj.shape[0])))
    k = k.astype(np.complex64)
    plan = with_scheme(ops, scheme, lambda: ops.Plan3(shape, traj.shape[0]).set_trajectory(traj))
    img, rss = plan.adjoint(torch.from_numpy(k).cuda(), rss=True, coils_per_pass=2)  # passes of 2 and 1 coils
    torch.cuda.synchronize()
    t = traj.cpu().numpy().astype(np.float64)
    dcf = O.ramp_sq_dcf3(t)
    ref = np.stack(per_coil(lambda c: O.kb_adjoint3(k[c], t, dcf, shape), coils), axis=-1)
    err = rel_l2(img.cpu().numpy(), ref)
    print(f"3-D {shape} grid {tuple(2 * n for n in shape)} scheme {scheme}: rel-L2 {err:.2e}")
    assert err <= TOL
    assert rel_l2(rss.cpu().numpy(), np.sqrt(np.sum(np.abs(ref) ** 2, axis=-1))) <= TOL
    plan.close()


@pytest.fixture(scope="module")
def full(ops):
    traj = ops.golden_means3(SPOKES, READOUT)
    g = torch.Generator(device="cuda").manual_seed(20261205)
    S = traj.shape[0]
    k = torch.complex(torch.randn((COILS, S), generator=g, device="cuda"),
                      torch.randn((COILS, S), generator=g, device="cuda"))
    plan = with_scheme(ops, 2, lambda: ops.Plan3((N, N, N), S).set_trajectory(traj))
    img, rss = plan.adjoint(k, rss=True, coils_per_pass=16)
    torch.cuda.synchronize()
    yield traj, k, plan, img, rss
    plan.close()


# 6 indices per axis: both edges, the centre pair and one point in each half.
LATTICE = [0, 57, 127, 128, 201, 255]


def test_full_volume_matches_the_oracle_on_a_voxel_lattice(ops, full):
    traj, k, _, img, _ = full
    t = traj.cpu().numpy().astype(np.float64)
    dcf = O.ramp_sq_dcf3(t)
    idx = torch.tensor(LATTICE, device="cuda")
    for c in (0, COILS - 1):
        got = img[..., c].index_select(0, idx).index_select(1, idx).index_select(2, idx).cpu().numpy()
        ref = O.kb_adjoint3_voxels(k[c].cpu().numpy(), t, dcf, (N, N, N), LATTICE, LATTICE, LATTICE, threads=16)
        err = rel_l2(got, ref)
        print(f"3-D 256^3 coil {c}: lattice rel-L2 {err:.2e}")
        assert err <= TOL, c


def test_full_volume_coil_passes_and_repeats_are_bit_identical(ops, full):
    _, k, plan, img, rss = full
    for cpp in (20, 8, 16):
        img2, rss2 = plan.adjoint(k, rss=True, coils_per_pass=cpp)
        torch.cuda.synchronize()
        assert torch.equal(img2, img) and torch.equal(rss2, rss), cpp
        del img2, rss2


def test_full_volume_rss_is_the_root_sum_of_squares(ops, full):
    _, _, _, img, rss = full
    ref = torch.sqrt(torch.sum(img.real.double() ** 2 + img.imag.double() ** 2, dim=-1))
    assert float(torch.linalg.vector_norm(rss.double() - ref) / torch.linalg.vector_norm(ref)) <= 1e-6


def test_full_volume_schedules_agree(ops, full):
    traj, k, _, img, _ = full
    plan = with_scheme(ops, 0, lambda: ops.Plan3((N, N, N), traj.shape[0]).set_trajectory(traj))
    b_img, _ = plan.adjoint(k, coils_per_pass=16)
    torch.cuda.synchronize()
    plan.close()
    num = torch.linalg.vector_norm(img - b_img, dim=(0, 1, 2))
    den = torch.linalg.vector_norm(b_img, dim=(0, 1, 2))
    assert float((num / den).max()) <= 2e-6


def test_full_volume_is_linear(ops, full):
    _, k, plan, img, _ = full
    k2 = torch.roll(k, shifts=1, dims=0)
    b, _ = plan.adjoint(k2, coils_per_pass=16)
    ab, _ = plan.adjoint(k + 2 * k2, coils_per_pass=16)
    torch.cuda.synchronize()
    ref = img + 2 * b
    del b
    assert float(torch.linalg.vector_norm(ab - ref) / torch.linalg.vector_norm(ref)) <= 1e-6


def test_full_volume_pruning_keeps_every_bit(ops, full):
    """Grid pruning (the cells outside the sampled ball are neither written by the gather
    nor read by the x pass) gives the same bits as writing and reading the whole grid."""
    _, k, plan, img, rss = full
    saved = ops.tune_get("grid_prune")
    ops.tune_set("grid_prune", 0)
    try:
        img2, rss2 = plan.adjoint(k, rss=True, coils_per_pass=16)
        torch.cuda.synchronize()
    finally:
        ops.tune_set("grid_prune", saved)
    assert torch.equal(img2, img) and torch.equal(rss2, rss)


@pytest.mark.parametrize("shape,spokes", [((22, 22, 22), 300), ((32, 32, 32), 700), ((32, 64, 96), 1200)])
def test_pruning_keeps_small_volumes(ops, shape, spokes):
    """The same at grid lengths that take the direct DFT (44) and the LDS passes."""
    traj = ops.golden_means3(spokes, 2 * max(shape))
    g = torch.Generator(device="cuda").manual_seed(spokes)
    k = torch.complex(torch.randn((3, traj.shape[0]), generator=g, device="cuda"),
                      torch.randn((3, traj.shape[0]), generator=g, device="cuda"))
    plan = with_scheme(ops, 2, lambda: ops.Plan3(shape, traj.shape[0]).set_trajectory(traj))
    saved = ops.tune_get("grid_prune")
    out = []
    try:
        for prune in (1, 0):
            ops.tune_set("grid_prune", prune)
            img, rss = plan.adjoint(k, rss=True, coils_per_pass=2)
            torch.cuda.synchronize()
            out.append((img, rss))
    finally:
        ops.tune_set("grid_prune", saved)
        plan.close()
    assert torch.equal(out[0][0], out[1][0]) and torch.equal(out[0][1], out[1][1])
