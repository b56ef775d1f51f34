"""GPU parity for the non-Cartesian path: binning (bit-exact vs the CPU binning),
Kaiser-Bessel adjoint NUFFT (rel-L2 <= 1e-5 vs the FP64 oracle of the same
algorithm), the reference bilinear route, the exact direct NUDFTs, trajectories
and the ramp DCF.
"""
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


def dev(x):
    return torch.from_numpy(np.ascontiguousarray(x)).cuda()


def radial_scan(ops, n, coils, spokes, readout, seed, noise=True):
    """Golden-angle scan simulated with the FP64 GPU direct forward model (float32 trajectory)."""
    traj = ops.golden_angle2(spokes, readout)
    img = dev(synth.coil_images2(n, coils, seed))
    k = ops.nudft2_forward_f64(traj, img, (n - 1) / 2, (n - 1) / 2).cpu().numpy()
    if noise:
        k = synth.add_noise(k, seed)
    return traj, k.astype(np.complex64)


def oracle_kb(k, traj_np, n, width=4, oversampling=2.0, density=True):
    t = traj_np.astype(np.float64)
    dcf = O.ramp_dcf2(t) if density else None
    imgs = per_coil(lambda c: O.kb_adjoint2(k[c].astype(np.complex128), t, dcf, n, n, oversampling, width), k.shape[0])
    return np.stack(imgs, axis=-1)


def test_golden_angle_matches_host(ops):
    got = ops.golden_angle2(37, 64).cpu().numpy()
    ref = synth.golden_angle2(37, 64)
    assert np.abs(got.astype(np.float64) - ref).max() <= 4e-7
    assert np.array_equal(got, ref.astype(np.float32)) or np.mean(got != ref.astype(np.float32)) < 1e-3
    oref = O.golden_angle2(37, 64)
    assert np.abs(oref - ref).max() < 1e-15


def test_ramp_dcf_reference_kat(ops):
    # radial_gridding_tests.cpp:16-29: (0,0) -> 0, (3/5, 4/5) -> 1, (-pi, 0) -> pi.
    t = np.array([[0, 0], [0.6, 0.8], [-np.pi, 0]], np.float32)
    got = ops.ramp_dcf2(dev(t)).cpu().numpy()
    np.testing.assert_allclose(got, [0.0, 1.0, np.pi], rtol=1e-6, atol=1e-7)


@pytest.mark.parametrize("n,spokes,readout", [(32, 21, 64), (256, 402, 512), (320, 603, 640), (192, 101, 384)])
def test_binning_bit_exact(ops, n, spokes, readout):
    traj = ops.golden_angle2(spokes, readout)
    g = 2 * n
    plan = ops.Plan2(n, n, spokes * readout).set_trajectory(traj)
    start, perm = plan.export_bins()
    _, _, ref_start, ref_perm = O.bin2(traj.cpu().numpy(), g, g)
    assert np.array_equal(start.cpu().numpy().view(np.uint32), ref_start)
    assert np.array_equal(perm.cpu().numpy().view(np.uint32), ref_perm)


def test_binning_edges_bit_exact(ops):
    # +-pi (the inclusive Nyquist edge), exact zeros, and coordinates that round onto cell edges.
    rng = np.random.default_rng(1)
    t = rng.uniform(-np.pi, np.pi, size=(5000, 2)).astype(np.float32)
    t[:50] = 0.0
    t[50:60, 0] = np.float32(np.pi)
    t[60:70, 1] = -np.float32(np.pi)
    t[70:200] = (np.round(t[70:200] / (2 * np.pi / 64)) * (2 * np.pi / 64)).astype(np.float32)
    for g in (64, 96, 128):
        plan = ops.Plan2(g // 2, g // 2, t.shape[0]).set_trajectory(dev(t))
        start, perm = plan.export_bins()
        _, _, ref_start, ref_perm = O.bin2(t, g, g)
        assert np.array_equal(start.cpu().numpy().view(np.uint32), ref_start)
        assert np.array_equal(perm.cpu().numpy().view(np.uint32), ref_perm)


def test_trajectory_outside_nyquist_is_rejected(ops):
    from kspacejet_amd._lib import KsjGpuError
    t = np.zeros((10, 2), np.float32)
    t[3, 1] = 3.2
    plan = ops.Plan2(16, 16, 10)
    with pytest.raises(KsjGpuError):
        plan.set_trajectory(dev(t))
    t[3, 1] = np.nan
    with pytest.raises(KsjGpuError):
        plan.set_trajectory(dev(t))


@pytest.mark.parametrize("width", [4, 6, 8])
def test_kb_adjoint_small_matches_oracle(ops, width):
    n, coils = 32, 3
    traj, k = radial_scan(ops, n, coils, 21, 64, seed=7)
    plan = ops.Plan2(n, n, traj.shape[0], width=width).set_trajectory(traj)
    img, rss = plan.adjoint(dev(k), images=True, rss=True)
    ref = oracle_kb(k, traj.cpu().numpy(), n, width)
    for c in range(coils):
        assert rel_l2(img.cpu().numpy()[..., c], ref[..., c]) <= TOL
    assert rel_l2(rss.cpu().numpy(), O.rss(ref)) <= TOL


@pytest.mark.parametrize("item_visits", [1, 7, 64, 4096])
def test_kb_work_item_split_is_equivalent(ops, item_visits):
    # Every block split into single-visit items, odd items, the default, no split: all match the oracle.
    n, coils = 48, 2
    traj, k = radial_scan(ops, n, coils, 55, 96, seed=3)
    plan = ops.Plan2(n, n, traj.shape[0], item_visits=item_visits).set_trajectory(traj)
    img, _ = plan.adjoint(dev(k))
    ref = oracle_kb(k, traj.cpu().numpy(), n)
    assert rel_l2(img.cpu().numpy(), ref) <= TOL


@pytest.mark.parametrize("n,coils,spokes", [(128, 8, 101), (192, 12, 150), (320, 9, 203)])
def test_kb_adjoint_uneven_sizes(ops, n, coils, spokes):
    traj, k = radial_scan(ops, n, coils, spokes, 2 * n, seed=n)
    plan = ops.Plan2(n, n, traj.shape[0]).set_trajectory(traj)
    img, rss = plan.adjoint(dev(k), rss=True)
    ref = oracle_kb(k, traj.cpu().numpy(), n)
    errs = [rel_l2(img.cpu().numpy()[..., c], ref[..., c]) for c in range(coils)]
    assert max(errs) <= TOL, errs
    assert rel_l2(rss.cpu().numpy(), O.rss(ref)) <= TOL


def test_kb_adjoint_config2_full_size(ops):
    # 256x256, 16 coils, 402 golden-angle spokes x 512 samples, KB W=4 sigma=2 (G=512), ramp DCF, RSS.
    n, coils = 256, 16
    traj, k = radial_scan(ops, n, coils, 402, 512, seed=synth.scan_seed(2))
    plan = ops.Plan2(n, n, traj.shape[0]).set_trajectory(traj)
    img, rss = plan.adjoint(dev(k), rss=True)
    ref = oracle_kb(k, traj.cpu().numpy(), n)
    errs = [rel_l2(img.cpu().numpy()[..., c], ref[..., c]) for c in range(coils)]
    print("config2 per-coil rel-L2 max", max(errs), "rss", rel_l2(rss.cpu().numpy(), O.rss(ref)))
    assert max(errs) <= TOL
    assert rel_l2(rss.cpu().numpy(), O.rss(ref)) <= TOL


def test_kb_adjoint_is_deterministic(ops):
    n, coils = 64, 4
    traj, k = radial_scan(ops, n, coils, 51, 128, seed=9)
    plan = ops.Plan2(n, n, traj.shape[0], item_visits=8).set_trajectory(traj)
    a, ra = plan.adjoint(dev(k), rss=True)
    b, rb = plan.adjoint(dev(k), rss=True)
    assert torch.equal(a, b) and torch.equal(ra, rb)


def test_bilinear_route_matches_reference(ops):
    # The reference radial route (linear 2x2 taps on the unoversampled grid, origin phase,
    # unscaled inverse DFT): plan kernel=linear, grid=image, uncentred.
    n, coils, spokes, readout = 64, 3, 31, 128
    traj, k = radial_scan(ops, n, coils, spokes, readout, seed=4)
    plan = ops.Plan2(n, n, traj.shape[0], kernel="linear", width=2, oversampling=1.0, centred=False)
    plan.set_trajectory(traj)
    img, _ = plan.adjoint(dev(k))
    t = traj.cpu().numpy().astype(np.float64)
    dcf = O.ramp_dcf2(t)
    for c in range(coils):
        ref = O.radial_linear_gridding2(k[c].astype(np.complex128), t, dcf, n, n, (n - 1) / 2, (n - 1) / 2)
        assert rel_l2(img.cpu().numpy()[..., c], ref) <= TOL
        R = O.ref()
        if R is not None:  # the reference's own compiled float implementation, where it was built
            out = np.zeros((n, n), np.complex64)
            tf = traj.cpu().numpy()
            d32 = np.zeros(tf.shape[0], np.float32)
            assert R.ref_ramp_dcf2_f32(tf.shape[0], tf.reshape(-1), d32) == 0
            v = np.ascontiguousarray(k[c])
            assert R.ref_radial_linear_gridding2_f32(n, n, (n - 1) / 2, (n - 1) / 2, tf.shape[0],
                                                     v.view(np.float32), tf.reshape(-1), d32,
                                                     out.view(np.float32).reshape(-1)) == 0
            assert rel_l2(img.cpu().numpy()[..., c], out) <= 1e-4


def test_bilinear_reference_kat_off_bin(ops):
    # radial_gridding_tests.cpp:92-125: 2x2 grid, sample 2/pi at (pi/2, 0), ramp DCF, origin 0 -> 1,1,0,0.
    t = np.array([[np.pi / 2, 0.0]], np.float32)
    k = np.array([[2 / np.pi]], np.complex64)
    plan = ops.Plan2(2, 2, 1, kernel="linear", width=2, oversampling=1.0, centred=False, row_origin=0.0,
                     col_origin=0.0).set_trajectory(dev(t))
    img, _ = plan.adjoint(dev(k), layout="coil_major")
    np.testing.assert_allclose(img.cpu().numpy()[0], [[1, 1], [0, 0]], atol=1e-6)


def test_direct_adjoint_matches_oracle(ops):
    rng = np.random.default_rng(2)
    rows, cols, coils, S = 12, 10, 3, 300
    t = rng.uniform(-np.pi, np.pi, (S, 2)).astype(np.float32)
    y = (rng.standard_normal((coils, S)) + 1j * rng.standard_normal((coils, S))).astype(np.complex64)
    got = ops.nudft2_adjoint(dev(t), dev(y), rows, cols, (rows - 1) / 2, (cols - 1) / 2).cpu().numpy()
    for c in range(coils):
        ref = O.nudft2_adjoint(y[c], t.astype(np.float64), rows, cols, (rows - 1) / 2, (cols - 1) / 2)
        assert rel_l2(got[..., c], ref) <= 1e-5
    # reference KAT nufft_tests.cpp DirectAdjointWritesAllGridPoints: one DC sample 2-1i -> 2-1i everywhere
    one = ops.nudft2_adjoint(dev(np.zeros((1, 2), np.float32)), dev(np.array([[2 - 1j]], np.complex64)), 2, 3,
                             0.0, 0.0).cpu().numpy()
    np.testing.assert_allclose(one[..., 0], (2 - 1j) * np.ones((2, 3)), atol=1e-6)


def test_direct_forward_f64_matches_oracle(ops):
    rng = np.random.default_rng(8)
    rows, cols, coils, S = 9, 40, 5, 257
    t = rng.uniform(-np.pi, np.pi, (S, 2)).astype(np.float32)
    x = rng.standard_normal((coils, rows, cols)) + 1j * rng.standard_normal((coils, rows, cols))
    got = ops.nudft2_forward_f64(dev(t), dev(x), 4.0, 19.5).cpu().numpy()
    for c in range(coils):
        ref = O.nudft2_forward(x[c], t.astype(np.float64), 4.0, 19.5)
        assert rel_l2(got[c], ref) <= 1e-12
    # reference KAT nufft_tests.cpp DirectForwardUsesRowMajorPooledMatrix: 10 and -4
    img = np.array([[[1, 2], [3, 4]]], np.complex128)
    tt = np.array([[0, 0], [np.pi, 0]], np.float32)
    out = ops.nudft2_forward_f64(dev(tt), dev(img), 0.0, 0.0).cpu().numpy()[0]
    np.testing.assert_allclose(out, [10, -4], atol=1e-6)


@pytest.mark.parametrize("coils,count", [(12, 5), (16, 9), (8, 3)])
def test_batched_scans_equal_single_scans_bitwise(ops, coils, count):
    # Stacking scans of one trajectory along the coil axis changes no per-coil
    # arithmetic, so each scan's result is identical alone and in the batch.
    n = 64
    traj = ops.golden_angle2(41, 2 * n)
    plan = ops.Plan2(n, n, traj.shape[0]).set_trajectory(traj)
    scans = []
    for i in range(count):
        img = dev(synth.coil_images2(n, coils, 100 + i))
        scans.append(ops.nudft2_forward_f64(traj, img, (n - 1) / 2, (n - 1) / 2).to(torch.complex64))
    imgs, rsss = plan.adjoint_batch(scans, images=True, rss=True)
    for i, k in enumerate(scans):
        one_img, one_rss = plan.adjoint(k, images=True, rss=True)
        assert torch.equal(imgs[i], one_img) and torch.equal(rsss[i], one_rss)
    ref = oracle_kb(scans[-1].cpu().numpy(), traj.cpu().numpy(), n)
    assert rel_l2(imgs[-1].cpu().numpy(), ref) <= TOL


# ---------------------------------------------------------------- trajectory units (a13)

@pytest.mark.parametrize("units,lo,hi,enc", [("cycles_per_fov", -0.5, 0.5, (2, 2)),
                                             ("radians_per_pixel", -3.14159, 3.14159, (2, 2)),
                                             ("encoded_matrix_index", -40.0, 40.0, (96, 80))])
def test_trajectory_normalisation_is_bit_exact(ops, units, lo, hi, enc):
    rng = np.random.default_rng(4)
    raw = rng.uniform(lo, hi, (100000, 2)).astype(np.float32)
    raw[:3] = [[lo, hi], [0.0, 0.0], [hi, lo]]  # the inclusive bounds
    got = ops.normalize_trajectory2(torch.from_numpy(raw).cuda(), units, *enc).cpu().numpy()
    ref, rc = O.normalize_trajectory2(raw, units, *enc)
    assert rc == 0 and np.array_equal(got, ref)


def test_trajectory_normalisation_reports_the_first_failure(ops):
    from kspacejet_amd._lib import KsjGpuError
    raw = np.zeros((1000, 2), np.float32)
    raw[700, 0] = np.nan       # later sample, non-finite
    raw[300, 1] = 0.75         # first failure: sample 300's ky outside the cycles/FOV interval
    with pytest.raises(KsjGpuError, match="Nyquist"):
        ops.normalize_trajectory2(torch.from_numpy(raw).cuda(), "cycles_per_fov", 2, 2)
    raw[300, 1] = 0.0
    with pytest.raises(KsjGpuError, match="finite"):
        ops.normalize_trajectory2(torch.from_numpy(raw).cuda(), "cycles_per_fov", 2, 2)
    with pytest.raises(KsjGpuError, match="unspecified"):
        ops.normalize_trajectory2(torch.from_numpy(np.zeros((4, 2), np.float32)).cuda(), "unspecified", 2, 2)
    with pytest.raises(KsjGpuError, match="encoded"):
        ops.normalize_trajectory2(torch.from_numpy(np.zeros((4, 2), np.float32)).cuda(), "encoded_matrix_index", 1, 2)
