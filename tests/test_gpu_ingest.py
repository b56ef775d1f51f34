"""GPU parity for the acquisition -> k-space frame assembly of the non-Cartesian route
(ksj_gpu_assemble_acquisitions; the replay loop of noncartesian_rss_hdf5.cpp:1077-1083):
bit-identical to the oracle for uniform and ragged scans, and an ingest -> normalise ->
reconstruct chain that equals the same reconstruction from the oracle's frame."""
import numpy as np
import pytest

import oracle as O

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ops():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from kspacejet_amd import ops as K
    K.init(0)
    return K


def acquisitions(counts, channels, seed):
    rng = np.random.default_rng(seed)
    return [(rng.standard_normal(n * channels) + 1j * rng.standard_normal(n * channels)).astype(np.complex64)
            for n in counts]


def assemble(ops, acqs, counts, channels):
    packed = torch.from_numpy(np.concatenate(acqs)).cuda()
    out = ops.assemble_acquisitions(packed, counts, channels)
    torch.cuda.synchronize()
    return out.cpu().numpy()


CASES = [
    ([512] * 402, 16),                                     # the headline radial scan, one spoke per acquisition
    ([256, 257, 511, 512, 513, 1024, 1025, 2048, 2049, 4097, 1, 2], 3),  # every chunk form and its edges
    ([10000, 1, 3], 2),                                     # one long acquisition among short ones
    ([128] * 33, 64),
    ([1] * 1000, 1),
    ([777], 5),
    ([4096, 2, 600, 8200], 3),                              # even counts: 16-byte accesses, ragged, long rows
]


@pytest.mark.parametrize("counts,channels", CASES)
def test_frame_is_bit_identical_to_the_oracle(ops, counts, channels):
    print(f"assemble: {len(counts)} acquisitions, {channels} channels", flush=True)
    acqs = acquisitions(counts, channels, seed=len(counts) + channels)
    got = assemble(ops, acqs, counts, channels)
    want, rc = O.assemble_acquisitions(acqs, channels)
    assert rc == 0 and got.shape == want.shape
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))


def test_buffers_off_a_16_byte_boundary(ops):
    """Even counts with the acquisitions starting 8 bytes past an aligned address: the 8-byte form."""
    counts, channels = [512] * 20 + [256, 1024], 4
    acqs = acquisitions(counts, channels, seed=13)
    flat = np.concatenate(acqs)
    buf = torch.zeros(flat.size + 1, dtype=torch.complex64, device="cuda")
    buf[1:] = torch.from_numpy(flat).cuda()
    packed = buf[1:]
    assert packed.data_ptr() % 16 == 8
    got = ops.assemble_acquisitions(packed, counts, channels).cpu().numpy()
    want, _ = O.assemble_acquisitions(acqs, channels)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))


@pytest.mark.parametrize("rows", [1, 2])
def test_rows_per_workgroup_give_the_same_frame(ops, rows):
    """assemble_rows 1 and 2 (the default 4 runs above; each falls back to a divisor of the channel
    count) against the oracle."""
    try:
        ops.tune_set("assemble_rows", rows)
        for counts, channels in CASES + [([100, 7, 2048], 6), ([513] * 9, 12)]:
            acqs = acquisitions(counts, channels, seed=channels)
            got = assemble(ops, acqs, counts, channels)
            want, _ = O.assemble_acquisitions(acqs, channels)
            assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), (rows, counts[:4], channels)
    finally:
        ops.tune_set("assemble_rows", 4)


def test_consecutive_calls_on_one_stream(ops):
    """A ragged call (offsets in the workspace) and then uniform and ragged ones of other shapes."""
    for counts, channels in [([5, 900, 33], 4), ([64] * 10, 8), ([2, 3000], 16)]:
        acqs = acquisitions(counts, channels, seed=7)
        got = assemble(ops, acqs, counts, channels)
        want, _ = O.assemble_acquisitions(acqs, channels)
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32))


def test_ingest_normalise_reconstruct_chain(ops):
    """Spokes as ISMRMRD acquisitions: samples [C][ns] and raw [ns][kx, ky] cycles/FOV each; the
    GPU-assembled frame and GPU-normalised trajectory reconstruct to the same bits as the
    oracle's frame and trajectory."""
    n, coils, spokes, readout = 64, 4, 101, 128
    traj = O.golden_angle2(spokes, readout)  # [S][ky, kx] rad/px, FP64
    raw = (traj[:, ::-1] / (2 * np.pi)).astype(np.float32)  # ISMRMRD order [kx, ky], cycles/FOV
    counts = [readout] * spokes
    acqs = acquisitions(counts, coils, seed=11)
    frame_ref, _ = O.assemble_acquisitions(acqs, coils)
    traj_ref, rc = O.normalize_trajectory2(raw, "cycles_per_fov", n, n)
    assert rc == 0

    packed = torch.from_numpy(np.concatenate(acqs)).cuda()
    frame = ops.assemble_acquisitions(packed, counts, coils)
    traj_gpu = ops.normalize_trajectory2(torch.from_numpy(np.ascontiguousarray(raw)).cuda(), "cycles_per_fov", n, n)
    plan = ops.Plan2(n, n, spokes * readout).set_trajectory(traj_gpu)
    _, rss = plan.adjoint(frame, images=False, rss=True)

    plan_ref = ops.Plan2(n, n, spokes * readout).set_trajectory(torch.from_numpy(traj_ref).cuda())
    _, rss_ref = plan_ref.adjoint(torch.from_numpy(frame_ref).cuda(), images=False, rss=True)
    torch.cuda.synchronize()
    assert np.array_equal(traj_gpu.cpu().numpy(), traj_ref)
    assert np.array_equal(rss.cpu().numpy().view(np.uint32), rss_ref.cpu().numpy().view(np.uint32))
    assert float(rss_ref.abs().max()) > 0
