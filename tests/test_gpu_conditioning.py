"""The k-space conditioning kernels (ksj_gpu_channel_mix, ksj_gpu_phase_correct,
ksj_gpu_readout_crop) against the oracle: the reference's known answers exactly, the
channel mix within rel-L2 1e-6 of the FP64 product (FP32 accumulation over at most 64
channels) and bit-exact for an identity matrix, the phase correction and the crop
bit-exact at every shape, up to the operators' largest frames."""
import numpy as np
import pytest

import oracle as O
from kspacejet_amd import provider_host as H
from tests.helpers import rel_l2

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

TOL = 1e-6


@pytest.fixture(scope="module")
def ops():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from kspacejet_amd import ops as K
    K.init(0)
    return K


@pytest.fixture(params=[1, 0], ids=["matrix_cores", "vector"])
def scheme(request, ops):
    """Both channel-mix schedules (tuning key mix_scheme)."""
    saved = ops.tune_get("mix_scheme")
    ops.tune_set("mix_scheme", request.param)
    yield request.param
    ops.tune_set("mix_scheme", saved)


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def crandn(rng, shape):
    return (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(np.complex64)


def test_reference_kats(ops, scheme):
    x = np.arange(1, 9, dtype=np.float32).reshape(2, 2, 2).astype(np.complex64)
    y = ops.channel_mix(dev(np.diag([2, 3]).astype(np.complex64)), dev(x)).cpu().numpy()
    assert np.array_equal(y.reshape(-1), np.array([2, 4, 6, 8, 15, 18, 21, 24], np.complex64))
    y = ops.channel_mix(dev(np.array([[1, 2]], np.complex64)), dev(x)).cpu().numpy()
    assert np.array_equal(y.reshape(-1), np.array([11, 14, 17, 20], np.complex64))
    xn = np.arange(1, 7, dtype=np.float32).reshape(2, 3).astype(np.complex64)
    assert np.array_equal(ops.channel_mix(dev(np.diag([2, 3]).astype(np.complex64)), dev(xn)).cpu().numpy().reshape(-1),
                          np.array([2, 4, 6, 12, 15, 18], np.complex64))
    assert np.array_equal(ops.channel_mix(dev(np.array([[1, 2]], np.complex64)), dev(xn)).cpu().numpy().reshape(-1),
                          np.array([9, 12, 15], np.complex64))
    k = np.array([1, 2, 3, 4, 1 + 1j, 2 + 1j, 3 + 1j, 4 + 1j], np.complex64).reshape(2, 2, 2)
    p = np.array([1j, -1j, 1, -1], np.complex64).reshape(2, 2)
    assert np.array_equal(ops.phase_correct(dev(k), dev(p)).cpu().numpy(), O.phase_correct(k, p))
    f = np.array([[[100 * c + 10 * r + q for q in range(4)] for r in range(2)] for c in range(2)], np.complex64)
    assert np.array_equal(ops.readout_crop(dev(f), 2, 1).cpu().numpy().reshape(-1).real,
                          [1, 2, 11, 12, 101, 102, 111, 112])


@pytest.mark.parametrize("out_ch,in_ch,samples", [
    (1, 1, 1), (2, 3, 255), (4, 4, 257), (5, 8, 1000), (16, 16, 205824), (8, 32, 4099), (12, 17, 3000),
    (33, 40, 777), (64, 64, 65536), (1, 64, 300)])
def test_channel_mix_matches_the_oracle(ops, scheme, out_ch, in_ch, samples):
    rng = np.random.default_rng(out_ch * 131 + in_ch)
    w = crandn(rng, (out_ch, in_ch))
    x = crandn(rng, (in_ch, samples))
    got = ops.channel_mix(dev(w), dev(x)).cpu().numpy()
    ref = O.channel_mix(w, x)
    assert rel_l2(got, ref) <= TOL
    # every output channel on its own, so one wrong row cannot hide in the total
    for o in range(out_ch):
        assert rel_l2(got[o], ref[o]) <= 4 * TOL, o


@pytest.mark.parametrize("channels", [1, 7, 16, 64])
def test_identity_matrix_copies_bits(ops, scheme, channels):
    rng = np.random.default_rng(channels)
    x = crandn(rng, (channels, 96, 80))
    got = ops.channel_mix(dev(np.eye(channels, dtype=np.complex64)), dev(x)).cpu().numpy()
    assert np.array_equal(got.view(np.uint32), x.view(np.uint32))


def test_cartesian_frame_mix_at_the_largest_frame(ops, scheme):
    rng = np.random.default_rng(11)
    w = crandn(rng, (16, 64))
    x = crandn(rng, (64, 512, 512))
    got = ops.channel_mix(dev(w), dev(x)).cpu().numpy()
    ref = O.channel_mix(w[:, :], x[:, :8, :])  # the oracle on the first 8 lines of every channel
    assert rel_l2(got[:, :8, :], ref) <= TOL
    again = ops.channel_mix(dev(w), dev(x)).cpu().numpy()
    assert np.array_equal(got, again)


@pytest.mark.parametrize("shape", [(1, 2, 2), (2, 2, 3), (8, 33, 65), (64, 512, 512), (3, 1, 512)])
def test_phase_correct_is_bit_exact(ops, shape):
    rng = np.random.default_rng(sum(shape))
    x = crandn(rng, shape)
    p = np.exp(1j * rng.uniform(-np.pi, np.pi, (shape[0], shape[2]))).astype(np.complex64)
    got = ops.phase_correct(dev(x), dev(p)).cpu().numpy()
    assert np.array_equal(got.view(np.uint32), O.phase_correct(x, p).view(np.uint32))


@pytest.mark.parametrize("shape,out_cols,offset", [((2, 2, 4), 2, 1), ((8, 64, 512), 256, 128), ((1, 1, 2), 2, 0),
                                                   ((64, 512, 512), 256, 0), ((5, 7, 300), 1, 299)])
def test_readout_crop_is_bit_exact(ops, shape, out_cols, offset):
    rng = np.random.default_rng(out_cols + offset)
    x = crandn(rng, shape)
    got = ops.readout_crop(dev(x), out_cols, offset).cpu().numpy()
    assert np.array_equal(got.view(np.uint32), O.readout_crop(x, out_cols, offset).view(np.uint32))


def test_argument_checks(ops):
    x = dev(np.zeros((65, 4), np.complex64))
    with pytest.raises(RuntimeError, match="channel counts"):
        ops.channel_mix(dev(np.zeros((1, 65), np.complex64)), x)
    with pytest.raises(ValueError, match="inside the readout"):  # the wrapper checks first
        ops.readout_crop(dev(np.zeros((2, 2, 4), np.complex64)), 3, 2)
    # the library checks too, for callers of the C ABI
    from kspacejet_amd import _lib
    src, dst = dev(np.zeros((2, 2, 4), np.complex64)), dev(np.zeros((2, 2, 3), np.complex64))
    with pytest.raises(RuntimeError, match="crop range"):
        _lib.call("ksj_gpu_readout_crop", src.data_ptr(), dst.data_ptr(), 4, 4, 3, 2,
                  torch.cuda.current_stream().cuda_stream)


# ---------------------------------------------------------------- through the Provider ABI


@pytest.fixture(scope="module")
def module(ops):
    return H.ProviderModule("kspace_conditioning")


def fire(module, op_id, config, dyn, dyn_type, cal=None, cal_type=None, out_bytes=0, keys=(5, 9)):
    st, op, msg = module.create(op_id, config)
    assert st == H.OK, msg
    assert op.open(0) == H.OK
    batches = [H.Batch(0, [H.Item(H.aligned_copy(dyn), dyn_type, *keys)])]
    if cal is not None:
        batches.append(H.Batch(1, [H.Item(H.aligned_copy(cal), cal_type, 0, 0)]))
    ls = H.Lease(batches, H.OutputSpec(out_bytes, dyn_type), scratch_bytes=0)
    st, r, msg = op.process(ls)
    assert st == H.OK, msg
    # one sealed output with the k-space item's keys and empty metadata; every input consumed
    assert r.outcome == H.PROCESS_DONE and r.sealed_output_count == 1 and ls.seal_calls == 1
    assert r.consumed_input_item_count == len(batches) and r.terminal_epoch == ls.terminal_epoch
    assert (ls.sealed.semantic_key_hash, ls.sealed.order_key) == keys and ls.sealed_metadata == b""
    op.close()
    return ls.output.view(np.complex64)


def test_provider_reference_kats(module):
    """kspace_conditioning_provider_tests.cpp:539-729, exact."""
    x = np.arange(1, 9, dtype=np.float32).astype(np.complex64)
    w = np.array([2, 0, 0, 3], np.complex64)
    cfg = '{"channel_count":2,"cols":2,"rows":2}'
    assert np.array_equal(fire(module, "noise_prewhiten", cfg, x, "ksj.kspace-frame", w, "ksj.noise-model", 64),
                          np.array([2, 4, 6, 8, 15, 18, 21, 24], np.complex64))
    k = np.array([1, 2, 3, 4, 1 + 1j, 2 + 1j, 3 + 1j, 4 + 1j], np.complex64)
    p = np.array([1j, -1j, 1, -1], np.complex64)
    assert np.array_equal(fire(module, "phase_correct", cfg, k, "ksj.kspace-frame", p, "ksj.phase-model", 64),
                          np.array([1j, -2j, 3j, -4j, 1 + 1j, -2 - 1j, 3 + 1j, -4 - 1j], np.complex64))
    got = fire(module, "coil_compress", '{"cols":2,"physical_channel_count":2,"rows":2,"virtual_channel_count":1}', x,
               "ksj.kspace-frame", np.array([1, 2], np.complex64), "ksj.coil-compression-basis", 32)
    assert np.array_equal(got, np.array([11, 14, 17, 20], np.complex64))
    xn = np.arange(1, 7, dtype=np.float32).astype(np.complex64)
    got = fire(module, "noncartesian_noise_prewhiten", '{"channel_count":2,"sample_count":3}', xn,
               "ksj.noncartesian-kspace-frame", w, "ksj.noise-model", 48)
    assert np.array_equal(got, np.array([2, 4, 6, 12, 15, 18], np.complex64))
    got = fire(module, "noncartesian_coil_compress",
               '{"physical_channel_count":2,"sample_count":3,"virtual_channel_count":1}', xn,
               "ksj.noncartesian-kspace-frame", np.array([1, 2], np.complex64), "ksj.coil-compression-basis", 24)
    assert np.array_equal(got, np.array([9, 12, 15], np.complex64))
    f = np.array([[[100 * c + 10 * r + q for q in range(4)] for r in range(2)] for c in range(2)], np.complex64)
    got = fire(module, "readout_oversampling_remove",
               '{"channel_count":2,"input_cols":4,"output_cols":2,"readout_offset":1,"rows":2}', f.reshape(-1),
               "ksj.kspace-frame", out_bytes=64)
    assert np.array_equal(got.real, [1, 2, 11, 12, 101, 102, 111, 112])


def test_provider_operators_at_scan_sizes_match_the_oracle(module):
    rng = np.random.default_rng(21)
    C, V, R, K, S = 16, 8, 96, 256, 65536
    x = crandn(rng, (C, R, K))
    wn = crandn(rng, (C, C))
    got = fire(module, "noise_prewhiten", f'{{"channel_count":{C},"cols":{K},"rows":{R}}}', x, "ksj.kspace-frame", wn,
               "ksj.noise-model", C * R * K * 8).reshape(C, R, K)
    assert rel_l2(got, O.channel_mix(wn, x)) <= TOL
    b = crandn(rng, (V, C))
    got = fire(module, "coil_compress",
               f'{{"cols":{K},"physical_channel_count":{C},"rows":{R},"virtual_channel_count":{V}}}', x,
               "ksj.kspace-frame", b, "ksj.coil-compression-basis", V * R * K * 8).reshape(V, R, K)
    assert rel_l2(got, O.channel_mix(b, x)) <= TOL
    p = np.exp(1j * rng.uniform(-np.pi, np.pi, (C, K))).astype(np.complex64)
    got = fire(module, "phase_correct", f'{{"channel_count":{C},"cols":{K},"rows":{R}}}', x, "ksj.kspace-frame", p,
               "ksj.phase-model", C * R * K * 8).reshape(C, R, K)
    assert np.array_equal(got.view(np.uint32), O.phase_correct(x, p).view(np.uint32))
    got = fire(module, "readout_oversampling_remove",
               f'{{"channel_count":{C},"input_cols":{K},"output_cols":{K // 2},"readout_offset":{K // 4},'
               f'"rows":{R}}}', x, "ksj.kspace-frame", out_bytes=C * R * K // 2 * 8).reshape(C, R, K // 2)
    assert np.array_equal(got, O.readout_crop(x, K // 2, K // 4))
    xn = crandn(rng, (C, S))
    got = fire(module, "noncartesian_noise_prewhiten", f'{{"channel_count":{C},"sample_count":{S}}}', xn,
               "ksj.noncartesian-kspace-frame", wn, "ksj.noise-model", C * S * 8).reshape(C, S)
    assert rel_l2(got, O.channel_mix(wn, xn)) <= TOL
    got = fire(module, "noncartesian_coil_compress",
               f'{{"physical_channel_count":{C},"sample_count":{S},"virtual_channel_count":{V}}}', xn,
               "ksj.noncartesian-kspace-frame", b, "ksj.coil-compression-basis", V * S * 8).reshape(V, S)
    assert rel_l2(got, O.channel_mix(b, xn)) <= TOL
