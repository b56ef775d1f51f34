"""Oracle pins for the acquisition -> k-space frame assembly of the non-Cartesian route
(noncartesian_rss_hdf5.cpp:1077-1083), and the C-ABI's argument checks (no GPU needed)."""
import ctypes as C

import numpy as np
import pytest

import oracle as O


def labelled(counts, channels):
    """Acquisition a, channel c, sample s holds a + i (4096 c + s): every value names its origin
    exactly in complex64."""
    out = []
    for a, n in enumerate(counts):
        c, s = np.meshgrid(np.arange(channels), np.arange(n), indexing="ij")
        out.append((a + 1j * (4096 * c + s)).astype(np.complex64).reshape(-1))
    return out


@pytest.mark.parametrize("counts,channels", [([512] * 7, 16), ([1, 5, 4095, 2, 300], 3), ([17], 1),
                                             ([3, 1, 2], 64)])
def test_frame_columns_name_their_acquisition_channel_and_sample(counts, channels):
    frame, rc = O.assemble_acquisitions(labelled(counts, channels), channels)
    assert rc == 0 and frame.shape == (channels, sum(counts))
    # Column j of the frame belongs to the acquisition whose sample range holds j (found by
    # search over the running totals, not by the copy loop).
    ends = np.cumsum(counts)
    j = np.arange(sum(counts))
    a = np.searchsorted(ends, j, side="right")
    s = j - (ends[a] - np.asarray(counts)[a])
    c = np.arange(channels)[:, None]
    assert np.array_equal(frame.real, np.broadcast_to(a, frame.shape).astype(np.float32))
    assert np.array_equal(frame.imag, (4096 * c + s[None, :]).astype(np.float32))


def test_frame_is_the_concatenation_of_the_channel_major_acquisitions():
    rng = np.random.default_rng(3)
    counts, channels = [64, 1, 129, 7], 5
    acqs = [(rng.standard_normal(n * channels) + 1j * rng.standard_normal(n * channels)).astype(np.complex64)
            for n in counts]
    frame, rc = O.assemble_acquisitions(acqs, channels)
    assert rc == 0
    want = np.concatenate([a.reshape(channels, n) for a, n in zip(acqs, counts)], axis=1)
    assert np.array_equal(frame.view(np.uint32), want.view(np.uint32))


def test_oracle_rejects_what_the_route_rejects():
    assert O.assemble_acquisitions([], 4)[1] == 1
    assert O.assemble_acquisitions([np.zeros(8, np.complex64)], 0)[1] == 1
    assert O.assemble_acquisitions([np.zeros(65, np.complex64)], 65)[1] == 1
    assert O.assemble_acquisitions([np.zeros(8, np.complex64), np.zeros(0, np.complex64)], 4)[1] == 1


def test_abi_rejects_bad_arguments_before_any_device_call():
    from kspacejet_amd import _lib
    lib = _lib.load()
    assert lib.ksj_gpu_assemble_workspace_bytes(402) == 4 * 403
    counts = (C.c_uint32 * 3)(512, 0, 512)
    good = (C.c_uint32 * 3)(512, 512, 512)
    fake = C.c_void_p(0x1000)  # never dereferenced: every case fails its checks first
    assert lib.ksj_gpu_assemble_acquisitions(fake, good, 0, 16, fake, None, 0, None) == 1
    assert lib.ksj_gpu_assemble_acquisitions(fake, counts, 3, 16, fake, None, 0, None) == 1
    assert lib.ksj_gpu_assemble_acquisitions(fake, good, 3, 0, fake, None, 0, None) == 1
    assert lib.ksj_gpu_assemble_acquisitions(fake, good, 3, 65, fake, None, 0, None) == 1
    assert lib.ksj_gpu_assemble_acquisitions(None, good, 3, 16, fake, None, 0, None) == 1
    assert lib.ksj_gpu_assemble_acquisitions(fake, None, 3, 16, fake, None, 0, None) == 1
    ragged = (C.c_uint32 * 3)(512, 1, 512)
    assert lib.ksj_gpu_assemble_acquisitions(fake, ragged, 3, 16, fake, None, 0, None) == 1
    assert lib.ksj_gpu_assemble_acquisitions(fake, ragged, 3, 16, fake, fake, 15, None) == 1
    big = (C.c_uint32 * 2)(0xFFFFFFFF, 1)
    assert lib.ksj_gpu_assemble_acquisitions(fake, big, 2, 1, fake, fake, 64, None) == 1


def test_ops_wrapper_rejects_bad_shapes_before_any_launch():
    torch = pytest.importorskip("torch")
    from kspacejet_amd import ops
    x = torch.zeros(16 * 1024, dtype=torch.complex64)
    with pytest.raises(ValueError):
        ops.assemble_acquisitions(x, [], 16)
    with pytest.raises(ValueError):
        ops.assemble_acquisitions(x, [512, 0, 512], 16)
    with pytest.raises(ValueError):
        ops.assemble_acquisitions(x, [512, 512], 65)
    with pytest.raises(ValueError):  # one sample short
        ops.assemble_acquisitions(x, [512, 511], 16)
    with pytest.raises(ValueError):  # right size, but host memory
        ops.assemble_acquisitions(x, [512, 512], 16)


def test_rows_per_workgroup_key_accepts_1_2_4_only():
    from kspacejet_amd import _lib
    lib = _lib.load()
    v = C.c_int64()
    assert lib.ksj_gpu_tune_get(b"assemble_rows", C.byref(v)) == 0 and v.value == 4
    assert lib.ksj_gpu_tune_set(b"assemble_rows", 3) != 0
    assert lib.ksj_gpu_tune_set(b"assemble_rows", 1) == 0
    assert lib.ksj_gpu_tune_set(b"assemble_rows", 4) == 0


def test_c_copy_loop_of_the_cpu_baseline_equals_the_oracle():
    counts, channels = [512, 1, 300, 4096, 7], 6
    acqs = [(np.arange(n * channels) * (1 + 1j) + a).astype(np.complex64) for a, n in enumerate(counts)]
    want, _ = O.assemble_acquisitions(acqs, channels)
    got, secs = O.assemble_acquisitions_c(np.concatenate(acqs), counts, channels, frames=3)
    assert secs >= 0 and np.array_equal(got.view(np.uint32), want.view(np.uint32))
    assert O.assemble_acquisitions_c(np.zeros(8, np.complex64), [4, 0], 1)[0] is None
    assert O.assemble_acquisitions_c(np.zeros(65, np.complex64), [1], 65)[0] is None
