"""Host-memory requests of the continuous-batching server: a 2-D request whose buffers lie
in host memory is staged through the server's pinned staging, sent to HBM on its lane
and copied back on completion.  Its results equal the same request served from HBM bit
for bit (radial and Cartesian, coil images and RSS), host and device requests of one
trajectory share one plan (the host hash equals the device hash) and stack in one
launch, a staging far smaller than the stream applies back-pressure without changing a
result, and the requests it cannot take are refused: larger than the staging, 3-D, an
unknown memory kind.  A request struct of the layout before the memory field is served
as device memory."""
import ctypes as C

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


def radial_scan(ops, traj, n, coils, seed):
    img = torch.from_numpy(synth.coil_images2(n, coils, seed)).cuda()
    k = ops.nudft2_forward_f64(traj, img, (n - 1) / 2, (n - 1) / 2).cpu().numpy()
    return torch.from_numpy(synth.add_noise(k, seed).astype(np.complex64))


def served(server, kspace, traj, n, images, host):
    """One radial request; returns (ticket, images or None, rss) in the request's memory."""
    dev = "cpu" if host else "cuda"
    k = kspace if host else kspace.cuda()
    t = traj.cpu() if host else traj
    ch = kspace.shape[0]
    img = torch.empty((n, n, ch), dtype=torch.complex64, device=dev) if images else None
    rss = torch.empty((n, n), dtype=torch.float32, device=dev)
    return server.submit_radial(k, t, n, images=img, rss=rss), img, rss


def test_host_radial_requests_equal_device_requests(ops):
    n, coils = 96, 8
    traj = ops.golden_angle2(41, 192)
    scans = [radial_scan(ops, traj, n, coils, 50 + i) for i in range(6)]
    server = ops.Server(lanes=2, max_batch_scans=4)
    want = []
    for k in scans:
        t, img, rss = served(server, k, traj, n, True, host=False)
        server.wait(t)
        want.append((img.cpu().numpy(), rss.cpu().numpy()))
    got = [served(server, k, traj, n, i % 2 == 0, host=True) for i, k in enumerate(scans)]
    for i, (t, img, rss) in enumerate(got):
        assert server.wait(t) > 0
        assert np.array_equal(rss.numpy(), want[i][1]), i
        if img is not None:
            assert np.array_equal(img.numpy(), want[i][0]), i
    st = server.stats()
    assert st["plan_builds"] == 1 and st["failed"] == 0  # the host hash found the device plan
    ref = O.kb_adjoint2_coils(scans[0].numpy(), traj.cpu().numpy().astype(np.float64),
                              O.ramp_dcf2(traj.cpu().numpy().astype(np.float64)), n)
    assert rel_l2(got[0][1].numpy(), ref) <= 1e-5
    server.close()


def test_host_and_device_requests_stack_in_one_launch(ops):
    n, coils = 64, 4
    traj = ops.golden_angle2(33, 128)
    scans = [radial_scan(ops, traj, n, coils, 60 + i) for i in range(6)]
    server = ops.Server(lanes=1, max_batch_scans=8)
    alone = []
    for i, k in enumerate(scans):
        t, img, rss = served(server, k, traj, n, False, host=i % 2 == 0)
        server.wait(t)
        alone.append(rss.cpu().numpy())
    batches = server.stats()["batches"]
    server.hold(True)
    mixed = [served(server, k, traj, n, False, host=i % 2 == 1) for i, k in enumerate(scans)]
    server.hold(False)
    for i, (t, _, rss) in enumerate(mixed):
        server.wait(t)
        assert np.array_equal(rss.cpu().numpy(), alone[i]), i
    st = server.stats()
    assert st["batches"] == batches + 1 and st["max_batch_seen"] == 6 and st["plan_builds"] == 1
    server.close()


def test_host_cartesian_requests_equal_device_requests(ops):
    server = ops.Server(lanes=2)
    for i, (n, c, centred) in enumerate([(64, 4, False), (96, 8, True), (128, 16, False)]):
        k = torch.from_numpy(synth.cartesian_kspace2(n, c, 70 + i, centred=centred).astype(np.complex64))
        img_d = torch.empty((n, n, c), dtype=torch.complex64, device="cuda")
        rss_d = torch.empty((n, n), dtype=torch.float32, device="cuda")
        server.wait(server.submit_cartesian(k.cuda(), images=img_d, rss=rss_d, centred=centred))
        img_h = torch.empty((n, n, c), dtype=torch.complex64)
        rss_h = torch.empty((n, n), dtype=torch.float32)
        server.wait(server.submit_cartesian(k, images=img_h, rss=rss_h, centred=centred))
        assert torch.equal(img_h, img_d.cpu()) and torch.equal(rss_h, rss_d.cpu()), (n, c, centred)
        rss_only = torch.empty((n, n), dtype=torch.float32)
        server.wait(server.submit_cartesian(k, rss=rss_only, centred=centred))
        assert torch.equal(rss_only, rss_d.cpu())
    server.close()


def test_pinned_and_large_host_buffers(ops):
    """8 MiB of k-space and of coil images per request: pageable buffers go through the
    staging by the parallel copy (and by one thread when serve_copy_threads is 0), pinned
    ones by DMA to and from the caller's memory itself, a pinned input with pageable
    outputs by both; every form gives the device-fed bits."""
    n, c = 256, 16
    k = torch.from_numpy(synth.cartesian_kspace2(n, c, 77, centred=False).astype(np.complex64))
    servers = {"copy_threads_default": ops.Server(lanes=2)}
    saved = ops.tune_get("serve_copy_threads")
    ops.tune_set("serve_copy_threads", 0)
    servers["copy_threads_0"] = ops.Server(lanes=2)
    ops.tune_set("serve_copy_threads", saved)
    img_d = torch.empty((n, n, c), dtype=torch.complex64, device="cuda")
    rss_d = torch.empty((n, n), dtype=torch.float32, device="cuda")
    server = servers["copy_threads_default"]
    server.wait(server.submit_cartesian(k.cuda(), images=img_d, rss=rss_d))
    want_img, want_rss = img_d.cpu(), rss_d.cpu()
    forms = {"pageable": (False, False), "pinned": (True, True), "pinned_input": (True, False)}
    for name, server in servers.items():
        for form, (pin_in, pin_out) in forms.items():
            kk = k.pin_memory() if pin_in else k.clone()
            img = torch.empty((n, n, c), dtype=torch.complex64, pin_memory=pin_out)
            rss = torch.empty((n, n), dtype=torch.float32, pin_memory=pin_out)
            tickets = [server.submit_cartesian(kk, images=img, rss=rss)]
            server.wait(tickets[0])
            assert torch.equal(img, want_img) and torch.equal(rss, want_rss), (name, form)
        server.close()


def test_a_small_staging_applies_back_pressure(ops):
    """1 MiB of staging holds about two of these requests (k-space 256 KiB, images 128 KiB,
    RSS 16 KiB each); submitting 16 blocks until earlier ones complete, and every result
    is still the device-fed one."""
    n, coils = 64, 4
    traj = ops.golden_angle2(64, 128)
    scans = [radial_scan(ops, traj, n, coils, 80 + i) for i in range(4)]
    server = ops.Server(lanes=2, host_staging_mib=1)
    want = []
    for k in scans:
        t, img, rss = served(server, k, traj, n, True, host=False)
        server.wait(t)
        want.append((img.cpu().numpy(), rss.cpu().numpy()))
    got = [served(server, scans[i % 4], traj, n, True, host=True) for i in range(16)]
    for i, (t, img, rss) in enumerate(got):
        server.wait(t)
        assert np.array_equal(img.numpy(), want[i % 4][0]) and np.array_equal(rss.numpy(), want[i % 4][1]), i
    assert server.stats()["failed"] == 0
    server.close()


def test_requests_the_staging_cannot_take_are_refused(ops):
    from kspacejet_amd import _lib
    from kspacejet_amd._lib import KsjGpuError
    server = ops.Server(lanes=1, host_staging_mib=1)
    big = torch.zeros((16, 256, 256), dtype=torch.complex64)  # 8 MiB of k-space
    with pytest.raises(KsjGpuError) as e:
        server.submit_cartesian(big, rss=torch.empty((256, 256), dtype=torch.float32))
    assert e.value.code == 1 and "staging" in str(e.value)
    traj3 = ops.golden_means3(10, 8).cpu()
    with pytest.raises(ValueError, match="GPU tensor"):
        server.submit_radial3d(torch.zeros((2, traj3.shape[0]), dtype=torch.complex64), traj3, (4, 4, 4),
                               rss=torch.empty((4, 4, 4)))
    # 3-D in host memory and an unknown memory kind, straight through the C ABI
    lib = _lib.load()
    k = torch.zeros((2, traj3.shape[0]), dtype=torch.complex64)
    out = torch.empty((4, 4, 4), dtype=torch.float32)
    for kind, memory in ((_lib.SERVE_RADIAL3D, _lib.SERVE_HOST_MEMORY), (_lib.SERVE_CARTESIAN2D, 7)):
        r = _lib.ServeRequest(C.sizeof(_lib.ServeRequest), kind, 4, 4, 2, traj3.shape[0], 4,
                              _lib.DENSITY_NONE, 0, 4, 0, traj3.data_ptr(), k.data_ptr(), None, out.data_ptr(),
                              memory, 0)
        t = C.c_uint64(0)
        assert lib.ksj_gpu_serve_submit(server._handle, C.byref(r), C.byref(t)) == 1
    assert server.stats()["submitted"] == 0
    server.close()


def test_a_request_struct_without_the_memory_field_is_device_memory(ops):
    from kspacejet_amd import _lib
    n, c = 64, 4
    k = torch.from_numpy(synth.cartesian_kspace2(n, c, 90, centred=False).astype(np.complex64)).cuda()
    rss = torch.empty((n, n), dtype=torch.float32, device="cuda")
    server = ops.Server(lanes=1)
    r = _lib.ServeRequest(C.sizeof(_lib.ServeRequest), _lib.SERVE_CARTESIAN2D, n, n, c, 0, 0, 0, 0, 0, 0, None,
                          k.data_ptr(), None, rss.data_ptr(), 5, 0)  # memory 5 lies past the shorter struct
    r.struct_size = _lib.ServeRequest.memory.offset
    t = C.c_uint64(0)
    _lib.check(_lib.load().ksj_gpu_serve_submit(server._handle, C.byref(r), C.byref(t)))
    server.wait(t.value)
    assert rel_l2(rss.cpu().numpy(), O.rss(O.ifft2_coil_images(k.cpu().numpy()))) <= 1e-5
    r.struct_size = _lib.ServeRequest.memory.offset - 8
    assert _lib.load().ksj_gpu_serve_submit(server._handle, C.byref(r), C.byref(t)) == 1
    server.close()
