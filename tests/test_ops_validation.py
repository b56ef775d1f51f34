"""Caller tensors are checked before any pointer reaches a kernel (CPU only, no GPU):
a wrongly shaped trajectory, image or RSS buffer raises ValueError naming the shape,
and a correctly shaped CPU tensor still fails as "must be a GPU tensor" where only
device memory is served -- so the library is never called with a buffer that does not
match the launch.  A 2-D server request whose buffers are all on the CPU is a
host-memory request; one that mixes the two sides is refused."""
import types

import pytest

torch = pytest.importorskip("torch")
from kspacejet_amd import ops  # noqa: E402

c64, f32 = torch.complex64, torch.float32


def z(shape, dtype):
    return torch.zeros(shape, dtype=dtype)


def fake_plan(n, samples):
    return types.SimpleNamespace(desc=types.SimpleNamespace(image_rows=n, image_cols=n, samples=samples))


def server():
    return object.__new__(ops.Server)  # validation runs before the server handle is used


def test_nudft2_adjoint_checks_the_trajectory_rows():
    with pytest.raises(ValueError, match="shape"):
        ops.nudft2_adjoint(z((99, 2), f32), z((3, 100), c64), 8, 8, 3.5, 3.5)
    with pytest.raises(ValueError, match="GPU tensor"):
        ops.nudft2_adjoint(z((100, 2), f32), z((3, 100), c64), 8, 8, 3.5, 3.5)


@pytest.mark.parametrize("bad", ["traj", "images", "rss", "none"])
def test_submit_radial_checks_every_buffer(bad):
    n, ch, S = 16, 3, 200
    args = dict(kspace=z((ch, S), c64), traj=z((S, 2), f32), n=n, images=z((n, n, ch), c64), rss=z((n, n), f32))
    if bad == "traj":
        args["traj"] = z((S - 1, 2), f32)
    elif bad == "images":
        args["images"] = z((n, n, ch + 1), c64)
    elif bad == "rss":
        args["rss"] = z((n, n + 1), f32)
    else:
        args["images"] = args["rss"] = None
    with pytest.raises(ValueError, match="shape|images and/or rss"):
        server().submit_radial(**args)


def captured(monkeypatch):
    got = {}
    monkeypatch.setattr(ops.Server, "_submit", lambda self, req, keep, host: got.update(req=req, host=host) or 1)
    return got


def test_matching_cpu_buffers_form_a_host_memory_request(monkeypatch):
    from kspacejet_amd import _lib
    n, ch, S = 16, 3, 200
    got = captured(monkeypatch)
    server().submit_radial(z((ch, S), c64), z((S, 2), f32), n, images=z((n, n, ch), c64), rss=z((n, n), f32))
    assert got["host"] and got["req"].memory == _lib.SERVE_HOST_MEMORY
    server().submit_cartesian(z((ch, 8, 6), c64), rss=z((8, 6), f32))
    assert got["host"] and got["req"].memory == _lib.SERVE_HOST_MEMORY and got["req"].kind == _lib.SERVE_CARTESIAN2D
    with pytest.raises(ValueError, match="GPU tensor"):  # 3-D volumes are served from HBM only
        server().submit_radial3d(z((2, 50), c64), z((50, 3), f32), (4, 5, 6), rss=z((4, 5, 6), f32))


@pytest.mark.parametrize("bad", ["traj", "images", "rss"])
def test_submit_radial3d_checks_every_buffer(bad):
    shape, ch, S = (4, 5, 6), 2, 50
    traj = z((S, 2 if bad == "traj" else 3), f32)
    images = z((*shape, ch + (bad == "images")), c64)
    rss = z((4, 5, 6 + (bad == "rss")), f32)
    with pytest.raises(ValueError, match="shape"):
        server().submit_radial3d(z((ch, S), c64), traj, shape, images=images, rss=rss)


@pytest.mark.parametrize("bad", ["images", "rss"])
def test_submit_cartesian_checks_its_outputs(bad):
    images = z((8, 9, 2 + (bad == "images")), c64)
    rss = z((8, 9 + (bad == "rss")), f32)
    with pytest.raises(ValueError, match="shape"):
        server().submit_cartesian(z((2, 8, 9), c64), images=images, rss=rss)


@pytest.mark.parametrize("bad", ["samples", "images", "rss", "count"])
def test_adjoint_multi_checks_every_item(bad):
    plans = [fake_plan(8, 100), fake_plan(12, 150)]
    ks = [z((3, 100), c64), z((5, 150 - (bad == "samples")), c64)]
    imgs = [z((8, 8, 3), c64), z((12, 12, 5 + (bad == "images")), c64)]
    rss = [z((8, 8), f32), z((12, 11 if bad == "rss" else 12), f32)]
    if bad == "count":
        rss = rss[:1]
    with pytest.raises(ValueError, match="shape|one entry per plan"):
        ops.adjoint_multi(plans, ks, rss=rss, images=imgs, workspace=z(16, torch.uint8))


def test_adjoint_multi_checks_the_coil_major_layout():
    plans = [fake_plan(8, 100)]
    with pytest.raises(ValueError, match="shape"):
        ops.adjoint_multi(plans, [z((3, 100), c64)], images=[z((8, 8, 3), c64)], layout="coil_major",
                          workspace=z(16, torch.uint8))
    with pytest.raises(ValueError, match="GPU tensor"):
        ops.adjoint_multi(plans, [z((3, 100), c64)], images=[z((3, 8, 8), c64)], layout="coil_major",
                          workspace=z(16, torch.uint8))


def test_phase_correct_and_crop_check_their_shapes():
    with pytest.raises(ValueError, match="shape"):
        ops.phase_correct(z((2, 3, 4), c64), z((2, 5), c64))
    with pytest.raises(ValueError, match="inside the readout"):
        ops.readout_crop(z((2, 3, 8), c64), 6, 3)


class _OnGpu(torch.Tensor):
    """A CPU tensor that reports itself as a GPU tensor (only the device check reads it)."""
    @property
    def is_cuda(self):
        return True


def on_gpu(t):
    return torch.Tensor._make_subclass(_OnGpu, t)


def test_a_host_memory_request_keeps_every_buffer_on_the_host(monkeypatch):
    """A 2-D request with CPU k-space is a host-memory request: a GPU buffer in it is
    refused before the library sees a pointer, and a GPU request refuses CPU buffers."""
    captured(monkeypatch)
    with pytest.raises(ValueError, match="CPU tensor"):
        server().submit_radial(z((4, 100), c64), z((100, 2), f32), 8, rss=on_gpu(z((8, 8), f32)))
    with pytest.raises(ValueError, match="CPU tensor"):
        server().submit_cartesian(z((4, 8, 8), c64), images=on_gpu(z((8, 8, 4), c64)), rss=z((8, 8), f32))
    with pytest.raises(ValueError, match="GPU tensor"):
        server().submit_cartesian(on_gpu(z((4, 8, 8), c64)), rss=z((8, 8), f32))
    with pytest.raises(ValueError, match="shape"):
        server().submit_cartesian(z((4, 8, 8), c64), rss=z((8, 9), f32))
