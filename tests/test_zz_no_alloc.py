"""ksj_gpu_adjoint2_execute_static (the adjoint for callers that cannot allocate) against
ksj_gpu_adjoint2_execute, through the C ABI: two calls in a row, the second on a smaller
scan, each bit-identical to the workspace-taking entry point; no device memory is taken
by the calls; a scan whose workspace exceeds the static scratch is rejected before
anything is enqueued."""

import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ops():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from kspacejet_amd import ops as K
    K.init(0)
    return K


def scan(ops, n, coils, spokes, seed):
    traj = ops.golden_angle2(spokes, 2 * n)
    g = torch.Generator(device="cuda").manual_seed(seed)
    k = torch.complex(torch.randn((coils, traj.shape[0]), generator=g, device="cuda"),
                      torch.randn((coils, traj.shape[0]), generator=g, device="cuda"))
    plan = ops.Plan2(n, n, traj.shape[0]).set_trajectory(traj)
    return plan, k


def run_static(lib, plan, k, img, rss):
    stream = torch.cuda.current_stream().cuda_stream
    return lib.ksj_gpu_adjoint2_execute_static(plan._handle, k.data_ptr(), k.shape[0], img.data_ptr(), 0,
                                               rss.data_ptr(), stream)


def test_static_entry_matches_the_workspace_entry_on_two_calls(ops):
    from kspacejet_amd._lib import load
    lib = load()
    shapes = [(256, 16, 402, 1), (128, 8, 101, 2)]  # the second call on a smaller scan
    cases = []
    for n, coils, spokes, seed in shapes:
        plan, k = scan(ops, n, coils, spokes, seed)
        ref_img, ref_rss = plan.adjoint(k, rss=True)
        img = torch.full((n, n, coils), float("nan"), dtype=torch.complex64, device="cuda")
        rss = torch.full((n, n), float("nan"), dtype=torch.float32, device="cuda")
        cases.append((plan, k, ref_img, ref_rss, img, rss))
    torch.cuda.synchronize()
    free_before = torch.cuda.mem_get_info()[0]
    for plan, k, _, _, img, rss in cases:
        assert run_static(lib, plan, k, img, rss) == 0, lib.ksj_gpu_last_error().decode()
    torch.cuda.synchronize()
    assert torch.cuda.mem_get_info()[0] == free_before  # the calls took no device memory
    for plan, k, ref_img, ref_rss, img, rss in cases:
        assert torch.equal(img, ref_img) and torch.equal(rss, ref_rss)
        plan.close()


def test_static_entry_rejects_a_workspace_beyond_the_scratch(ops):
    from kspacejet_amd import _lib
    lib = _lib.load()
    plan, k = scan(ops, 256, 64, 402, 3)
    assert plan.workspace_bytes(64) > _lib.STATIC_SCRATCH_BYTES
    img = torch.zeros((256, 256, 64), dtype=torch.complex64, device="cuda")
    rss = torch.zeros((256, 256), dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    rc = run_static(lib, plan, k, img, rss)
    assert rc == 4  # KSJ_GPU_RESOURCE_EXHAUSTED
    assert "static scratch" in lib.ksj_gpu_last_error().decode()
    torch.cuda.synchronize()
    assert not img.any() and not rss.any()  # nothing was enqueued
    plan.close()
