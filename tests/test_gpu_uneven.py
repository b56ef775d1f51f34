"""Uneven batches: items of different matrix, coil count and trajectory in one
multi-plan call (ksj_gpu_adjoint2_execute_multi), fused (one launch per stage and
shape class, with or without the gridding work queue) or over lanes, eagerly and as a
captured HIP graph.  Each item must
equal its single-plan result bit for bit and the FP64 oracle to rel-L2 <= 1e-5; every
graph replay must equal the eager result."""
import numpy as np
import pytest

import oracle as O
from kspacejet_amd import synth
from tests.helpers import per_coil, rel_l2

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

TOL = 1e-5
# (matrix, coils, spokes) of each item; readout = 2 x matrix.
ITEMS = [(32, 3, 21), (48, 7, 33), (64, 12, 45), (40, 5, 27), (56, 9, 39), (64, 4, 61)]


@pytest.fixture(scope="module")
def ops():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from kspacejet_amd import ops as K
    K.init(0)
    return K


@pytest.fixture(scope="module", params=[2, 0], ids=["matrix", "blocks"])
def batch(ops, request):
    """The items' plans, built with the gather schedule of the parameter."""
    saved = ops.tune_get("gather_scheme")
    ops.tune_set("gather_scheme", request.param)
    plans, ks, trajs = [], [], []
    for i, (n, c, spokes) in enumerate(ITEMS):
        traj = ops.golden_angle2(spokes, 2 * n)
        img = torch.from_numpy(synth.coil_images2(n, c, synth.scan_seed(3, i))).cuda()
        k = ops.nudft2_forward_f64(traj, img, (n - 1) / 2, (n - 1) / 2).cpu().numpy()
        k = torch.from_numpy(synth.add_noise(k, synth.scan_seed(3, i)).astype(np.complex64)).cuda()
        plans.append(ops.Plan2(n, n, traj.shape[0]).set_trajectory(traj))
        ks.append(k)
        trajs.append(traj)
    ops.tune_set("gather_scheme", saved)
    return plans, ks, trajs


def outputs(batch):
    plans, ks, _ = batch
    rss = [torch.zeros((p.desc.image_rows, p.desc.image_cols), dtype=torch.float32, device="cuda") for p in plans]
    imgs = [torch.zeros((p.desc.image_rows, p.desc.image_cols, k.shape[0]), dtype=torch.complex64, device="cuda")
            for p, k in zip(plans, ks)]
    return rss, imgs


def test_items_match_single_plans_and_the_oracle(ops, batch):
    plans, ks, trajs = batch
    rss, imgs = outputs(batch)
    ws = torch.empty(ops.adjoint_multi_workspace_bytes(plans, [k.shape[0] for k in ks]), dtype=torch.uint8,
                     device="cuda")
    ops.adjoint_multi(plans, ks, rss=rss, images=imgs, workspace=ws)
    torch.cuda.synchronize()
    for i, (plan, k, traj) in enumerate(zip(plans, ks, trajs)):
        img1, rss1 = plan.adjoint(k, rss=True)
        assert torch.equal(imgs[i], img1) and torch.equal(rss[i], rss1), i
        n = ITEMS[i][0]
        t = traj.cpu().numpy().astype(np.float64)
        kk = k.cpu().numpy()
        dcf = O.ramp_dcf2(t)
        ref = np.stack(per_coil(lambda c: O.kb_adjoint2(kk[c], t, dcf, n, n), kk.shape[0]), axis=-1)
        assert rel_l2(imgs[i].cpu().numpy(), ref) <= TOL, i
        assert rel_l2(rss[i].cpu().numpy(), O.rss(ref)) <= TOL, i


def test_fused_and_lane_forms_give_identical_results(ops, batch):
    """The work-queue chain (multi_fused = 2), one launch per stage and shape class (1)
    and items spread over lanes (0) run the same per-item arithmetic: bit-identical
    images and RSS, in both image layouts."""
    plans, ks, _ = batch
    saved = ops.tune_get("multi_fused")
    results = []
    try:
        for fused in (2, 1, 0):
            ops.tune_set("multi_fused", fused)
            for layout in ("coil_last", "coil_major"):
                rss = [torch.zeros((p.desc.image_rows, p.desc.image_cols), dtype=torch.float32, device="cuda")
                       for p in plans]
                shape = (lambda p, k: (p.desc.image_rows, p.desc.image_cols, k.shape[0])) if layout == "coil_last" \
                    else (lambda p, k: (k.shape[0], p.desc.image_rows, p.desc.image_cols))
                imgs = [torch.zeros(shape(p, k), dtype=torch.complex64, device="cuda") for p, k in zip(plans, ks)]
                ws = torch.empty(ops.adjoint_multi_workspace_bytes(plans, [k.shape[0] for k in ks]),
                                 dtype=torch.uint8, device="cuda")
                ops.adjoint_multi(plans, ks, rss=rss, images=imgs, workspace=ws, layout=layout)
                torch.cuda.synchronize()
                results.append([r.cpu().numpy() for r in rss] + [m.cpu().numpy() for m in imgs])
    finally:
        ops.tune_set("multi_fused", saved)
    for form in (1, 2):
        for a, b in zip(results[:2], results[2 * form:2 * form + 2]):
            assert all(np.array_equal(x, y) for x, y in zip(a, b)), form
    assert ops.chain_faults() == 0


def test_fused_form_with_staged_ffts_matches_single_plans(ops, batch):
    """With the LDS-staged FFT everywhere (fft_register = 0) the fused form's coil-last
    column pass (coil-fastest lines, column weights by plane) still gives every item's
    single-plan bits, at LDS and direct-DFT lengths alike."""
    plans, ks, _ = batch
    saved = ops.tune_get("multi_fused"), ops.tune_get("fft_register")
    try:
        ops.tune_set("fft_register", 0)
        ops.tune_set("multi_fused", 1)
        rss, imgs = outputs(batch)
        ws = torch.empty(ops.adjoint_multi_workspace_bytes(plans, [k.shape[0] for k in ks]), dtype=torch.uint8,
                         device="cuda")
        ops.adjoint_multi(plans, ks, rss=rss, images=imgs, workspace=ws)
        torch.cuda.synchronize()
        for i, (plan, k) in enumerate(zip(plans, ks)):
            img1, rss1 = plan.adjoint(k, rss=True)
            assert torch.equal(imgs[i], img1) and torch.equal(rss[i], rss1), i
    finally:
        ops.tune_set("multi_fused", saved[0])
        ops.tune_set("fft_register", saved[1])


def test_chain_form_runs_its_work_queue(ops, batch, request):
    """multi_fused = 2 grids matrix-schedule items in the work-queue launch (its stage is
    recorded as multi_grid_chain) and falls back to the per-class launches for the 2x2-block
    schedule; no bounded wait runs out.  A single item goes through the chain too."""
    plans, ks, _ = batch
    matrix = request.node.callspec.id == "matrix"
    saved = ops.tune_get("multi_fused"), ops.tune_get("multi_lanes")
    ops.tune_set("multi_fused", 2)
    try:
        for lanes in (1, 3):
            ops.tune_set("multi_lanes", lanes)
            ops.profile_reset()
            ops.profile_enable(True)
            rss, imgs = outputs(batch)
            ws = torch.empty(ops.adjoint_multi_workspace_bytes(plans, [k.shape[0] for k in ks]), dtype=torch.uint8,
                             device="cuda")
            ops.adjoint_multi(plans, ks, rss=rss, images=imgs, workspace=ws)
            torch.cuda.synchronize()
            ops.profile_enable(False)
            chain = ops.profile_read("multi_grid_chain")[1]
            staged = ops.profile_read("multi_grid")[1]
            assert (chain, staged) == ((lanes, 0) if matrix else (0, lanes)), (lanes, chain, staged)
            for i, (plan, k) in enumerate(zip(plans, ks)):
                img1, rss1 = plan.adjoint(k, rss=True)
                assert torch.equal(imgs[i], img1) and torch.equal(rss[i], rss1), (lanes, i)
        one_rss = [torch.zeros_like(rss[2])]
        one_img = [torch.zeros_like(imgs[2])]
        ws = torch.empty(ops.adjoint_multi_workspace_bytes(plans[2:3], [ks[2].shape[0]]), dtype=torch.uint8,
                         device="cuda")
        ops.adjoint_multi(plans[2:3], ks[2:3], rss=one_rss, images=one_img, workspace=ws)
        torch.cuda.synchronize()
        img1, rss1 = plans[2].adjoint(ks[2], rss=True)
        assert torch.equal(one_img[0], img1) and torch.equal(one_rss[0], rss1)
    finally:
        ops.profile_enable(False)
        ops.tune_set("multi_fused", saved[0])
        ops.tune_set("multi_lanes", saved[1])
    assert ops.chain_faults() == 0


def test_item_lanes_do_not_change_results(ops, batch):
    plans, ks, _ = batch
    saved = ops.tune_get("multi_lanes"), ops.tune_get("multi_fused")
    ops.tune_set("multi_fused", 0)
    results = []
    try:
        for lanes in (1, 3, 8):
            ops.tune_set("multi_lanes", lanes)
            rss, imgs = outputs(batch)
            ws = torch.empty(ops.adjoint_multi_workspace_bytes(plans, [k.shape[0] for k in ks]), dtype=torch.uint8,
                             device="cuda")
            ops.adjoint_multi(plans, ks, rss=rss, images=imgs, workspace=ws)
            torch.cuda.synchronize()
            results.append([r.cpu().numpy() for r in rss] + [m.cpu().numpy() for m in imgs])
    finally:
        ops.tune_set("multi_lanes", saved[0])
        ops.tune_set("multi_fused", saved[1])
    for res in results[1:]:
        assert all(np.array_equal(a, b) for a, b in zip(res, results[0]))


@pytest.mark.parametrize("fused", [0, 1, 2])
def test_captured_graph_replays_equal_the_eager_result(ops, batch, fused):
    plans, ks, _ = batch
    saved = ops.tune_get("multi_fused")
    ops.tune_set("multi_fused", fused)
    stream = torch.cuda.Stream()
    rss, imgs = outputs(batch)
    ws = torch.empty(ops.adjoint_multi_workspace_bytes(plans, [k.shape[0] for k in ks]), dtype=torch.uint8,
                     device="cuda")
    torch.cuda.synchronize()  # outputs were zeroed on the default stream
    with torch.cuda.stream(stream):
        ops.adjoint_multi(plans, ks, rss=rss, images=imgs, workspace=ws)  # eager run (also warms every FFT length)
    stream.synchronize()
    eager = [r.clone() for r in rss] + [m.clone() for m in imgs]
    graph = ops.Graph(stream)
    with graph:
        ops.adjoint_multi(plans, ks, rss=rss, images=imgs, workspace=ws)
    for replay in range(3):
        for t in rss + imgs:  # the program's own eager work between replays
            t.zero_()
        torch.cuda.synchronize()
        graph.replay()
        stream.synchronize()
        got = rss + imgs
        assert all(torch.equal(a, b) for a, b in zip(got, eager)), replay
    graph.close()
    ops.tune_set("multi_fused", saved)
