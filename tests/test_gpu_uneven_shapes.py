"""The uneven radial batch at its configured shapes: 64 items drawn from seeds (matrix
128/192/256/320, so grids of 256/384/512/640; 8-32 coils, so every padded coil class
8/16/32; 101-603 golden-angle spokes, readout 2 x matrix), simulated as the benchmark
simulates them (ellipse phantom x Gaussian coils, exact FP64 forward DFT, noise
1e-3 max|k|).  One multi-plan call runs them in each form -- the gridding work queue
(one persistent launch for every item's prep, gather and reduce), fused (one launch per
stage and shape class) and over lanes -- eagerly and as replays of a captured HIP graph:

* every item equals its single-plan result bit for bit, in both forms;
* two coils of every item (the first and the last) are within rel-L2 1e-5 of the FP64
  oracle; the per-item errors are printed;
* every graph replay equals the eager result.

Both gather schedules (4x4 blocks on the matrix cores, 2x2 blocks on the VALU) run."""
import numpy as np
import pytest

import oracle as O
from kspacejet_amd import synth
from tests.helpers import rel_l2

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

TOL = 1e-5
COUNT = 64
_refs = {}  # oracle images per (item, coil): the same data under both schedules


@pytest.fixture(scope="module")
def ops():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from kspacejet_amd import ops as K
    K.init(0)
    return K


@pytest.fixture(scope="module")
def data(ops):
    shapes = synth.uneven_items2(COUNT)
    trajs, ks = [], []
    for i, (m, c, spokes) in enumerate(shapes):
        seed = synth.scan_seed(3, i)
        traj = ops.golden_angle2(spokes, 2 * m)
        img = torch.from_numpy(synth.coil_images2(m, c, seed)).cuda()
        k = ops.nudft2_forward_f64(traj, img, (m - 1) / 2, (m - 1) / 2).cpu().numpy()
        ks.append(torch.from_numpy(synth.add_noise(k, seed).astype(np.complex64)).cuda())
        trajs.append(traj)
        del img
    torch.cuda.synchronize()
    return shapes, trajs, ks


@pytest.fixture(scope="module", params=[2, 0], ids=["matrix", "blocks"])
def plans(ops, data, request):
    shapes, trajs, _ = data
    saved = ops.tune_get("gather_scheme")
    ops.tune_set("gather_scheme", request.param)
    try:
        ps = [ops.Plan2(m, m, t.shape[0]).set_trajectory(t) for (m, _, _), t in zip(shapes, trajs)]
    finally:
        ops.tune_set("gather_scheme", saved)
    yield ps
    for p in ps:
        p.close()


def outputs(shapes):
    rss = [torch.zeros((m, m), dtype=torch.float32, device="cuda") for m, _, _ in shapes]
    imgs = [torch.zeros((m, m, c), dtype=torch.complex64, device="cuda") for m, c, _ in shapes]
    return rss, imgs


def run_multi(ops, plans, ks, shapes, fused):
    saved = ops.tune_get("multi_fused")
    ops.tune_set("multi_fused", fused)
    try:
        rss, imgs = outputs(shapes)
        ws = torch.empty(ops.adjoint_multi_workspace_bytes(plans, [k.shape[0] for k in ks]), dtype=torch.uint8,
                         device="cuda")
        ops.adjoint_multi(plans, ks, rss=rss, images=imgs, workspace=ws)
        torch.cuda.synchronize()
    finally:
        ops.tune_set("multi_fused", saved)
    return rss, imgs


def oracle_image(i, c, k, traj, m):
    key = (i, c)
    if key not in _refs:
        t = traj.cpu().numpy().astype(np.float64)
        _refs[key] = O.kb_adjoint2(k[c].cpu().numpy(), t, O.ramp_dcf2(t), m, m)
    return _refs[key]


def test_every_shape_class_is_present():
    shapes = synth.uneven_items2(COUNT)
    assert {m for m, _, _ in shapes} == set(synth.UNEVEN_MATRICES)
    pads = {8 if c <= 8 else 16 if c <= 16 else 32 for _, c, _ in shapes}
    assert pads == {8, 16, 32}


def test_forms_match_single_plans_and_the_oracle(ops, data, plans):
    import concurrent.futures as cf
    shapes, trajs, ks = data
    chain = run_multi(ops, plans, ks, shapes, 2)
    fused = run_multi(ops, plans, ks, shapes, 1)
    lanes = run_multi(ops, plans, ks, shapes, 0)
    for i, (plan, k) in enumerate(zip(plans, ks)):
        img1, rss1 = plan.adjoint(k, rss=True)
        for rss, imgs in (chain, fused, lanes):
            assert torch.equal(imgs[i], img1) and torch.equal(rss[i], rss1), i
    assert ops.chain_faults() == 0
    torch.cuda.synchronize()
    jobs = [(i, c) for i, (_, cc, _) in enumerate(shapes) for c in (0, cc - 1)]
    with cf.ThreadPoolExecutor(max_workers=16) as ex:
        refs = list(ex.map(lambda j: oracle_image(j[0], j[1], ks[j[0]], trajs[j[0]], shapes[j[0]][0]), jobs))
    worst = 0.0
    for (i, c), ref in zip(jobs, refs):
        m, cc, spokes = shapes[i]
        err = rel_l2(fused[1][i][..., c].cpu().numpy(), ref)
        worst = max(worst, err)
        print(f"item {i:2d} matrix {m} coils {cc:2d} spokes {spokes:3d} coil {c:2d}: rel-L2 {err:.2e}")
        assert err <= TOL, (i, c)
    print(f"worst rel-L2 over {len(jobs)} coil images: {worst:.2e}")


@pytest.mark.parametrize("fused", [2, 1, 0])
def test_graph_replays_equal_the_eager_result(ops, data, plans, fused):
    shapes, _, ks = data
    saved = ops.tune_get("multi_fused")
    ops.tune_set("multi_fused", fused)
    try:
        stream = torch.cuda.Stream()
        rss, imgs = outputs(shapes)
        ws = torch.empty(ops.adjoint_multi_workspace_bytes(plans, [k.shape[0] for k in ks]), dtype=torch.uint8,
                         device="cuda")
        torch.cuda.synchronize()
        with torch.cuda.stream(stream):
            ops.adjoint_multi(plans, ks, rss=rss, images=imgs, workspace=ws)
        stream.synchronize()
        eager = [r.clone() for r in rss] + [m.clone() for m in imgs]
        graph = ops.Graph(stream)
        with graph:
            ops.adjoint_multi(plans, ks, rss=rss, images=imgs, workspace=ws)
        for replay in range(3):
            for t in rss + imgs:
                t.zero_()
            torch.cuda.synchronize()
            graph.replay()
            stream.synchronize()
            assert all(torch.equal(a, b) for a, b in zip(rss + imgs, eager)), replay
        graph.close()
    finally:
        ops.tune_set("multi_fused", saved)
