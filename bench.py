#!/usr/bin/env python3
"""Benchmark of the KSpaceJet 2-D radial reconstruction hot path on one MI355X.

Workload (the headline configuration): 2-D golden-angle radial adjoint NUFFT of a
256x256 image from 16 coils x 402 spokes x 512 samples (2x readout oversampling),
Ram-Lak ramp density compensation, Kaiser-Bessel width 4 with Beatty beta on a
512x512 grid (sigma 2), FP32 gridding/FFT, deapodize, crop, root-sum-of-squares.

A step is one burst of `--batch` scans, already resident in HBM (k-space and
trajectory), queued on the continuous-batching server (ksj_gpu_serve_*); the K
bursts are queued back to back and the server is drained once at the end, so the
timed region is the server running at saturation.  Every scan's output is its RSS
image.  `value` is served images per second over the timed steps.  Beside it the line reports:
  * latency: p50/p99 submit-to-completion latency of an open-loop Poisson stream of
    the same scans at `--load` x the measured served rate;
  * plain: the same scans through the per-scan operator (ksj_gpu_adjoint2_execute,
    one synchronous call per scan, data resident in HBM), images/s and ms per scan;
  * plain_provider: the drop-in Provider operator (gpu-recon kb_gridding_reconstruct_rss
    through ksj_provider_query) fed host-pageable payloads as the reference's host
    feeds them, one synchronous firing per scan (PCIe staging included), and over
    device edges (payloads in HBM);
  * host_fed: host-memory requests through the server (pageable and pinned caller
    buffers, PCIe staging included), the measured PCIe rates, and the like-for-like
    ratio to the host-fed Provider;
  * end_to_end: k-space in + RSS out per image at the served rate, against HBM peak;
  * uneven: 64 seeded items of mixed matrix / coil / spoke counts, plain per item vs
    one autotuned multi-plan call vs that call replayed as a captured HIP graph;
  * stages / roofline: per-kernel HIP-event timings of the batched pass (one pass
    lane, so kernels do not overlap) and the gather kernel against HBM bandwidth;
  * mixed / mixed_poisson: the reference's mixed stream (60 % radial scans of the
    headline shape, 30 % Cartesian scans of 128/256/384 matrix with 8-32 coils, 10 % 3-D
    kooshball volumes at half the 3-D spoke count), served vs plain, as a burst and
    under open-loop Poisson arrivals at several rates (p50/p99 per rate);
  * cpu_baseline: the reference's own radial route compiled from its sources and the
    FP32 restatement of this route, on every core of the affinity mask and on one;
    volume.cpu_baseline and cartesian.cpu_baseline: FP32 restatements of the 3-D route
    and of the Cartesian inverse FFT + RSS, timed the same way;
  * cartesian: the reference's CPU-runnable Cartesian case (256 x 256, 8 coils, centred
    inverse FFT per coil + RSS) through the plain GPU operator.
  * ingest: the headline scan's 402 ISMRMRD acquisitions assembled into the route's
    k-space frame in HBM (ksj_gpu_assemble_acquisitions), beside a plain device copy of
    the same bytes; ingest.cpu_baseline: the replay's copy loop in C on one core.
Every ratio compares like with like: both sides in HBM, or both fed from host memory.

    python bench.py [--steps K] [--warmup W] [--batch B] [--gpus 1]

Prints ONE JSON line.  Synthetic data: seeded 10-ellipse phantom x Gaussian coil
maps, k-space from an exact FP64 direct NUDFT on the GPU plus complex Gaussian
noise (sigma = 1e-3 max|k|), stored complex64.
"""
from __future__ import annotations

import argparse
import ctypes as C
import hashlib
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

METRIC = "2D radial recon images/s + p50/p99 scan latency, batched-served vs plain, 1 GPU"
HBM_PEAK_GBS = 8000.0
N, COILS, SPOKES, READOUT, WIDTH, SIGMA = 256, 16, 402, 512, 4, 2.0
STAGES = ("grid_prep", "grid_gather", "grid_reduce", "fft_rows", "fft_cols_deapod", "coil_combine")
TRAFFIC_PROFILE = os.path.join(HERE, "profiles", "radial256_c16_kb4_hbm_traffic.txt")
COMPULSORY_BYTES = COILS * SPOKES * READOUT * 8 + N * N * 4  # k-space in, RSS out


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_scans(ops, torch, traj, count, distinct):
    """`count` device-resident scans built from `distinct` seeded scans (copies are separate buffers)."""
    from kspacejet_amd import synth
    base = []
    for i in range(distinct):
        seed = synth.scan_seed(2, i)
        img = torch.from_numpy(synth.coil_images2(N, COILS, seed)).cuda()
        k = ops.nudft2_forward_f64(traj, img, (N - 1) / 2, (N - 1) / 2).cpu().numpy()
        base.append(synth.add_noise(k, seed).astype(np.complex64))
    scans = [torch.from_numpy(base[i % distinct]).cuda() for i in range(count)]
    return scans, base


def algorithmic_bytes(stage: str, fused_columns: bool = True) -> float:
    """Compulsory HBM bytes of one stage per image of the headline workload.  With the
    column epilogue (fused_columns) the column pass reads the coil-fastest row output and
    writes only the RSS image; without it, it writes coil planes that coil_combine reads."""
    S = SPOKES * READOUT
    G = int(N * SIGMA)
    return {
        "grid_gather": 8.0 * S * COILS + 12.0 * S + 8.0 * G * G * COILS,
        "grid_prep": 16.0 * S * COILS + 12.0 * S,
        "grid_reduce": 0.0,  # partial sums of blocks cut into several items: no compulsory bytes
        "fft_rows": 8.0 * G * G * COILS + 8.0 * G * N * COILS,
        "fft_cols_deapod": 8.0 * G * N * COILS + (4.0 * N * N if fused_columns else 8.0 * N * N * COILS),
        "coil_combine": 8.0 * N * N * COILS + 4.0 * N * N,
    }[stage]


def gather_kernel_label() -> str:
    """Template instance of the 64-coil gather kernel the current tuning launches."""
    from kspacejet_amd import ops
    if ops.tune_get("gather_scheme") == 2:
        taps = ops.tune_get("matrix_weights") == 1
        share = taps and ops.tune_get("matrix_tap_share") == 1
        pipeline = 1 if taps else ops.tune_get("matrix_pipeline")  # the packed-visit loop keeps one batch in flight
        return (f"gather_matrix_kernel<64,{ops.tune_get('matrix_batch') // 4},{pipeline},"
                f"{ops.tune_get('matrix_block_width') // 4},{str(taps).lower()},{str(share).lower()}>")
    two = ops.tune_get("gather_coils_per_lane") == 2
    return f"gather_item_kernel<{32 if two else 64},{ops.tune_get('gather_batch')},{2 if two else 1}>"


def measured_traffic(label: str, scans_per_launch: int):
    """HBM bytes per launch of the batched gather kernel, from the committed rocprofv3
    FETCH_SIZE / WRITE_SIZE passes of this workload (already scaled by the copy-kernel
    calibration), or None when that summary is absent or was taken of another kernel or
    stacking."""
    try:
        for line in open(TRAFFIC_PROFILE):
            if line.startswith("traffic_bytes_per_launch"):
                _, key, value = line.split()
                if key == f"{label}@{scans_per_launch}scans":
                    return float(value)
    except OSError:
        pass
    return None


def host_cores():
    """Cores this process may run on (its affinity mask) and the cgroup CPU quota, if any."""
    cores = len(os.sched_getaffinity(0))
    quota = None
    try:
        q, period = open("/sys/fs/cgroup/cpu.max").read().split()[:2]
        if q != "max":
            quota = round(int(q) / int(period), 2)
    except (OSError, ValueError):
        pass
    return cores, quota


def cpu_baseline(scan_np, traj_np, seconds_hint: float):
    """Two CPU routes of the headline scan, each on every core the process may run on and
    on one core: the reference's own radial route compiled from its sources (oracle/_ref:
    ramp DCF + bilinear gridding on the 256 x 256 image grid + radix-2 inverse FFT + RSS),
    and the like-for-like FP32 restatement of the GPU route (oracle/ksj_kb_route.c: KB
    W = 4 on the 512 x 512 grid, sigma 2, deapodize, crop, RSS).  Each leg times whole
    scans over about a quarter of `seconds_hint`."""
    import oracle as O
    cores, quota = host_cores()
    leg = max(1.0, seconds_hint / 4)

    def timed(run, threads):
        """images/s of `run(scans, threads) -> seconds` over about `leg` seconds."""
        probe = run(threads, threads)
        scans = threads * max(1, int(leg / max(probe, 1e-3)))
        return scans / run(scans, threads), scans

    def ref_run(scans, threads):
        t0 = time.perf_counter()
        O.ref_radial_route_rss(scan_np, traj_np, N, N, scans=scans, threads=threads)
        return time.perf_counter() - t0

    def port_run(scans, threads):
        return O.kb_route_rss_f32(scan_np, traj_np, N, width=WIDTH, scans=scans, threads=threads, timed=True)[1]

    port_all, port_scans = timed(port_run, cores)
    port_one, _ = timed(port_run, 1)
    like = {"value": round(port_all, 2), "unit": "images/s", "cores": cores, "kind": "port",
            "one_core_value": round(port_one, 2),
            "sample": f"{port_scans} scans through the FP32 restatement of the GPU route (ramp DCF, KB W={WIDTH} "
                      f"beta Beatty on the {int(N * SIGMA)}^2 grid, radix-2 inverse FFT of the rows and of the "
                      f"{N} kept columns, deapodize, crop, RSS), scan-parallel on {cores} threads; per-trajectory "
                      f"tables made once, outside the timing, as the GPU's resident plan"}
    if not O.ref_available():
        out = dict(like)
        out["cpu_quota_cores"] = quota
        return out
    ref_all, ref_scans = timed(ref_run, cores)
    ref_one, _ = timed(ref_run, 1)
    return {"value": round(ref_all, 2), "unit": "images/s", "cores": cores, "kind": "reference",
            "one_core_value": round(ref_one, 2), "cpu_quota_cores": quota,
            "sample": f"{ref_scans} scans of the 256x256x16-coil 402x512 radial shape through the reference route "
                      f"(ramp DCF + bilinear gridding on 256x256 + radix-2 inverse FFT per coil, then RSS) "
                      f"compiled from the reference's eigen_radial_gridding.cpp, scan-parallel on {cores} threads "
                      f"(every core of the affinity mask); sample-count envelope lifted",
            "like_for_like": like}


def percentile(xs, q):
    return float(np.percentile(np.asarray(xs), q)) if len(xs) else None


def ms3(x):
    """Milliseconds to three decimals; None (no samples) stays None."""
    return None if x is None else round(x, 3)


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--batch", type=int, default=32, help="scans per step")
    ap.add_argument("--distinct", type=int, default=4, help="distinct seeded scans behind the batch")
    ap.add_argument("--lanes", type=int, default=3, help="server lanes (batches in flight)")
    ap.add_argument("--max-batch", type=int, default=64, help="scans stacked into one server launch")
    ap.add_argument("--load", type=float, default=0.7, help="Poisson arrival rate as a fraction of served rate")
    ap.add_argument("--latency-requests", type=int, default=300)
    ap.add_argument("--mixed-requests", type=int, default=48, help="0 skips the mixed-stream burst section")
    ap.add_argument("--poisson-requests", type=int, default=300,
                    help="requests per arrival rate of the mixed-stream Poisson sweep; 0 skips it")
    ap.add_argument("--poisson-fractions", default="0.3,0.5,0.7,0.9",
                    help="arrival rates of the Poisson sweep as fractions of the served burst capacity")
    ap.add_argument("--host-fed-steps", type=int, default=10, help="bursts of host-memory requests; 0 skips them")
    ap.add_argument("--uneven-items", type=int, default=64, help="items of the uneven batch; 0 skips that section")
    ap.add_argument("--conditioning-reps", type=int, default=20,
                    help="timed launches per k-space conditioning kernel; 0 skips that section")
    ap.add_argument("--volume-coils", type=int, default=20,
                    help="coils of the 3-D kooshball volume (256^3, 16384 x 512 spokes); 0 skips that section")
    ap.add_argument("--volume-coils-per-pass", type=int, default=16)
    ap.add_argument("--volume-random", action="store_true",
                    help="3-D section on seeded complex Gaussian k-space instead of the simulated phantom")
    ap.add_argument("--cartesian-steps", type=int, default=10,
                    help="steps of 32 scans of the Cartesian 256x256 8-coil case (plain GPU operator and CPU "
                         "restatement); 0 skips it")
    ap.add_argument("--ingest-reps", type=int, default=50,
                    help="launches of the acquisition -> k-space frame assembly on the headline scan's 402 "
                         "acquisitions; 0 skips it")
    ap.add_argument("--provider-firings", type=int, default=40,
                    help="firings of the drop-in Provider operator (host-fed plain form); 0 skips it")
    ap.add_argument("--serve-fft-schedule", type=int, default=-1,
                    help="fft_schedule while the server runs the served and latency sections (-1, the default, "
                         "keeps the process setting; 3 takes half the threads per line at length 512).  The plain, "
                         "stage and other sections run with the process setting.")
    ap.add_argument("--tune-seconds", type=float, default=20.0, help="autotuner budget for the uneven batch")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--tune", action="append", default=[], metavar="KEY=VALUE",
                    help="launch parameter for ksj_gpu_tune_set (repeatable), e.g. gather_map=0")
    ap.add_argument("--copy-probe", action="store_true",
                    help="also run one 256 MiB device-to-device copy (calibrates FETCH_SIZE/WRITE_SIZE counters)")
    args = ap.parse_args(argv)
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1 (multi-GPU sharding is "
                                   "described in DESIGN.md but not launched)"}))
        return 2

    import torch
    from kspacejet_amd import ops, _lib
    from kspacejet_amd._lib import load

    ops.init(0)
    lib = load()
    for kv in args.tune:
        key, value = kv.split("=", 1)
        if lib.ksj_gpu_tune_set(key.encode(), int(value)) != 0:
            raise SystemExit(f"--tune {kv}: {lib.ksj_gpu_last_error().decode()}")
    traj = ops.golden_angle2(SPOKES, READOUT)
    S = SPOKES * READOUT
    traj_key = int.from_bytes(hashlib.sha256(traj.cpu().numpy().tobytes()).digest()[:8], "little") | 1
    log("generating synthetic scans ...")
    scans, base = make_scans(ops, torch, traj, args.batch, min(args.distinct, args.batch))
    scan0 = base[0]
    outs = [torch.empty((N, N), dtype=torch.float32, device="cuda") for _ in range(args.batch)]
    stream = lambda: torch.cuda.current_stream().cuda_stream

    # ---------------------------------------------------------------- served
    server = ops.Server(lanes=args.lanes, max_batch_scans=args.max_batch)
    Req = _lib.ServeRequest * args.batch
    reqs = Req(*[_lib.ServeRequest(C.sizeof(_lib.ServeRequest), _lib.SERVE_RADIAL2D, N, N, COILS, S, WIDTH,
                                   _lib.DENSITY_RAMP, 0, 0, traj_key, traj.data_ptr(), k.data_ptr(), None,
                                   o.data_ptr()) for k, o in zip(scans, outs)])
    tickets = (C.c_uint64 * args.batch)()

    def served_step():
        """Queues one burst; the server works on it while the next burst is queued."""
        rc = lib.ksj_gpu_serve_submit_batch(server._handle, args.batch, reqs, tickets)
        if rc != 0:
            raise RuntimeError(lib.ksj_gpu_last_error().decode())
        return list(tickets)

    fft_schedule_saved = ops.tune_get("fft_schedule")
    if args.serve_fft_schedule >= 0:
        ops.tune_set("fft_schedule", args.serve_fft_schedule)
    log("served bursts ...")
    for _ in range(args.warmup):
        for t in served_step():
            server.wait(t)
    torch.cuda.synchronize()
    step_tickets = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step_tickets += served_step()
    server.drain()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    burst_lat = [server.wait(t) for t in step_tickets]
    images = args.batch * args.steps
    value = images / elapsed

    # Open-loop Poisson stream at a fraction of the served rate: per-scan latency.
    log("open-loop latency stream ...")
    rate = args.load * value
    rng = np.random.default_rng(7)
    arrivals = np.cumsum(rng.exponential(1.0 / rate, args.latency_requests))
    one = _lib.ServeRequest * 1
    stream_tickets = []
    t_start = time.perf_counter()
    for i, t_arr in enumerate(arrivals):
        while time.perf_counter() - t_start < t_arr:
            pass
        r = one(reqs[i % args.batch])
        t = C.c_uint64(0)
        if lib.ksj_gpu_serve_submit_batch(server._handle, 1, r, C.byref(t)) != 0:
            raise RuntimeError(lib.ksj_gpu_last_error().decode())
        stream_tickets.append(t.value)
    stream_lat = [server.wait(t) for t in stream_tickets]
    served_stats = server.stats()

    ops.tune_set("fft_schedule", fft_schedule_saved)

    # ---------------------------------------------------------------- plain
    log("plain per-scan operator ...")
    plan = ops.Plan2(N, N, S, oversampling=SIGMA, width=WIDTH).set_trajectory(traj)
    ws = torch.empty(plan.workspace_bytes(COILS), dtype=torch.uint8, device="cuda")

    def plain_step():
        for k, out in zip(scans, outs):
            lib.ksj_gpu_adjoint2_execute(plan._handle, k.data_ptr(), COILS, None, 0, out.data_ptr(),
                                         ws.data_ptr(), ws.numel(), stream())
            torch.cuda.synchronize()  # one synchronous call per scan, as a per-firing operator runs

    for _ in range(args.warmup):
        plain_step()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        plain_step()
    plain_value = images / (time.perf_counter() - t0)

    # ---------------------------------------------------------------- stages and roofline
    P = C.c_void_p * args.batch
    k_ptrs = P(*[k.data_ptr() for k in scans])
    r_ptrs = P(*[o.data_ptr() for o in outs])
    lanes_saved = ops.tune_get("pass_lanes")
    ops.tune_set("pass_lanes", 1)  # kernels back to back: per-kernel durations without overlap
    ws_batch = torch.empty(plan.batch_workspace_bytes(COILS, args.batch), dtype=torch.uint8, device="cuda")

    def run_batched():
        rc = lib.ksj_gpu_adjoint2_execute_batch(plan._handle, args.batch, k_ptrs, COILS, None, 0, r_ptrs,
                                                ws_batch.data_ptr(), ws_batch.numel(), stream())
        if rc != 0:
            raise RuntimeError(lib.ksj_gpu_last_error().decode())

    run_batched()
    lib.ksj_gpu_profile_enable(1)
    lib.ksj_gpu_profile_reset()
    torch.cuda.synchronize()
    for _ in range(max(1, args.steps // 2)):
        run_batched()
    torch.cuda.synchronize()
    stages = {}
    for name in STAGES:
        tot, cnt = C.c_double(0), C.c_uint64(0)
        lib.ksj_gpu_profile_read(name.encode(), C.byref(tot), C.byref(cnt))
        stages[name] = {"avg_us": 1e3 * tot.value / max(cnt.value, 1), "launches": cnt.value}
    lib.ksj_gpu_profile_enable(0)
    ops.tune_set("pass_lanes", lanes_saved)
    del ws_batch
    per_pass = max(1, 64 // COILS)
    scans_per_launch = min(per_pass, args.batch)
    g = stages["grid_gather"]
    per_launch_bytes = algorithmic_bytes("grid_gather") * scans_per_launch
    achieved = per_launch_bytes / (g["avg_us"] * 1e-6) / 1e9
    label = gather_kernel_label()
    roofline = {"kernel": f"{label} (grid_gather)", "bound": "hbm", "achieved": round(achieved, 1),
                "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                "traffic": measured_traffic(label, scans_per_launch), "scans_per_launch": scans_per_launch,
                "algorithmic_bytes_per_launch": per_launch_bytes, "avg_us": round(g["avg_us"], 2)}
    fused_columns = stages["coil_combine"]["launches"] == 0
    stage_report = {k: {"avg_us": round(v["avg_us"], 2),
                        "gbs": round(algorithmic_bytes(k, fused_columns) * scans_per_launch
                                     / max(v["avg_us"] * 1e-6, 1e-12) / 1e9, 1)}
                    for k, v in stages.items() if v["launches"]}
    pipeline_bytes = sum(algorithmic_bytes(k, fused_columns) for k, v in stages.items() if v["launches"])

    # ---------------------------------------------------------------- mixed stream
    mixed = None
    if args.mixed_requests > 0:
        log("mixed stream ...")
        mixed = {"with_volumes": run_mixed(torch, lib, _lib, server, plan, scans, traj, traj_key, args.mixed_requests,
                                           stream, 0.1),
                 "2d_only": run_mixed(torch, lib, _lib, server, plan, scans, traj, traj_key, args.mixed_requests,
                                      stream, 0.0)}
    poisson = None
    if args.poisson_requests > 0:
        log("mixed stream under Poisson arrivals ...")
        poisson = run_mixed_poisson(torch, lib, _lib, server, plan, scans, traj, traj_key, args.poisson_requests,
                                    stream, [float(x) for x in args.poisson_fractions.split(",")])
    server.close()
    host_fed = None
    if args.host_fed_steps > 0:
        log("host-memory requests ...")
        host_fed = run_host_fed(torch, ops, lib, _lib, base, traj, traj_key, args.batch, args.host_fed_steps,
                                args.warmup, args.lanes, args.max_batch)
    del scans, outs, ws
    provider = None
    if args.provider_firings > 0:
        log("drop-in Provider operator ...")
        provider = run_provider_plain(scan0, traj.cpu().numpy(), args.provider_firings, args.warmup)
        if host_fed is not None:  # like for like: both fed from host-pageable memory
            host_fed["plain_provider_host_images_per_s"] = provider["images_per_s"]
            host_fed["served_over_plain_provider_host"] = round(
                host_fed["forms"]["pageable"]["images_per_s"] / provider["images_per_s"], 2)
    volume = None
    if args.volume_coils > 0:
        log("3-D kooshball volume ...")
        volume = run_volume(torch, ops, lib, args.volume_coils, args.volume_coils_per_pass, max(1, args.steps // 4),
                            1, simulate=not args.volume_random, cpu=not args.no_cpu_baseline)
    cartesian = None
    if args.cartesian_steps > 0:
        log("Cartesian 256x256 8-coil case ...")
        cartesian = run_cartesian(torch, ops, lib, args.cartesian_steps, cpu=not args.no_cpu_baseline)
    ingest = None
    if args.ingest_reps > 0:
        log("acquisition -> k-space frame assembly ...")
        ingest = run_ingest(torch, ops, lib, args.ingest_reps, cpu=not args.no_cpu_baseline)
    conditioning = None
    if args.conditioning_reps > 0:
        log("k-space conditioning kernels ...")
        conditioning = run_conditioning(torch, ops, args.conditioning_reps)
    uneven = None
    if args.uneven_items > 0:
        log("uneven batch ...")
        uneven = run_uneven(torch, ops, lib, args.uneven_items, args.steps, args.warmup, args.tune_seconds)

    if args.copy_probe:
        src = torch.empty(256 << 20, dtype=torch.uint8, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    cpu = None
    if not args.no_cpu_baseline:
        log("timing the CPU reference route ...")
        cpu = cpu_baseline(scan0, traj.cpu().numpy(), args.cpu_seconds)
    line = {
        "metric": METRIC, "value": round(value, 1), "unit": "images/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(1e3 * elapsed / args.steps, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic: seeded 10-ellipse phantom x 16 Gaussian coils, exact FP64 NUDFT k-space + noise, complex64",
        "config": {"workload": "2-D golden-angle radial adjoint NUFFT + RSS, served: bursts of device-resident scans "
                               "through the continuous-batching server (same-trajectory scans stacked along the coil "
                               "axis, resident plan)",
                   "image": N, "coils": COILS, "spokes": SPOKES, "readout": READOUT, "kernel": "kaiser_bessel",
                   "kernel_width": WIDTH, "oversampling": SIGMA, "grid": int(N * SIGMA), "batch": args.batch,
                   "server_lanes": args.lanes, "server_max_batch": args.max_batch,
                   "server_fft_schedule": args.serve_fft_schedule if args.serve_fft_schedule >= 0
                   else fft_schedule_saved},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "latency": {"p50_ms": ms3(percentile(stream_lat, 50)), "p99_ms": ms3(percentile(stream_lat, 99)),
                    "arrival_rate_per_s": round(rate, 1), "requests": len(stream_lat),
                    "process": "open-loop Poisson arrivals, one scan per request",
                    "saturated_p50_ms": ms3(percentile(burst_lat, 50)),
                    "saturated_p99_ms": ms3(percentile(burst_lat, 99))},
        "plain": {"images_per_s": round(plain_value, 1), "ms_per_scan": round(1e3 / plain_value, 3),
                  "operator": "ksj_gpu_adjoint2_execute, one synchronous call per scan, plan resident"},
        "served_over_plain": round(value / plain_value, 2),
        "end_to_end": {"compulsory_bytes_per_image": COMPULSORY_BYTES,
                       "pipeline_bytes_per_image": round(pipeline_bytes),
                       "achieved_GBps": round(value * COMPULSORY_BYTES / 1e9, 1),
                       "frac": round(value * COMPULSORY_BYTES / 1e9 / HBM_PEAK_GBS, 4),
                       "note": "k-space in + RSS out per image over the served rate, against 8 TB/s"},
        "plain_provider": provider,
        "served_over_plain_provider_device": round(
            value / provider["device_edges"]["device_trajectory"]["images_per_s"], 2) if provider else None,
        "host_fed": host_fed,
        "stages": stage_report,
        "served_stats": served_stats,
        "mixed": mixed,
        "mixed_poisson": poisson,
        "uneven": uneven,
        "volume": volume,
        "cartesian": cartesian,
        "conditioning": conditioning,
        "ingest": ingest,
    }
    print(json.dumps(line))
    return 0


def run_conditioning(torch, ops, reps):
    """The k-space conditioning kernels on device-resident data, timed as `reps` launches
    replayed from one captured HIP graph (the per-launch average): noise
    prewhitening and coil compression (channel mix) of the headline scan's non-Cartesian
    k-space (205,824 samples; 16x16, 32->16 and 64x64 matrices), and per-line phase
    correction and 2x readout-oversampling removal of a 16-channel 256 x 512 Cartesian
    frame.  Algorithmic bytes: every input value read once and every output value written
    once; flops of the mix: 8 per complex multiply-add."""
    g = torch.Generator(device="cuda").manual_seed(11)

    def crandn(*shape):
        return torch.complex(torch.randn(shape, generator=g, device="cuda"), torch.randn(shape, generator=g, device="cuda"))

    def timed(fn):
        # `reps` launches captured back to back in one HIP graph (no host gaps between
        # them), the replay bracketed by one event pair: the average launch time.  The
        # kernels record no memsets, so they are safe to capture.
        fn()
        torch.cuda.synchronize()
        graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(graph):
            for _ in range(reps):
                fn()
        graph.replay()
        torch.cuda.synchronize()
        start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        start.record()
        graph.replay()
        stop.record()
        torch.cuda.synchronize()
        del graph
        return start.elapsed_time(stop) * 1e3 / reps  # us per launch

    res = {}
    S = SPOKES * READOUT
    for name, out_ch, in_ch in (("prewhiten_16", 16, 16), ("compress_32_to_16", 16, 32), ("prewhiten_64", 64, 64)):
        x, w = crandn(in_ch, S), crandn(out_ch, in_ch)
        us = timed(lambda: ops.channel_mix(w, x))
        nbytes = (in_ch + out_ch) * S * 8 + out_ch * in_ch * 8
        if ops.tune_get("mix_scheme") == 1:
            rb = 2 if out_ch <= 16 else 4 if out_ch <= 32 else 8
            pairs = 4 if in_ch <= 16 else 8 if in_ch <= 32 else 16
            kernel = f"channel_mix_mfma_kernel<{rb}, {pairs}>"
        else:
            kernel = f"channel_mix_kernel<{4 if in_ch <= 4 else 8 if in_ch <= 8 else 16 if in_ch <= 16 else 32 if in_ch <= 32 else 64}>"
        res[name] = {"kernel": kernel, "samples": S, "us": round(us, 2),
                     "gbs": round(nbytes / (us * 1e-6) / 1e9, 1), "frac_hbm": round(nbytes / (us * 1e-6) / 8e12, 3),
                     "tflops": round(8.0 * out_ch * in_ch * S / (us * 1e-6) / 1e12, 2)}
        del x, w
    nc, nr, nk = 16, 256, 512
    x = crandn(nc, nr, nk)
    p = crandn(nc, nk)
    us = timed(lambda: ops.phase_correct(x, p))
    nbytes = 2 * nc * nr * nk * 8 + nc * nk * 8
    res["phase_correct"] = {"kernel": "phase_correct_kernel", "frame": [nc, nr, nk], "us": round(us, 2),
                            "gbs": round(nbytes / (us * 1e-6) / 1e9, 1),
                            "frac_hbm": round(nbytes / (us * 1e-6) / 8e12, 3)}
    us = timed(lambda: ops.readout_crop(x, nk // 2, nk // 4))
    nbytes = 2 * nc * nr * (nk // 2) * 8
    res["readout_crop"] = {"kernel": "readout_crop_kernel", "frame": [nc, nr, nk], "out_cols": nk // 2, "us": round(us, 2),
                           "gbs": round(nbytes / (us * 1e-6) / 1e9, 1),
                           "frac_hbm": round(nbytes / (us * 1e-6) / 8e12, 3)}
    return res


def run_volume(torch, ops, lib, coils, coils_per_pass, steps, warmup, simulate=True, cpu=True):
    """3-D radial (kooshball) adjoint of one 256^3 volume: 16384 golden-means spokes x
    512 samples, KB width 4 on a 512^3 grid (sigma 2), |k|^2 density, coils in passes
    of `coils_per_pass`, RSS volume out.  The k-space is simulated as the workload
    defines it: the 10-ellipsoid phantom times 3-D Gaussian coil sensitivities (sigma 90
    voxels, centres on a sphere of radius 140), through the FP64 Kaiser-Bessel forward
    NUFFT of width 6 on the 2x grid, plus complex Gaussian noise of sigma 1e-3 max|k|,
    stored complex64 (`simulate` False: seeded complex Gaussian data of the same shape)."""
    n, spokes, readout = 256, 16384, 512
    traj = ops.golden_means3(spokes, readout)
    S = traj.shape[0]
    t_sim = time.perf_counter()
    if simulate:
        from kspacejet_amd import synth
        k = synth.kooshball_kspace3(ops, torch, n, coils, traj, synth.scan_seed(4), width=6)
    else:
        gen = torch.Generator(device="cuda").manual_seed(0x4B53504A ^ (4 << 32))
        k = torch.randn((coils, S), dtype=torch.complex64, device="cuda", generator=gen)
    torch.cuda.synchronize()
    sim_s = time.perf_counter() - t_sim
    log(f"  3-D k-space {'simulated (FP64 KB W=6 forward NUFFT)' if simulate else 'random'} in {sim_s:.1f} s")
    t0 = time.perf_counter()
    plan = ops.Plan3((n, n, n), S, width=WIDTH).set_trajectory(traj)
    torch.cuda.synchronize()
    build_ms = 1e3 * (time.perf_counter() - t0)
    info = plan.info()
    ws = torch.empty(plan.workspace_bytes(coils, coils_per_pass), dtype=torch.uint8, device="cuda")
    rss = torch.empty((n, n, n), dtype=torch.float32, device="cuda")
    stream = torch.cuda.current_stream().cuda_stream

    def run():
        rc = lib.ksj_gpu_adjoint3_execute(plan._handle, k.data_ptr(), coils, coils_per_pass, None, rss.data_ptr(),
                                          ws.data_ptr(), ws.numel(), stream)
        if rc != 0:
            raise RuntimeError(lib.ksj_gpu_last_error().decode())

    for _ in range(warmup):
        run()
    torch.cuda.synchronize()
    lib.ksj_gpu_profile_enable(1)
    lib.ksj_gpu_profile_reset()
    t0 = time.perf_counter()
    for _ in range(steps):
        run()
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / steps
    stages = {}
    for name in ("grid3_prep", "grid3_gather", "grid3_reduce", "fft3_x", "fft3_y_deapod", "fft3_z_deapod",
                 "coil_combine3"):
        tot, cnt = C.c_double(0), C.c_uint64(0)
        lib.ksj_gpu_profile_read(name.encode(), C.byref(tot), C.byref(cnt))
        if cnt.value:
            stages[name] = round(tot.value / steps, 3)  # ms per volume (the profile totals are in ms)
    lib.ksj_gpu_profile_enable(0)
    G = 2 * n
    passes = -(-coils // coils_per_pass)
    # Algorithmic bytes of the gather per volume: samples in, sorted coordinates and
    # weights once per pass, every grid cell of every coil written once.
    gather_bytes = 8 * S * coils + 12 * S * passes + 8 * G ** 3 * coils
    gather_ms = stages.get("grid3_gather")
    res = {"volumes_per_s": round(1.0 / dt, 3), "ms_per_volume": round(1e3 * dt, 2), "plan_build_ms": round(build_ms, 1),
           "shape": [n, n, n], "grid": [G, G, G], "coils": coils, "coils_per_pass": coils_per_pass,
           "spokes": spokes, "readout": readout, "samples": S, "kernel_width": WIDTH, "density": "|k|^2",
           "plan_items": info["items"], "plan_visits": info["visits"], "plan_device_bytes": info["device_bytes"],
           "workspace_bytes": ws.numel(), "stages_ms_per_volume": stages,
           "gather_gbs": round(gather_bytes / (gather_ms * 1e-3) / 1e9, 1) if gather_ms else None,
           "data": ("10-ellipsoid phantom x 20 3-D Gaussian coil sensitivities, FP64 Kaiser-Bessel W=6 forward "
                    "NUFFT on the 2x grid, complex Gaussian noise 1e-3 max|k|, stored complex64"
                    if simulate else "seeded complex Gaussian k-space of the workload's shape"),
           "simulation_s": round(sim_s, 2)}
    plan.close()
    del ws, rss
    if cpu:
        res["cpu_baseline"] = volume_cpu_baseline(k, traj, n, coils)
    del k
    return res


def volume_cpu_baseline(k, traj, n, coils):
    """The FP32 CPU restatement of the 3-D route (oracle/ksj_kb_route.c: |k|^2 DCF, KB W=4,
    beta Beatty, on the (2n)^3 grid, inverse FFT pruned axis by axis, deapodization, RSS)
    on a bounded sample of the same volume: one coil per core of the affinity mask at once
    (each core its own 1 GiB grid), then one coil on one core; volumes/s = coils/s divided
    by the volume's coil count.  Per-trajectory tables are made outside the timing, as the
    GPU's resident plan."""
    import oracle as O
    cores, _ = host_cores()
    t = min(cores, coils)
    traj_np = traj.cpu().numpy().astype(np.float32)
    k_np = k[:t].cpu().numpy()
    _, secs = O.kb_route3_rss_f32(k_np, traj_np, n, width=WIDTH, threads=t, timed=True)
    _, one = O.kb_route3_rss_f32(k_np[:1], traj_np, n, width=WIDTH, threads=1, timed=True)
    return {"value": round(t / secs / coils, 4), "unit": "volumes/s", "cores": t, "kind": "port",
            "one_core_value": round(1.0 / one / coils, 5),
            "sample": f"{t} coils of the {n}^3 {coils}-coil volume at once on {t} threads ({secs:.1f} s), and one "
                      f"coil on one thread ({one:.1f} s), through the FP32 restatement of the GPU's 3-D route; "
                      f"volumes/s = coils/s / {coils}"}


def run_cartesian(torch, ops, lib, steps, cpu=True):
    """The reference's CPU-runnable Cartesian case: 256 x 256, 8 coils, k-space the FP64
    centred FFT of the 10-ellipse phantom times 2-D Gaussian coils plus complex Gaussian
    noise 1e-3 max|k|, stored complex64; recon = centred inverse 2-D FFT per coil, then
    RSS.  Plain per-scan GPU operator (ksj_gpu_ifft2_coil_images + ksj_gpu_rss, one
    synchronous call pair per scan, data in HBM) and, on the host, the FP32 CPU
    restatement (oracle/ksj_kb_route.c) scan-parallel on every core and on one core."""
    from kspacejet_amd import synth
    n, coils, batch = 256, 8, 32
    k_np = synth.cartesian_kspace2(n, coils, synth.scan_seed(1), centred=True).astype(np.complex64)
    k = torch.from_numpy(k_np).cuda()
    img = torch.empty((n, n, coils), dtype=torch.complex64, device="cuda")
    rss = torch.empty((n, n), dtype=torch.float32, device="cuda")
    ws = torch.empty(int(lib.ksj_gpu_ifft2_workspace_bytes(coils, n, n)), dtype=torch.uint8, device="cuda")
    stream = torch.cuda.current_stream().cuda_stream

    def one():
        lib.ksj_gpu_ifft2_coil_images(k.data_ptr(), img.data_ptr(), coils, n, n, 1, 0, ws.data_ptr(), ws.numel(),
                                      stream)
        lib.ksj_gpu_rss(img.data_ptr(), rss.data_ptr(), n * n, coils, stream)
        torch.cuda.synchronize()

    for _ in range(5):
        one()
    scans = max(1, steps) * batch
    t0 = time.perf_counter()
    for _ in range(scans):
        one()
    gpu = scans / (time.perf_counter() - t0)
    res = {"images_per_s": round(gpu, 1), "shape": [n, n], "coils": coils, "scans": scans,
           "operator": "ksj_gpu_ifft2_coil_images (centred, coil-last) + ksj_gpu_rss, one synchronous call pair per "
                       "scan, data in HBM",
           "data": "10-ellipse phantom x 8 2-D Gaussian coils, FP64 centred FFT, complex Gaussian noise 1e-3 max|k|, "
                   "stored complex64"}
    if cpu:
        import oracle as O
        cores, _ = host_cores()
        probe = O.cart_route_rss_f32(k_np, centred=True, scans=cores, threads=cores, timed=True)[1]
        reps = cores * max(1, int(3.0 / max(probe, 1e-3)))
        secs = O.cart_route_rss_f32(k_np, centred=True, scans=reps, threads=cores, timed=True)[1]
        reps1 = max(1, int(2.0 / max(probe, 1e-3)))
        secs1 = O.cart_route_rss_f32(k_np, centred=True, scans=reps1, threads=1, timed=True)[1]
        res["cpu_baseline"] = {"value": round(reps / secs, 1), "unit": "images/s", "cores": cores, "kind": "port",
                               "one_core_value": round(reps1 / secs1, 1),
                               "sample": f"{reps} scans on {cores} threads and {reps1} on one, through the FP32 "
                                         f"restatement (radix-2 inverse FFT per coil with Normalization::inverse, "
                                         f"centred, RSS); the reference's Eigen FFT is not in the image"}
    del k, img, rss, ws
    return res


def run_ingest(torch, ops, lib, reps, cpu=True):
    """ISMRMRD acquisitions -> the route's k-space frame (ksj_gpu_assemble_acquisitions) for the
    headline scan: 402 spokes, one acquisition each, 16 channels x 512 samples, packed in HBM
    in the ISMRMRD layout.  The uniform form (every spoke 512 samples) is timed as `reps`
    launches replayed from one captured HIP graph; the ragged form (random 384-640 samples per
    spoke) per call on the host clock, its offset upload and stream synchronisation included.
    Algorithmic bytes per frame: the frame read once and written once, 2 x 8 x S x C.
    cpu_baseline: the replay's copy loop in C (one memcpy per acquisition and channel, as
    the route's replay) on one core."""
    S = SPOKES * READOUT
    g = torch.Generator(device="cuda")
    g.manual_seed(5)
    packed = torch.complex(torch.randn(COILS * S, generator=g, device="cuda"),
                           torch.randn(COILS * S, generator=g, device="cuda"))
    frame = torch.empty((COILS, S), dtype=torch.complex64, device="cuda")
    ws = torch.empty(int(lib.ksj_gpu_assemble_workspace_bytes(SPOKES)), dtype=torch.uint8, device="cuda")
    counts = (C.c_uint32 * SPOKES)(*([READOUT] * SPOKES))
    stream = torch.cuda.current_stream().cuda_stream

    def uniform():
        rc = lib.ksj_gpu_assemble_acquisitions(packed.data_ptr(), counts, SPOKES, COILS, frame.data_ptr(),
                                               ws.data_ptr(), ws.numel(), torch.cuda.current_stream().cuda_stream)
        assert rc == 0

    uniform()
    torch.cuda.synchronize()
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph):
        for _ in range(reps):
            uniform()
    graph.replay()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    graph.replay()
    e1.record()
    torch.cuda.synchronize()
    del graph
    us = e0.elapsed_time(e1) * 1e3 / reps
    nbytes = 2 * 8 * S * COILS
    gbs = nbytes / (us * 1e-6) / 1e9
    # The same bytes through a plain device-to-device copy (torch copy_), for scale.
    flat = frame.view(-1)
    flat.copy_(packed)
    torch.cuda.synchronize()
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph):
        for _ in range(reps):
            flat.copy_(packed)
    graph.replay()
    torch.cuda.synchronize()
    e0.record()
    graph.replay()
    e1.record()
    torch.cuda.synchronize()
    del graph
    copy_us = e0.elapsed_time(e1) * 1e3 / reps
    rng = np.random.default_rng(9)
    ragged_counts = rng.integers(384, 641, SPOKES)
    rtotal = int(ragged_counts.sum())
    rpacked = torch.complex(torch.randn(COILS * rtotal, generator=g, device="cuda"),
                            torch.randn(COILS * rtotal, generator=g, device="cuda"))
    rframe = torch.empty((COILS, rtotal), dtype=torch.complex64, device="cuda")
    rc_counts = (C.c_uint32 * SPOKES)(*[int(x) for x in ragged_counts])

    def ragged():
        rc = lib.ksj_gpu_assemble_acquisitions(rpacked.data_ptr(), rc_counts, SPOKES, COILS, rframe.data_ptr(),
                                               ws.data_ptr(), ws.numel(), stream)
        assert rc == 0

    ragged()
    t0 = time.perf_counter()
    for _ in range(reps):
        ragged()
    ragged_us = (time.perf_counter() - t0) * 1e6 / reps
    res = {"us_per_frame": round(us, 2), "GBps": round(gbs, 1), "frac": round(gbs / HBM_PEAK_GBS, 4),
           "plain_copy_us": round(copy_us, 2), "plain_copy_GBps": round(nbytes / (copy_us * 1e-6) / 1e9, 1),
           "bytes_per_frame": nbytes, "ragged_us_per_call": round(ragged_us, 2), "ragged_samples": rtotal,
           "shape": {"acquisitions": SPOKES, "channels": COILS, "samples_per_acquisition": READOUT},
           "operator": "ksj_gpu_assemble_acquisitions: ISMRMRD acquisitions [C][ns] packed in HBM -> frame [C][S]"}
    if cpu:
        import oracle as O
        host = packed.cpu().numpy()
        host_counts = [READOUT] * SPOKES
        probe = O.assemble_acquisitions_c(host, host_counts, COILS, frames=5)[1] / 5
        frames = max(5, int(3.0 / max(probe, 1e-5)))
        secs = O.assemble_acquisitions_c(host, host_counts, COILS, frames=frames)[1]
        res["cpu_baseline"] = {"value": round(frames / secs, 1), "unit": "frames/s", "cores": 1, "kind": "port",
                               "sample": f"{frames} frames through the replay's copy loop in C (oracle/ksj_kb_route.c: "
                                         f"one memcpy per acquisition and channel, as the reference's replay) on one "
                                         f"core"}
        res["gpu_frames_per_s"] = round(1e6 / us, 1)
    del packed, frame, ws, rpacked, rframe
    return res


def run_provider_plain(scan_np, traj_np, firings, warmup):
    """The drop-in Provider operator (module gpu-recon, kb_gridding_reconstruct_rss,
    reached through ksj_provider_query): one synchronous firing per scan.  Three edge
    forms of the same firing, same bits out:
      * host: host-pageable, 64-byte aligned k-space and trajectory payloads and a host
        output grant, as the reference's host supplies them; the operator stages them
        through its pinned bounce buffers and copies the RSS image back;
      * device: k-space payload and output grant in HBM (KSJ_PROVIDER_MEMORY_DEVICE,
        an upstream GPU operator's output and a downstream one's input), trajectory on
        the host, compared bytewise with the resident plan's on the host;
      * device_trajectory: the trajectory in HBM as well, compared with the resident
        plan's and value-checked on the GPU; nothing crosses PCIe but a flag word.
    The resident plan is reused in every form (same trajectory bytes)."""
    import torch
    from kspacejet_amd import provider_host as H
    S = traj_np.shape[0]
    cfg = (f'{{"channels":{COILS},"density_compensation":"radial_analytic_ramp","image_cols":{N},"image_rows":{N},'
           f'"kernel_width":{WIDTH},"oversampling":2,"sample_count":{S},"trajectory_units":"radians_per_pixel"}}')
    module = H.ProviderModule("gpu_recon")
    st, op, msg = module.create("kb_gridding_reconstruct_rss", cfg)
    if st != H.OK or op.open(0) != H.OK:
        raise RuntimeError(f"provider operator: {msg}")
    k_np = np.ascontiguousarray(scan_np, np.complex64)
    t_np = np.ascontiguousarray(traj_np, np.float32)

    def lease(k_dev, t_dev, o_dev):
        def item(a, type_ref, dev):
            if dev:
                return H.Item(torch.from_numpy(a).cuda(), type_ref, memory_domain=H.MEMORY_DEVICE)
            return H.Item(H.aligned_copy(a), type_ref)
        ls = H.Lease([H.Batch(0, [item(k_np, "ksj.noncartesian-kspace-frame", k_dev)]),
                      H.Batch(1, [item(t_np, "ksj.trajectory-frame", t_dev)])],
                     H.OutputSpec(N * N * 4, "ksj.image-frame", np.float32,
                                  memory_domain=H.MEMORY_DEVICE if o_dev else H.MEMORY_HOST_PAGEABLE))
        torch.cuda.synchronize()  # device payloads complete before the operator's stream reads them
        return ls

    def timed(ls):
        def fire():
            st, r, msg = op.process(ls)
            if st != H.OK or r.sealed_output_count != 1:
                raise RuntimeError(f"provider firing failed: {msg}")
        for _ in range(max(1, warmup)):
            fire()
        t0 = time.perf_counter()
        for _ in range(firings):
            fire()
        dt = (time.perf_counter() - t0) / firings
        out = ls.output.cpu().numpy() if hasattr(ls.output, "data_ptr") else ls.output
        return dt, hashlib.sha256(np.ascontiguousarray(out).tobytes()).hexdigest()

    host = lease(False, False, False)
    dt, digest = timed(host)
    forms = {}
    for name, dev in (("device", (True, False, True)), ("device_trajectory", (True, True, True))):
        ddt, ddigest = timed(lease(*dev))
        forms[name] = {"images_per_s": round(1.0 / ddt, 1), "ms_per_firing": round(1e3 * ddt, 3),
                       "same_bits_as_host": ddigest == digest}
    op.close()
    forms["chain"] = run_provider_chain(module, cfg, k_np, t_np, firings, warmup)
    return {"images_per_s": round(1.0 / dt, 1), "ms_per_firing": round(1e3 * dt, 3), "firings": firings,
            "operator": "org.kspacejet.gpu-recon kb_gridding_reconstruct_rss via ksj_provider_query, host-pageable "
                        "payloads staged by the operator, one synchronous firing per scan",
            "host_bytes_per_firing": int(host.batches[0].items[0].data.nbytes + t_np.nbytes + N * N * 4),
            "device_edges": forms}


def run_provider_chain(module, cfg, k_np, t_np, firings, warmup):
    """Two operators in a row, the coil-image edge between them the one the device domain
    removes from PCIe: kb_gridding_reconstruct (coil images, 256 x 256 x 16 complex64 =
    8.4 MB) then coil_combine_sensitivity with seeded Gaussian maps.  Host edges: every
    payload host-pageable, so the coil images go D2H out of the first operator and H2D
    into the second.  Device edges: k-space, coil images, maps and output in HBM (the
    trajectory stays on the host).  Same bits out."""
    import torch
    from kspacejet_amd import provider_host as H
    from kspacejet_amd import synth
    st, rec, msg = module.create("kb_gridding_reconstruct", cfg)
    if st != H.OK or rec.open(0) != H.OK:
        raise RuntimeError(f"provider operator: {msg}")
    st, comb, msg = module.create("coil_combine_sensitivity", f'{{"channels":{COILS},"cols":{N},"rows":{N}}}')
    if st != H.OK or comb.open(0) != H.OK:
        raise RuntimeError(f"provider operator: {msg}")
    maps = np.ascontiguousarray(synth.coil_maps2(N, COILS, synth.scan_seed(1, 0)).transpose(1, 2, 0), np.complex64)
    coil_bytes = N * N * COILS * 8

    def item(a, type_ref, dev, keys=(0, 0)):
        if dev:
            a = a if hasattr(a, "data_ptr") else torch.from_numpy(a).cuda()
            return H.Item(a, type_ref, *keys, memory_domain=H.MEMORY_DEVICE)
        return H.Item(a if a.ctypes.data % 64 == 0 else H.aligned_copy(a), type_ref, *keys)

    def chain(dev):
        dom = H.MEMORY_DEVICE if dev else H.MEMORY_HOST_PAGEABLE
        first = H.Lease([H.Batch(0, [item(k_np, "ksj.noncartesian-kspace-frame", dev)]),
                         H.Batch(1, [item(t_np, "ksj.trajectory-frame", False)])],
                        H.OutputSpec(coil_bytes, "ksj.coil-image-frame", np.complex64, memory_domain=dom))
        imgs = first.output if dev else first.output.view(np.complex64)
        second = H.Lease([H.Batch(0, [item(imgs, "ksj.coil-image-frame", dev, (5, 6))]),
                          H.Batch(1, [item(maps, "ksj.sensitivity-map", dev, (1, 2))])],
                         H.OutputSpec(N * N * 8, "ksj.complex-image-frame", np.complex64, memory_domain=dom))
        torch.cuda.synchronize()

        def fire():
            for op, ls in ((rec, first), (comb, second)):
                st, r, msg = op.process(ls)
                if st != H.OK or r.sealed_output_count != 1:
                    raise RuntimeError(f"provider firing failed: {msg}")
        for _ in range(max(1, warmup)):
            fire()
        t0 = time.perf_counter()
        for _ in range(firings):
            fire()
        dt = (time.perf_counter() - t0) / firings
        out = second.output.cpu().numpy() if dev else second.output
        return dt, hashlib.sha256(np.ascontiguousarray(out).tobytes()).hexdigest()

    host_dt, host_digest = chain(False)
    dev_dt, dev_digest = chain(True)
    rec.close()
    comb.close()
    return {"operators": "kb_gridding_reconstruct -> coil_combine_sensitivity",
            "host_ms_per_image": round(1e3 * host_dt, 3), "device_ms_per_image": round(1e3 * dev_dt, 3),
            "coil_image_bytes_off_pcie": 2 * coil_bytes, "same_bits": host_digest == dev_digest}


def run_uneven(torch, ops, lib, count, steps, warmup, tune_seconds):
    """Uneven radial batch: `count` items drawn from seeds (matrix 128/192/256/320, 8-32
    coils, 101-603 golden-angle spokes, readout 2 x matrix, KB width 4, sigma 2, ramp
    density), every item's RSS image per step.  Three forms over the same resident data:
    plain (one synchronous ksj_gpu_adjoint2_execute per item), one multi-plan call
    (ksj_gpu_adjoint2_execute_multi) after autotuning its launch keys, and that call
    captured once in a HIP graph and replayed."""
    from kspacejet_amd import autotune, synth
    shapes = synth.uneven_items2(count)
    t0 = time.perf_counter()
    plans, ks, outs = [], [], []
    for i, (m, c, spokes) in enumerate(shapes):
        seed = synth.scan_seed(3, i)
        traj = ops.golden_angle2(spokes, 2 * m)
        img = torch.from_numpy(synth.coil_images2(m, c, seed)).cuda()
        k = ops.nudft2_forward_f64(traj, img, (m - 1) / 2, (m - 1) / 2).cpu().numpy()
        ks.append(torch.from_numpy(synth.add_noise(k, seed).astype(np.complex64)).cuda())
        plans.append(ops.Plan2(m, m, traj.shape[0], oversampling=SIGMA, width=WIDTH).set_trajectory(traj))
        outs.append(torch.empty((m, m), dtype=torch.float32, device="cuda"))
        del img, traj
    torch.cuda.synchronize()
    log(f"  {count} items generated in {time.perf_counter() - t0:.1f} s")
    channels = [k.shape[0] for k in ks]
    stream = torch.cuda.Stream()
    sid = stream.cuda_stream

    # plain: one synchronous call per item
    ws1 = torch.empty(max(p.workspace_bytes(c) for p, c in zip(plans, channels)), dtype=torch.uint8, device="cuda")

    def plain():
        for p, k, c, o in zip(plans, ks, channels, outs):
            rc = lib.ksj_gpu_adjoint2_execute(p._handle, k.data_ptr(), c, None, 0, o.data_ptr(), ws1.data_ptr(),
                                              ws1.numel(), sid)
            if rc != 0:
                raise RuntimeError(lib.ksj_gpu_last_error().decode())
            stream.synchronize()

    for _ in range(warmup):
        plain()
    t0 = time.perf_counter()
    for _ in range(steps):
        plain()
    plain_s = (time.perf_counter() - t0) / steps
    del ws1

    # one multi-plan call; workspace sized for the largest lane counts the tuner tries
    saved = {k: ops.tune_get(k) for k in ("multi_lanes", "pass_lanes")}
    ops.tune_set("multi_lanes", max(autotune.EXACT_KEYS["multi_lanes"]))
    ops.tune_set("pass_lanes", max(autotune.EXACT_KEYS["pass_lanes"]))
    ws = torch.empty(ops.adjoint_multi_workspace_bytes(plans, channels), dtype=torch.uint8, device="cuda")
    for k, v in saved.items():
        ops.tune_set(k, v)

    def multi():
        with torch.cuda.stream(stream):
            ops.adjoint_multi(plans, ks, rss=outs, workspace=ws)

    def timed(run):
        for _ in range(warmup):
            run()
        stream.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            run()
        stream.synchronize()
        return (time.perf_counter() - t0) / steps

    untuned_s = timed(multi)
    with torch.cuda.stream(stream):
        tuned = autotune.autotune(multi, f"uneven{count}", budget_s=tune_seconds, rounding=True)
    multi_s = timed(multi)

    # the same call captured once in a HIP graph
    graph = ops.Graph(stream)
    with graph:
        multi()
    graph_s = timed(graph.replay)
    graph.close()
    # both multi-plan forms with the other tuned keys, each captured and replayed
    by_form = {}
    tuned_fused = ops.tune_get("multi_fused")
    for name, form in (("lanes", 0), ("fused", 1), ("chain", 2)):
        ops.tune_set("multi_fused", form)
        g = ops.Graph(stream)
        with g:
            multi()
        by_form[name] = round(1e3 * timed(g.replay), 3)
        g.close()
    ops.tune_set("multi_fused", tuned_fused)
    chain_faults = ops.chain_faults()
    if chain_faults:
        raise RuntimeError(f"the uneven batch's work queue reported {chain_faults} timed-out waits")
    res = {
        "items": count, "shapes": "matrix {128,192,256,320} x coils 8-32 x spokes 101-603, readout 2 x matrix",
        "matrix_counts": {str(m): sum(1 for s in shapes if s[0] == m) for m in synth.UNEVEN_MATRICES},
        "coil_images": int(sum(channels)), "samples": int(sum(c * k.shape[1] for c, k in zip(channels, ks))),
        "plain_images_per_s": round(count / plain_s, 1),
        "multi_untuned_images_per_s": round(count / untuned_s, 1),
        "multi_images_per_s": round(count / multi_s, 1),
        "graph_images_per_s": round(count / graph_s, 1),
        "graph_over_plain": round(plain_s / graph_s, 2),
        "ms_per_batch": {"plain": round(1e3 * plain_s, 3), "multi": round(1e3 * multi_s, 3),
                         "graph": round(1e3 * graph_s, 3)},
        "graph_ms_by_form": by_form,
        "tuned": tuned["settings"], "tune_trials": len(tuned["trials"]),
        "workspace_bytes": ws.numel(),
    }
    for p in plans:
        p.close()
    return res


class MixedStream:
    """Request stream in the reference's mixed proportions: 60 % radial scans of the
    headline shape, 30 % Cartesian scans (matrix 128/256/384, 8-32 coils, centred), 10 %
    3-D kooshball volumes at half the 3-D spoke count (256^3, 20 coils, 8192 x 512
    golden-means spokes, |k|^2 density), in a seeded order.  `served` queues requests on
    the server; `plain` runs one request as a synchronous plain call with resident plans.
    A 2-D image and a 3-D volume each count as one request."""

    def __init__(self, torch, lib, _lib, plan, radial_scans, traj, traj_key, count, stream, volume_fraction,
                 seed=11):
        from kspacejet_amd import ops, synth
        self.torch, self.lib, self.stream = torch, lib, stream
        rng = np.random.default_rng(seed)
        S = SPOKES * READOUT
        cart_shapes = [(int(rng.choice([128, 256, 384])), int(rng.integers(8, 33))) for _ in range(6)]
        cart = [torch.from_numpy(synth.cartesian_kspace2(n, c, synth.scan_seed(5, i)).astype(np.complex64)).cuda()
                for i, (n, c) in enumerate(cart_shapes)]
        n3 = round(volume_fraction * count)
        nc = round(0.3 * count)
        self.kinds = rng.permutation(["radial"] * (count - n3 - nc) + ["cartesian"] * nc + ["volume"] * n3)
        self.vol_n, self.vol_coils, vol_spokes = 256, 20, 8192
        vol_n = self.vol_n
        self.plan3 = None
        if n3:
            traj3 = ops.golden_means3(vol_spokes, READOUT)
            S3 = traj3.shape[0]
            gen = torch.Generator(device="cuda").manual_seed(0x4B53504A ^ (5 << 32))
            self.k3 = torch.randn((self.vol_coils, S3), dtype=torch.complex64, device="cuda", generator=gen)
            traj3_key = int.from_bytes(hashlib.sha256(traj3.cpu().numpy().tobytes()).digest()[:8], "little") | 1
            self.plan3 = ops.Plan3((vol_n, vol_n, vol_n), S3, width=WIDTH).set_trajectory(traj3)
        reqs, self.jobs = [], []
        for i, kind in enumerate(self.kinds):
            if kind == "radial":
                k = radial_scans[i % len(radial_scans)]
                o = torch.empty((N, N), dtype=torch.float32, device="cuda")
                reqs.append(_lib.ServeRequest(C.sizeof(_lib.ServeRequest), _lib.SERVE_RADIAL2D, N, N, COILS, S, WIDTH,
                                              _lib.DENSITY_RAMP, 0, 0, traj_key, traj.data_ptr(), k.data_ptr(), None,
                                              o.data_ptr()))
            elif kind == "cartesian":
                k = cart[i % len(cart)]
                c, n, _ = k.shape
                o = torch.empty((n, n), dtype=torch.float32, device="cuda")
                reqs.append(_lib.ServeRequest(C.sizeof(_lib.ServeRequest), _lib.SERVE_CARTESIAN2D, n, n, c, 0, 0, 0, 1,
                                              0, 0, None, k.data_ptr(), None, o.data_ptr()))
            else:
                k = self.k3
                o = torch.empty((vol_n, vol_n, vol_n), dtype=torch.float32, device="cuda")
                reqs.append(_lib.ServeRequest(C.sizeof(_lib.ServeRequest), _lib.SERVE_RADIAL3D, vol_n, vol_n,
                                              self.vol_coils, S3, WIDTH, _lib.DENSITY_RAMP_SQUARED, 0, vol_n, traj3_key,
                                              traj3.data_ptr(), k.data_ptr(), None, o.data_ptr()))
            self.jobs.append((kind, k, o))
        self.count = count
        self.reqs = (_lib.ServeRequest * count)(*reqs)
        self.ws3 = torch.empty(self.plan3.workspace_bytes(self.vol_coils, 8) if n3 else 1, dtype=torch.uint8,
                               device="cuda")
        self.plan = plan
        self.ws = torch.empty(max(plan.workspace_bytes(COILS),
                                  max(lib.ksj_gpu_ifft2_workspace_bytes(c, n, n) for n, c in cart_shapes)),
                              dtype=torch.uint8, device="cuda")
        self.coil_buf = torch.empty(max(n * n * c for n, c in cart_shapes), dtype=torch.complex64, device="cuda")
        self.cart_shapes = cart_shapes
        self.vol_spokes = vol_spokes

    def describe(self):
        return {"requests": self.count, "radial": int((self.kinds == "radial").sum()),
                "cartesian": int((self.kinds == "cartesian").sum()), "volumes": int((self.kinds == "volume").sum()),
                "cartesian_shapes": [f"{n}x{n}x{c}" for n, c in self.cart_shapes],
                "volume_shape": f"{self.vol_n}^3 x {self.vol_coils} coils, {self.vol_spokes}x{READOUT} spokes"}

    def plain(self, i):
        kind, k, o = self.jobs[i]
        lib, stream = self.lib, self.stream
        if kind == "radial":
            lib.ksj_gpu_adjoint2_execute(self.plan._handle, k.data_ptr(), COILS, None, 0, o.data_ptr(),
                                         self.ws.data_ptr(), self.ws.numel(), stream())
        elif kind == "cartesian":
            c, n, _ = k.shape
            lib.ksj_gpu_ifft2_coil_images(k.data_ptr(), self.coil_buf.data_ptr(), c, n, n, 1, 0, self.ws.data_ptr(),
                                          self.ws.numel(), stream())
            lib.ksj_gpu_rss(self.coil_buf.data_ptr(), o.data_ptr(), n * n, c, stream())
        else:
            lib.ksj_gpu_adjoint3_execute(self.plan3._handle, k.data_ptr(), self.vol_coils, 8, None, o.data_ptr(),
                                         self.ws3.data_ptr(), self.ws3.numel(), stream())
        self.torch.cuda.synchronize()

    def burst(self, server):
        """Every request queued at once; returns (seconds, latencies)."""
        lib = self.lib
        tickets = (C.c_uint64 * self.count)()
        t0 = time.perf_counter()
        server.hold(True)
        if lib.ksj_gpu_serve_submit_batch(server._handle, self.count, self.reqs, tickets) != 0:
            raise RuntimeError(lib.ksj_gpu_last_error().decode())
        server.hold(False)
        server.drain()
        dt = time.perf_counter() - t0
        return dt, [server.wait(t) for t in tickets]

    def plain_all(self):
        t0 = time.perf_counter()
        for i in range(self.count):
            self.plain(i)
        return time.perf_counter() - t0

    def poisson_served(self, server, rate, seed):
        """Open-loop Poisson arrivals at `rate` requests/s, each submitted at its arrival;
        latency = arrival to completion."""
        lib = self.lib
        arrivals = np.cumsum(np.random.default_rng(seed).exponential(1.0 / rate, self.count))
        one = _one_request_array()
        tickets = []
        t_start = time.perf_counter()
        for i, t_arr in enumerate(arrivals):
            while time.perf_counter() - t_start < t_arr:
                pass
            r = one(self.reqs[i])
            t = C.c_uint64(0)
            if lib.ksj_gpu_serve_submit_batch(server._handle, 1, r, C.byref(t)) != 0:
                raise RuntimeError(lib.ksj_gpu_last_error().decode())
            tickets.append(t.value)
        lat = [server.wait(t) for t in tickets]
        return time.perf_counter() - t_start, lat

    def poisson_plain(self, rate, seed):
        """The same arrivals served one at a time, in order, by synchronous plain calls
        (a plain per-request operator loop); latency = arrival to completion."""
        arrivals = np.cumsum(np.random.default_rng(seed).exponential(1.0 / rate, self.count))
        lat = []
        t_start = time.perf_counter()
        for i, t_arr in enumerate(arrivals):
            while time.perf_counter() - t_start < t_arr:
                pass
            self.plain(i)
            lat.append(1e3 * (time.perf_counter() - t_start - t_arr))
        return time.perf_counter() - t_start, lat

    def close(self):
        if self.plan3 is not None:
            self.plan3.close()
        self.jobs = None


def _one_request_array():
    from kspacejet_amd import _lib
    return _lib.ServeRequest * 1


def run_mixed(torch, lib, _lib, server, plan, radial_scans, traj, traj_key, count, stream, volume_fraction=0.1):
    """The mixed stream queued at once (a burst) on the server vs one synchronous plain
    call per request: requests/s and burst latencies."""
    ms = MixedStream(torch, lib, _lib, plan, radial_scans, traj, traj_key, count, stream, volume_fraction)
    ms.burst(server)
    ms.plain_all()
    t_served, lat = ms.burst(server)
    t_plain = ms.plain_all()
    out = ms.describe()
    out.update({"served_requests_per_s": round(count / t_served, 1), "plain_requests_per_s": round(count / t_plain, 1),
                "served_over_plain": round(t_plain / t_served, 2), "burst_p50_ms": round(percentile(lat, 50), 3),
                "burst_p99_ms": round(percentile(lat, 99), 3)})
    ms.close()
    return out


def run_mixed_poisson(torch, lib, _lib, server, plan, radial_scans, traj, traj_key, count, stream, fractions):
    """The mixed stream (10 % volumes) under open-loop Poisson arrivals at several rates,
    each a fraction of the served burst capacity of the stream: per rate, p50/p99
    latency and achieved requests/s, served (the server) vs plain (one synchronous plain
    call per request in arrival order)."""
    ms = MixedStream(torch, lib, _lib, plan, radial_scans, traj, traj_key, count, stream, 0.1, seed=13)
    ms.burst(server)
    t_burst, _ = ms.burst(server)
    t_plain = ms.plain_all()
    capacity = count / t_burst
    rates = []
    for j, f in enumerate(fractions):
        lam = f * capacity
        ts, ls = ms.poisson_served(server, lam, 100 + j)
        tp, lp = ms.poisson_plain(lam, 100 + j)
        entry = {"fraction_of_served_capacity": f, "arrival_rate_per_s": round(lam, 1),
                 "served_p50_ms": round(percentile(ls, 50), 3), "served_p99_ms": round(percentile(ls, 99), 3),
                 "plain_p50_ms": round(percentile(lp, 50), 3), "plain_p99_ms": round(percentile(lp, 99), 3),
                 "served_requests_per_s": round(count / ts, 1), "plain_requests_per_s": round(count / tp, 1),
                 "p99_plain_over_served": round(percentile(lp, 99) / percentile(ls, 99), 2)}
        for cls, sel in (("2d", ms.kinds != "volume"), ("3d", ms.kinds == "volume")):
            if sel.any():
                a, b = np.asarray(ls)[sel], np.asarray(lp)[sel]
                entry[f"by_class_{cls}"] = {"served_p50_ms": round(percentile(a, 50), 3),
                                            "served_p99_ms": round(percentile(a, 99), 3),
                                            "plain_p50_ms": round(percentile(b, 50), 3),
                                            "plain_p99_ms": round(percentile(b, 99), 3)}
        rates.append(entry)
        log(f"  poisson {f:.2f} x capacity: served p99 {rates[-1]['served_p99_ms']} ms, "
            f"plain p99 {rates[-1]['plain_p99_ms']} ms")
    out = ms.describe()
    out.update({"served_capacity_requests_per_s": round(capacity, 1),
                "plain_capacity_requests_per_s": round(count / t_plain, 1),
                "capacity_served_over_plain": round(t_plain / t_burst, 2), "rates": rates,
                "process": "open-loop Poisson arrivals of the seeded mixed order; served: each request submitted at "
                           "its arrival; plain: one synchronous call per request in arrival order"})
    ms.close()
    return out


def run_host_fed(torch, ops, lib, _lib, base_np, traj, traj_key, batch, steps, warmup, lanes, max_batch):
    """Host-memory requests of the headline scan through the server: k-space in host
    memory, RSS back to host memory, bursts of `batch` distinct scans (host buffers of
    26.3 MB each, so nothing is cache resident).  Two forms: pageable caller buffers
    (copied into the server's pinned staging by its copy threads) and pinned caller
    buffers (DMA from and to them).  Beside them the measured PCIe rates (256 MiB copies)
    and the rate the H2D link alone would allow."""
    S = SPOKES * READOUT
    traj_h = traj.cpu()
    in_bytes = COILS * S * 8
    out_bytes = N * N * 4
    server = ops.Server(lanes=lanes, max_batch_scans=max_batch)
    forms = {}
    for form in ("pageable", "pinned"):
        pin = form == "pinned"
        ks = [torch.from_numpy(base_np[i % len(base_np)].copy()) for i in range(batch)]
        if pin:
            ks = [k.pin_memory() for k in ks]
        outs = [torch.empty((N, N), dtype=torch.float32, pin_memory=pin) for _ in range(batch)]
        reqs = (_lib.ServeRequest * batch)(*[
            _lib.ServeRequest(C.sizeof(_lib.ServeRequest), _lib.SERVE_RADIAL2D, N, N, COILS, S, WIDTH,
                              _lib.DENSITY_RAMP, 0, 0, traj_key, traj_h.data_ptr(), k.data_ptr(), None, o.data_ptr(),
                              _lib.SERVE_HOST_MEMORY, 0) for k, o in zip(ks, outs)])
        tickets = (C.c_uint64 * batch)()

        def step():
            if lib.ksj_gpu_serve_submit_batch(server._handle, batch, reqs, tickets) != 0:
                raise RuntimeError(lib.ksj_gpu_last_error().decode())
            return list(tickets)

        for _ in range(max(1, warmup)):
            for t in step():
                server.wait(t)
        all_t = []
        t0 = time.perf_counter()
        for _ in range(steps):
            all_t += step()
        server.drain()
        dt = time.perf_counter() - t0
        lat = [server.wait(t) for t in all_t]
        forms[form] = {"images_per_s": round(batch * steps / dt, 1), "saturated_p50_ms": round(percentile(lat, 50), 3),
                       "saturated_p99_ms": round(percentile(lat, 99), 3), "images": batch * steps}
        del ks, outs, reqs
    server.close()
    # PCIe rates of plain copies, timed with events on torch's stream (where torch copies).
    big = 256 << 20
    dev = torch.empty(big, dtype=torch.uint8, device="cuda")
    rates = {}
    for name, host, h2d in (("h2d_pinned", torch.empty(big, dtype=torch.uint8, pin_memory=True), True),
                            ("d2h_pinned", torch.empty(big, dtype=torch.uint8, pin_memory=True), False),
                            ("h2d_pageable", torch.empty(big, dtype=torch.uint8), True)):
        host.fill_(1)
        for _ in range(2):
            (dev.copy_(host, non_blocking=True) if h2d else host.copy_(dev, non_blocking=True))
        torch.cuda.synchronize()
        reps = 5
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            (dev.copy_(host, non_blocking=True) if h2d else host.copy_(dev, non_blocking=True))
        e1.record()
        e1.synchronize()
        rates[name + "_GBps"] = round(reps * big / (e0.elapsed_time(e1) * 1e-3) / 1e9, 1)
        del host
    del dev
    link_bound = rates["h2d_pinned_GBps"] * 1e9 / (in_bytes + 0.0)
    for f in forms.values():
        f["fraction_of_h2d_link_bound"] = round(f["images_per_s"] / link_bound, 3)
    return {"forms": forms, "pcie": rates, "host_bytes_per_image": in_bytes + out_bytes,
            "h2d_link_bound_images_per_s": round(link_bound, 1), "batch": batch, "steps": steps,
            "server": f"{lanes} lanes, up to {max_batch} scans per launch, 1024 MiB pinned staging, "
                      f"{ops.tune_get('serve_copy_threads')} copy threads"}


if __name__ == "__main__":
    sys.exit(main())
