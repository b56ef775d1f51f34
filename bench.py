#!/usr/bin/env python3
"""Benchmark of the KSpaceJet 2-D radial reconstruction hot path on one MI355X.

Workload (the headline configuration): 2-D golden-angle radial adjoint NUFFT of a
256x256 image from 16 coils x 402 spokes x 512 samples (2x readout oversampling),
Ram-Lak ramp density compensation, Kaiser-Bessel width 4 with Beatty beta on a
512x512 grid (sigma 2), FP32 gridding/FFT, deapodize, crop, root-sum-of-squares.

A step is one pass of the hot path over one batch of `--batch` scans that are
already resident in HBM (k-space samples + trajectory); the output of every scan
is its RSS image.  `value` is images per second over the timed steps.

    python bench.py [--steps K] [--warmup W] [--batch B] [--gpus 1]

Prints ONE JSON line.  Synthetic data: seeded 10-ellipse phantom x Gaussian coil
maps, k-space from an exact FP64 direct NUDFT on the GPU plus complex Gaussian
noise (sigma = 1e-3 max|k|), stored complex64.
"""
from __future__ import annotations

import argparse
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
    return scans, base[0]


def algorithmic_bytes(stage: str) -> float:
    """Compulsory HBM bytes of one launch per image (cfg figures of the design notes)."""
    S = SPOKES * READOUT
    G = int(N * SIGMA)
    if stage == "grid_gather":
        return 8.0 * S * COILS + 12.0 * S + 8.0 * G * G * COILS
    if stage == "grid_prep":
        return 16.0 * S * COILS + 12.0 * S
    if stage == "fft_rows":
        return 8.0 * G * G * COILS + 8.0 * G * N * COILS
    if stage == "fft_cols_deapod":
        return 8.0 * G * N * COILS + 8.0 * N * N * COILS
    if stage == "coil_combine":
        return 8.0 * N * N * COILS + 4.0 * N * N
    if stage == "grid_reduce":
        return 0.0  # partial sums of blocks cut into several items: no compulsory bytes
    raise KeyError(stage)


TRAFFIC_PROFILE = os.path.join(HERE, "profiles", "radial256_c16_kb4_hbm_traffic.txt")


def measured_traffic(scans_per_launch: int):
    """HBM bytes per launch of the batched gather kernel, from the committed rocprofv3
    FETCH_SIZE / WRITE_SIZE passes of this workload (already scaled by the copy-kernel
    calibration), or None when that summary is absent or was taken at another stacking."""
    try:
        for line in open(TRAFFIC_PROFILE):
            if line.startswith("traffic_bytes_per_launch"):
                _, key, value = line.split()
                if key == f"gather_item_kernel<64,8>@{scans_per_launch}scans":
                    return float(value)
    except OSError:
        pass
    return None


def cpu_baseline(scan_np, traj_np, seconds_hint: float):
    """The reference's own radial route compiled from its sources (oracle/_ref), all host cores."""
    import oracle as O
    if not O.ref_available():
        return None
    threads = min(16, os.cpu_count() or 1)
    t0 = time.perf_counter()
    O.ref_radial_route_rss(scan_np, traj_np, N, N, scans=threads, threads=threads)
    one_round = time.perf_counter() - t0
    rounds = max(1, int(seconds_hint / max(one_round, 1e-3)))
    scans = rounds * threads
    t0 = time.perf_counter()
    O.ref_radial_route_rss(scan_np, traj_np, N, N, scans=scans, threads=threads)
    dt = time.perf_counter() - t0
    return {"value": scans / dt, "unit": "images/s", "cores": threads, "kind": "reference",
            "sample": f"{scans} scans of the 256x256x16-coil 402x512 radial shape through the reference route "
                      f"(ramp DCF + bilinear gridding on 256x256 + radix-2 inverse FFT per coil, then RSS) "
                      f"compiled from the reference's eigen_radial_gridding.cpp; sample-count envelope lifted; "
                      f"{dt:.1f} s wall"}


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--batch", type=int, default=32)
    ap.add_argument("--distinct", type=int, default=4, help="distinct seeded scans behind the batch")
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
    from kspacejet_amd import ops
    from kspacejet_amd._lib import load

    ops.init(0)
    lib = load()
    for kv in args.tune:
        key, value = kv.split("=", 1)
        if lib.ksj_gpu_tune_set(key.encode(), int(value)) != 0:
            raise SystemExit(f"--tune {kv}: {lib.ksj_gpu_last_error().decode()}")
    traj = ops.golden_angle2(SPOKES, READOUT)
    S = SPOKES * READOUT
    log("generating synthetic scans ...")
    scans, scan0 = make_scans(ops, torch, traj, args.batch, min(args.distinct, args.batch))
    plan = ops.Plan2(N, N, S, oversampling=SIGMA, width=WIDTH).set_trajectory(traj)
    ws = torch.empty(plan.workspace_bytes(COILS), dtype=torch.uint8, device="cuda")
    outs = [torch.empty((N, N), dtype=torch.float32, device="cuda") for _ in range(args.batch)]

    import ctypes as C

    stream = lambda: torch.cuda.current_stream().cuda_stream
    per_pass = max(1, 64 // COILS)  # scans stacked per gridding pass (64 coil lanes)
    ws_batch = torch.empty(plan.batch_workspace_bytes(COILS, args.batch), dtype=torch.uint8, device="cuda")
    P = C.c_void_p * args.batch
    k_ptrs = P(*[k.data_ptr() for k in scans])
    r_ptrs = P(*[o.data_ptr() for o in outs])

    def run_plain():
        for k, out in zip(scans, outs):
            lib.ksj_gpu_adjoint2_execute(plan._handle, k.data_ptr(), COILS, None, 0, out.data_ptr(),
                                         ws.data_ptr(), ws.numel(), stream())

    def run_batched():
        rc = lib.ksj_gpu_adjoint2_execute_batch(plan._handle, args.batch, k_ptrs, COILS, None, 0, r_ptrs,
                                                ws_batch.data_ptr(), ws_batch.numel(), stream())
        if rc != 0:
            raise RuntimeError(lib.ksj_gpu_last_error().decode())

    def timed(fn, steps):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            fn()
        torch.cuda.synchronize()
        return time.perf_counter() - t0

    def stage_times(fn, steps):
        lib.ksj_gpu_profile_enable(1)
        lib.ksj_gpu_profile_reset()
        timed(fn, steps)
        out = {}
        for name in ("grid_prep", "grid_gather", "grid_reduce", "fft_rows", "fft_cols_deapod", "coil_combine"):
            tot, cnt = C.c_double(0), C.c_uint64(0)
            lib.ksj_gpu_profile_read(name.encode(), C.byref(tot), C.byref(cnt))
            out[name] = {"avg_us": 1e3 * tot.value / max(cnt.value, 1), "launches": cnt.value}
        lib.ksj_gpu_profile_enable(0)
        return out

    if args.copy_probe:
        src = torch.empty(256 << 20, dtype=torch.uint8, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    images = args.batch * args.steps
    for _ in range(args.warmup):
        run_plain()
        run_batched()
    plain_value = images / timed(run_plain, args.steps)
    elapsed = timed(run_batched, args.steps)
    value = images / elapsed
    stages_plain = stage_times(run_plain, max(1, args.steps // 2))
    stages = stage_times(run_batched, max(1, args.steps // 2))
    scans_per_launch = min(per_pass, args.batch)
    dominant = max(stages, key=lambda s: stages[s]["avg_us"] * stages[s]["launches"])
    g = stages["grid_gather"]
    per_launch_bytes = algorithmic_bytes("grid_gather") * scans_per_launch
    achieved = per_launch_bytes / (g["avg_us"] * 1e-6) / 1e9
    roofline = {"kernel": "grid_gather", "bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": measured_traffic(scans_per_launch),
                "scans_per_launch": scans_per_launch, "algorithmic_bytes_per_launch": per_launch_bytes,
                "avg_us": round(g["avg_us"], 2)}
    stage_report = {k: {"avg_us": round(v["avg_us"], 2),
                        "gbs": round(algorithmic_bytes(k) * scans_per_launch / max(v["avg_us"] * 1e-6, 1e-12) / 1e9, 1)}
                    for k, v in stages.items() if v["launches"]}
    plain_report = {k: round(v["avg_us"], 2) for k, v in stages_plain.items()}
    cpu = None
    if not args.no_cpu_baseline:
        log("timing the CPU reference route ...")
        cpu = cpu_baseline(scan0, traj.cpu().numpy(), args.cpu_seconds)
    line = {
        "metric": METRIC, "value": round(value, 1), "unit": "images/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(1e3 * elapsed / args.steps, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic: seeded 10-ellipse phantom x 16 Gaussian coils, exact FP64 NUDFT k-space + noise, complex64",
        "config": {"workload": "2-D golden-angle radial adjoint NUFFT + RSS, batched (scans on one trajectory "
                               "stacked along the coil axis), device-resident",
                   "image": N, "coils": COILS, "spokes": SPOKES, "readout": READOUT, "kernel": "kaiser_bessel",
                   "kernel_width": WIDTH, "oversampling": SIGMA, "grid": int(N * SIGMA), "batch": args.batch},
        "roofline": roofline, "dominant_stage": dominant, "stages": stage_report, "cpu_baseline": cpu,
        "plain_images_per_s": round(plain_value, 1), "batched_over_plain": round(value / plain_value, 2),
        "plain_stage_us": plain_report,
    }
    print(json.dumps(line))
    return 0


if __name__ == "__main__":
    sys.exit(main())
