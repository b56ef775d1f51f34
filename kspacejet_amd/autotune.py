"""A small autotuner for the launch parameters of libksj_gpu.so (ksj_gpu_tune_set).

It times a caller-supplied workload -- a function that enqueues GPU work on torch's
current stream -- with HIP events, and walks the tuning keys one at a time
(coordinate descent): each key is set to each of its candidate values with the
others held at their best so far, and keeps the fastest.  The default keys leave the
results bit-identical; the FFT radix schedule (rounding only) and the work-item size
(summation order of new plans) are tuned only when asked for, and every value of
them is tested against the FP64 oracle, so tuning never trades accuracy for speed.  Chosen settings are kept
per workload signature in a JSON file when a path is given, so a later process can
apply them without re-measuring.

    best = autotune(lambda: run_batch(), signature="uneven64", budget_s=20)
"""
from __future__ import annotations

import json
import os
import time
from typing import Callable, Dict, Iterable, Optional

import torch

from . import ops

# Candidate values per key; every value gives bit-identical results.  The gather keys
# are those of the default (matrix) schedule; plans of the 2x2-block schedule tune
# gather_batch (4, 8, 16) and gather_coils_per_lane (1, 2) instead.
EXACT_KEYS: Dict[str, tuple] = {
    "matrix_batch": (16, 8, 4),
    "matrix_slab": (16, 1, 64),
    "matrix_pipeline": (1, 2),
    "matrix_run_cost": (12, 8, 16),
    "matrix_tap_share": (1, 0),
    "matrix_persistent": (0, 8),
    "gather_map": (0, 1),
    "fft_lines_per_block": (0, 8, 32),
    "pass_lanes": (2, 1, 3),
    "multi_lanes": (4, 2, 6, 8),
    "multi_fused": (0, 1, 2),
}
# Change rounding only: every FFT radix schedule is an exact factorisation of the DFT
# (tested against the FP64 oracle at every length); tuned when `rounding=True`.
ROUNDING_KEYS: Dict[str, tuple] = {"fft_schedule": (0, 3, 1, 2)}
# Changes the float summation order of the gather of new plans (results stay within
# tolerance); needs `rebuild`.
ORDER_KEYS: Dict[str, tuple] = {"matrix_item_visits": (128, 64, 256)}
# Keys read when a plan is built: tried only when a `rebuild` is given.
PLAN_KEYS = {"matrix_item_visits", "item_visits", "matrix_run_cost"}


def time_ms(run: Callable[[], None], repeats: int = 5) -> float:
    """Median wall time of `run` on torch's current stream, in ms, after one warm-up call."""
    run()
    torch.cuda.synchronize()
    samples = []
    for _ in range(repeats):
        start, end = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        start.record()
        run()
        end.record()
        end.synchronize()
        samples.append(start.elapsed_time(end))
    samples.sort()
    return samples[len(samples) // 2]


def autotune(run: Callable[[], None], signature: str, keys: Optional[Dict[str, Iterable[int]]] = None,
             repeats: int = 5, budget_s: float = 30.0, cache_path: Optional[str] = None,
             rebuild: Optional[Callable[[], None]] = None, rounding: bool = False) -> Dict[str, object]:
    """Tunes `keys` (default: every bit-exact key, plus the FFT radix schedule when
    `rounding`) for `run` and leaves the best values set.

    `rebuild` is called after a key that only affects new plans changes (PLAN_KEYS),
    so the workload picks it up; without it those keys are skipped.  Returns {"settings": {...}, "ms": best time,
    "baseline_ms": time with the starting settings, "trials": [...]}.
    """
    if cache_path and os.path.exists(cache_path):
        cache = json.load(open(cache_path))
        if signature in cache:
            for k, v in cache[signature]["settings"].items():
                ops.tune_set(k, v)
            return dict(cache[signature], cached=True)
    keys = dict(EXACT_KEYS if keys is None else keys)
    if rounding:
        keys.update(ROUNDING_KEYS)
    t_end = time.perf_counter() + budget_s
    best = {k: ops.tune_get(k) for k in keys}
    best_ms = baseline = time_ms(run, repeats)
    trials = [{"settings": dict(best), "ms": round(best_ms, 4)}]
    for key, values in keys.items():
        if key in PLAN_KEYS and rebuild is None:
            continue
        for v in values:
            if v == best[key] or time.perf_counter() > t_end:
                continue
            ops.tune_set(key, v)
            if key in PLAN_KEYS:
                rebuild()
            ms = time_ms(run, repeats)
            trials.append({"settings": dict(best, **{key: v}), "ms": round(ms, 4)})
            if ms < best_ms:
                best_ms, best[key] = ms, v
            else:
                ops.tune_set(key, best[key])
                if key in PLAN_KEYS:
                    rebuild()
    result = {"settings": best, "ms": round(best_ms, 4), "baseline_ms": round(baseline, 4), "trials": trials}
    if cache_path:
        cache = json.load(open(cache_path)) if os.path.exists(cache_path) else {}
        cache[signature] = {"settings": best, "ms": result["ms"], "baseline_ms": result["baseline_ms"]}
        json.dump(cache, open(cache_path, "w"), indent=1)
    return result
