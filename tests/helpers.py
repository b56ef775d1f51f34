"""Small shared helpers for the tests."""
from __future__ import annotations

import concurrent.futures as cf

import numpy as np


def rel_l2(x, ref) -> float:
    x = np.asarray(x)
    ref = np.asarray(ref)
    den = np.linalg.norm(ref.ravel())
    return float(np.linalg.norm((x - ref).ravel()) / (den if den > 0 else 1.0))


def per_coil(fn, n_coils: int, workers: int = 8):
    """Runs fn(c) for every coil on host threads (the oracle releases the GIL) and stacks the results."""
    with cf.ThreadPoolExecutor(max_workers=workers) as ex:
        return list(ex.map(fn, range(n_coils)))
