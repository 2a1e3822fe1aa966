"""Autotuner for libtsg's tuning parameters.

The reference fixes its thresholds at compile time (tile-bin limits and
thread counts, main_spgemm-col.cu:26-31; nsparse bin sizes and hash-table
limits, spgemm_nsparse_kernel.h:4-14).  Here the matching choices are
run-time parameters of a handle (tsg_set_tile_density, tsg_set_row_bins),
so they can be searched on the workload at hand:

* ``tile_density``   -- AUTO takes the tile path when A holds at least this
                        many entries per non-empty 16x16 tile;
* ``row_bins``       -- rows with at most (b1, b2, b3) intermediate products
                        go to the 16-, 32- and 64-lane LDS-hash kernels (hash
                        tables of 128, 512 and 2048 slots), longer rows to the
                        big-row bitmap path;
* ``block_pointer_budget`` -- -1: the library decides whether the big-row
                        path builds its B block-pointer table or searches B
                        rows; 6 GiB: the table is built whenever it fits
                        (a numeric phase replayed many times may prefer it).

Parameters change speed only: every setting gives bitwise the same C.  The
search is a coordinate descent from the defaults: one parameter at a time,
each candidate timed as the minimum of ``reps`` runs of one full product
(symbolic + numeric, outputs allocated as a caller would), a change kept when
it is at least ``min_gain`` faster; two sweeps at most.
"""
from __future__ import annotations

import json
import math
import os
import time
from typing import Callable, Dict, List, Optional

import numpy as np

DEFAULT = {"tile_density": 32.0, "row_bins": (64, 256, 1024), "block_pointer_budget": -1}

# candidate values of each coordinate (validity: b1 <= b2 <= b3, b1 <= 64,
# b2 <= 256, b3 <= 1024)
SPACE = {
    "bin3": (256, 512, 1024),
    "bin2": (64, 128, 256),
    "bin1": (16, 32, 64),
    "tile_density": (8.0, 16.0, 32.0, 64.0, 1e9),
    "block_pointer_budget": (-1, 6 << 30),
}


def valid(t: dict) -> bool:
    b1, b2, b3 = t["row_bins"]
    return 0 <= b1 <= 64 and b1 <= b2 <= 256 and b2 <= b3 <= 1024 and t["tile_density"] > 0


def _with(t: dict, key: str, v) -> dict:
    out = {"tile_density": t["tile_density"], "row_bins": tuple(t["row_bins"]),
           "block_pointer_budget": t.get("block_pointer_budget", -1)}
    if key == "tile_density":
        out["tile_density"] = float(v)
    elif key == "block_pointer_budget":
        out["block_pointer_budget"] = int(v)
    else:
        b = list(out["row_bins"])
        b[{"bin1": 0, "bin2": 1, "bin3": 2}[key]] = int(v)
        out["row_bins"] = tuple(b)
    return out


def coordinate_descent(measure: Callable[[dict], float], start: Optional[dict] = None,
                       space: Optional[Dict[str, tuple]] = None, sweeps: int = 2,
                       min_gain: float = 0.02, log: Optional[Callable[[str], None]] = None):
    """Minimize measure(tuning) over the space; returns (best, best_ms,
    trials) with trials a list of (tuning, ms)."""
    space = space or SPACE
    cur = {**DEFAULT, **(start or {})}
    cur["row_bins"] = tuple(cur["row_bins"])
    cache: Dict[str, float] = {}
    trials: List[tuple] = []

    def timed(t):
        key = json.dumps(t, sort_keys=True)
        if key not in cache:
            cache[key] = measure(t)
            trials.append((t, cache[key]))
            if log:
                log(f"[autotune] {t} -> {cache[key]:.3f} ms")
        return cache[key]

    best_ms = timed(cur)
    for _ in range(sweeps):
        changed = False
        for key, values in space.items():
            for v in values:
                cand = _with(cur, key, v)
                if not valid(cand) or cand == cur:
                    continue
                ms = timed(cand)
                if ms < best_ms * (1.0 - min_gain):
                    cur, best_ms, changed = cand, ms, True
        if not changed:
            break
    return cur, best_ms, trials


def measure_runs(handle, run: Callable[[], None], reps: int = 3) -> Callable[[dict], float]:
    """A measure() for coordinate_descent: applies the tuning to the handle,
    runs run() once untimed, then reps timed runs (device synchronized on
    both sides); minimum in milliseconds."""
    import torch

    def measure(t: dict) -> float:
        handle.set_tuning(t)
        run()
        torch.cuda.synchronize()
        best = math.inf
        for _ in range(reps):
            t0 = time.perf_counter()
            run()
            torch.cuda.synchronize()
            best = min(best, (time.perf_counter() - t0) * 1e3)
        return best

    return measure


def batch_run(handle, pairs):
    """One batched product of pairs (list of (A_i, B_i) DeviceCsr)."""
    import torch

    def run():
        arr, ic, per, roff, tot = handle.batch_symbolic(pairs)
        dev = ic.device
        jc = torch.empty(max(tot, 1), dtype=torch.int32, device=dev)
        c = torch.empty(max(tot, 1), dtype=torch.float64, device=dev)
        handle.batch_numeric(arr, ic, jc, c)

    return run


def batch_numeric_measure(handle, pairs, reps: int = 3) -> Callable[[dict], float]:
    """A measure() for the numeric phase alone -- what a captured HIP graph
    replays: per tuning, one batched symbolic phase (untimed), then the
    numeric phase once untimed and reps times timed; minimum in ms."""
    import torch

    def measure(t: dict) -> float:
        handle.set_tuning(t)
        arr, ic, per, roff, tot = handle.batch_symbolic(pairs)
        dev = ic.device
        jc = torch.empty(max(tot, 1), dtype=torch.int32, device=dev)
        c = torch.empty(max(tot, 1), dtype=torch.float64, device=dev)
        handle.batch_numeric(arr, ic, jc, c)
        torch.cuda.synchronize()
        best = math.inf
        for _ in range(reps):
            t0 = time.perf_counter()
            handle.batch_numeric(arr, ic, jc, c)
            torch.cuda.synchronize()
            best = min(best, (time.perf_counter() - t0) * 1e3)
        return best

    return measure


def single_run(handle, A, B):
    """One product C = A * B."""
    def run():
        handle.spgemm(A, B)
    return run


def signature(pairs) -> str:
    """Cache key of a workload: count, row-count and entries-per-row ranges
    (rounded to powers of two)."""
    rows = np.array([A.rows for A, _ in pairs])
    deg = np.array([A.nnz / max(A.rows, 1) for A, _ in pairs])
    lg = lambda x: int(round(math.log2(max(x, 1))))  # noqa: E731
    return (f"n{len(pairs)}_rows{lg(rows.min())}-{lg(rows.max())}"
            f"_deg{lg(deg.min() + 1)}-{lg(deg.max() + 1)}")


def tune(handle, run, reps: int = 3, cache_path: Optional[str] = None,
         key: Optional[str] = None, log=None,
         measure: Optional[Callable[[dict], float]] = None) -> dict:
    """Tune handle for run() (or for a measure(tuning) -> ms given instead,
    such as batch_numeric_measure); returns {"tile_density", "row_bins",
    "ms", "default_ms", "trials"} and leaves the best tuning applied.  With
    cache_path and key, a stored result is reused instead of searching."""
    if cache_path and key and os.path.exists(cache_path):
        with open(cache_path) as f:
            stored = json.load(f)
        if key in stored:
            res = stored[key]
            handle.set_tuning(res)
            return res
    if measure is None:
        measure = measure_runs(handle, run, reps)
    best, best_ms, trials = coordinate_descent(measure, log=log)
    default_ms = next(ms for t, ms in trials if t == {**DEFAULT,
                                                      "row_bins": tuple(DEFAULT["row_bins"])})
    res = {"tile_density": best["tile_density"], "row_bins": list(best["row_bins"]),
           "block_pointer_budget": best["block_pointer_budget"],
           "ms": best_ms, "default_ms": default_ms,
           "trials": [[t["tile_density"], list(t["row_bins"]), t["block_pointer_budget"], ms]
                      for t, ms in trials]}
    handle.set_tuning(best)
    if cache_path and key:
        stored = {}
        if os.path.exists(cache_path):
            with open(cache_path) as f:
                stored = json.load(f)
        stored[key] = res
        with open(cache_path, "w") as f:
            json.dump(stored, f, indent=1)
    return res
