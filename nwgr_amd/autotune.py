"""Autotuner for the sweep work decomposition of a ``PageRankGraph``.

The knobs (``nwgr_pr_tuning`` in include/nwgr_pagerank.h) decide how the
degree-sorted rows are cut into work units: the LDS-staged hot prefix of the
gathered vector, the hub threshold and chunk size, the unit size of the other
rows and the number of sweep workgroups per compute unit.  The tuner times a
fixed number of sweeps (tolerance 0, so every candidate does the same work)
for each candidate by coordinate descent, keeps the fastest, and caches the
result per (device architecture, graph signature) in a JSON file.

No reference counterpart: the reference has no tunable decomposition.
"""
from __future__ import annotations

import hashlib
import json
import os
import time

import numpy as np

from .page_rank import PageRankGraph

SEARCH_SPACE = {
    "lds_cache": [0, 8192, 16384, 36864],
    "unit_edges": [1024, 2048, 4096, 8192, 16384],
    "hub_threshold": [1024, 4096, 16384],
    "hub_chunk": [1024, 2048, 8192],
    "blocks_per_cu": [1, 2],
    "zero_rows_per_unit": [1024, 4096, 16384],
}


def graph_signature(in_degrees: np.ndarray, m: int) -> str:
    """Stable key of a graph's shape: n, m and a log2 histogram of in-degrees."""
    d = np.asarray(in_degrees, dtype=np.int64)
    hist = np.bincount(np.floor(np.log2(d + 1)).astype(np.int64), minlength=33)
    h = hashlib.sha1(hist.tobytes()).hexdigest()[:16]
    return f"n{len(d)}-m{m}-h{h}"


def time_sweeps(pg: PageRankGraph, sweeps: int = 8, reps: int = 2, **run_kw) -> float:
    """Average device time of one sweep launch (us) over `reps` solves of
    exactly `sweeps` sweeps."""
    kw = dict(damping=0.85, tol=0.0, max_iters=sweeps, dangling="uniform")
    kw.update(run_kw)
    out = np.empty(pg.n, np.float32)
    pg.run(out=out, **kw)  # capture the chunk graphs
    total_ms, launches = 0.0, 0
    for _ in range(reps):
        pg.run(out=out, **kw)
        prof = pg.last_profile()
        total_ms += prof["sweep_ms"]
        launches += prof["launches"]
    return total_ms / max(launches, 1) * 1e3


def _valid(t: dict) -> bool:
    return t["lds_cache"] * t["blocks_per_cu"] <= 36864


def autotune(pg: PageRankGraph, in_degrees=None, sweeps: int = 8, reps: int = 2, rounds: int = 2,
             cache_path: str | None = None, arch: str = "gfx950", log=None, **run_kw) -> dict:
    """Tune `pg` in place; returns {"tuning", "us_per_sweep", "trials"}."""
    key = None
    if cache_path and in_degrees is not None:
        key = f"{arch}:{graph_signature(in_degrees, pg.m)}"
        if os.path.exists(cache_path):
            with open(cache_path) as f:
                cache = json.load(f)
            if key in cache:
                pg.tuning = cache[key]["tuning"]
                return dict(cache[key], cached=True)
    best = pg.tuning
    best_us = time_sweeps(pg, sweeps, reps, **run_kw)
    trials = [(dict(best), best_us)]
    t0 = time.perf_counter()
    for _ in range(rounds):
        improved = False
        for name, values in SEARCH_SPACE.items():
            for v in values:
                if v == best[name]:
                    continue
                cand = dict(best, **{name: v})
                if not _valid(cand):
                    continue
                try:
                    pg.tuning = cand
                except Exception:  # rejected by the library's validation
                    continue
                us = time_sweeps(pg, sweeps, reps, **run_kw)
                trials.append((cand, us))
                if log:
                    log(f"autotune {name}={v}: {us:.1f} us/sweep")
                if us < best_us * 0.99:
                    best, best_us, improved = cand, us, True
        if not improved:
            break
    pg.tuning = best
    result = {"tuning": best, "us_per_sweep": best_us, "trials": len(trials),
              "seconds": time.perf_counter() - t0}
    if cache_path and key:
        cache = {}
        if os.path.exists(cache_path):
            with open(cache_path) as f:
                cache = json.load(f)
        cache[key] = {"tuning": best, "us_per_sweep": best_us}
        os.makedirs(os.path.dirname(os.path.abspath(cache_path)), exist_ok=True)
        with open(cache_path, "w") as f:
            json.dump(cache, f, indent=1)
    return result
