"""Autotuner for the sweep work decomposition.

``autotune(pg)`` tunes a ``PageRankGraph``: first the sweep engine (degree-
binned or tiled), then that engine's knobs (``nwgr_pr_tuning`` in
include/nwgr_pagerank.h) by coordinate descent --
  * tiled: the source tile width 2^tile_bits and the fill cost that cuts the
    chunk sequence into workgroup ranges;
  * binned: the LDS-staged hot prefix of the gathered vector, the hub
    threshold and chunk size, the unit size of the other rows and the number
    of sweep workgroups per compute unit.
``autotune_multi(mp)`` tunes a ``MultiPageRank`` batch: the largest graph whose
contributions are mirrored in LDS and the row length above which a row is
reduced by a wavefront.

Every candidate is timed on a fixed number of sweeps (tolerance 0, so all
candidates do the same work); the fastest is kept and, with a cache path,
stored per (device architecture, graph signature) in a JSON file.

No reference counterpart: the reference has no tunable decomposition.
"""
from __future__ import annotations

import hashlib
import json
import os
import time

import numpy as np

from .page_rank import PageRankGraph

ENGINE_BINNED, ENGINE_TILED = 0, 1

SEARCH_SPACE_TILED = {
    "tile_bits": [13, 14, 15],
    "tile_fill_cost": [0, 10, 40, 160],
    "tile_combine": [0, 1],
}

SEARCH_SPACE_BINNED = {
    "lds_cache": [0, 8192, 16384, 36864],
    "unit_edges": [1024, 2048, 4096, 8192, 16384],
    "hub_threshold": [1024, 4096, 16384],
    "hub_chunk": [1024, 2048, 8192],
    "blocks_per_cu": [1, 2],
    "zero_rows_per_unit": [1024, 4096, 16384],
}

SEARCH_SPACE_MULTI = {
    "lds_max_n": [0, 4096, 16384, 36864],
    "heavy_min": [64, 128, 256, 512, 1024],
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
    pg.run(out=out, **kw)  # capture the chunk graphs (and build the layout)
    total_ms, launches = 0.0, 0
    for _ in range(reps):
        pg.run(out=out, **kw)
        prof = pg.last_profile()
        total_ms += prof["sweep_ms"]
        launches += prof["launches"]
    return total_ms / max(launches, 1) * 1e3


def _valid(t: dict) -> bool:
    return t["lds_cache"] * t["blocks_per_cu"] <= 36864


def _cache_get(cache_path, key):
    if cache_path and key and os.path.exists(cache_path):
        with open(cache_path) as f:
            cache = json.load(f)
        return cache.get(key)
    return None


def _cache_put(cache_path, key, entry):
    if not (cache_path and key):
        return
    cache = {}
    if os.path.exists(cache_path):
        with open(cache_path) as f:
            cache = json.load(f)
    cache[key] = entry
    os.makedirs(os.path.dirname(os.path.abspath(cache_path)), exist_ok=True)
    with open(cache_path, "w") as f:
        json.dump(cache, f, indent=1)


def _descend(apply, measure, start: dict, space: dict, rounds: int, log, valid=lambda t: True):
    """Coordinate descent from `start` over `space`; `apply` sets a candidate
    (raising if the library rejects it), `measure` times the current one."""
    best = dict(start)
    apply(best)
    best_us = measure()
    trials = 1
    for _ in range(rounds):
        improved = False
        for name, values in space.items():
            for v in values:
                if v == best[name]:
                    continue
                cand = dict(best, **{name: v})
                if not valid(cand):
                    continue
                try:
                    apply(cand)
                except Exception:  # rejected by the library's validation
                    continue
                us = measure()
                trials += 1
                if log:
                    log(f"autotune {name}={v}: {us:.1f} us")
                if us < best_us * 0.99:
                    best, best_us, improved = cand, us, True
        if not improved:
            break
    apply(best)
    return best, best_us, trials


def autotune(pg: PageRankGraph, in_degrees=None, sweeps: int = 8, reps: int = 2, rounds: int = 2,
             cache_path: str | None = None, arch: str = "gfx950", log=None, engines=(ENGINE_BINNED, ENGINE_TILED),
             **run_kw) -> dict:
    """Tune `pg` in place; returns {"tuning", "us_per_sweep", "trials", "seconds"}."""
    key = f"{arch}:{graph_signature(in_degrees, pg.m)}" if in_degrees is not None else None
    hit = _cache_get(cache_path, key)
    if hit:
        pg.tuning = hit["tuning"]
        return dict(hit, cached=True)
    t0 = time.perf_counter()

    def apply(t):
        pg.tuning = t

    def measure():
        return time_sweeps(pg, sweeps, reps, **run_kw)

    best, best_us, trials = None, float("inf"), 0
    for eng in engines:
        start = dict(pg.tuning, engine=eng)
        space = SEARCH_SPACE_TILED if eng == ENGINE_TILED else SEARCH_SPACE_BINNED
        t, us, n = _descend(apply, measure, start, space, rounds, log,
                            valid=_valid if eng == ENGINE_BINNED else (lambda c: True))
        trials += n
        if log:
            log(f"autotune engine {eng}: {us:.1f} us/sweep")
        if us < best_us:
            best, best_us = t, us
    pg.tuning = best
    result = {"tuning": best, "us_per_sweep": best_us, "trials": trials, "seconds": time.perf_counter() - t0}
    _cache_put(cache_path, key, {"tuning": best, "us_per_sweep": best_us})
    return result


def autotune_multi(mp, sweeps: int = 8, reps: int = 2, rounds: int = 2, cache_path: str | None = None,
                   arch: str = "gfx950", log=None, **run_kw) -> dict:
    """Tune a ``MultiPageRank`` in place on `sweeps` sweeps of every graph;
    returns {"tuning", "ms_per_run", "trials", "seconds"}."""
    import torch
    key = f"{arch}:multi-n{mp.n}-g{mp.count}" if cache_path else None
    hit = _cache_get(cache_path, key)
    if hit:
        mp.tuning = hit["tuning"]
        return dict(hit, cached=True)
    kw = dict(damping=0.85, tol=0.0, max_iters=sweeps, dangling="uniform")
    kw.update(run_kw)
    out = torch.empty(mp.n, dtype=torch.float32, device="cuda:0")
    t0 = time.perf_counter()

    def apply(t):
        mp.tuning = t

    def measure():
        mp.run_device(out.data_ptr(), **kw)  # first run of a setting
        return sum(mp.run_device(out.data_ptr(), **kw).ms for _ in range(reps)) / reps * 1e3

    best, best_us, trials = _descend(apply, measure, mp.tuning, SEARCH_SPACE_MULTI, rounds, log)
    result = {"tuning": best, "ms_per_run": best_us / 1e3, "trials": trials, "seconds": time.perf_counter() - t0}
    _cache_put(cache_path, key, {"tuning": best, "ms_per_run": best_us / 1e3})
    return result
