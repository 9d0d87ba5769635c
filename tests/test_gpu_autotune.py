"""The autotuners (nwgr_amd.autotune) on the GPU: every setting they pick is a
valid one of their search space, tuning never changes what a solve computes
beyond the engines' own parity gates, and the cache returns the stored pick.
"""
import json

import numpy as np
import pytest

import nwgr_amd
from nwgr_amd import MultiPageRank, PageRankGraph
from nwgr_amd.autotune import (SEARCH_SPACE_BINNED, SEARCH_SPACE_MULTI, SEARCH_SPACE_TILED, autotune,
                               autotune_multi)
from tests.test_gpu_multi import check_graphs, sizes
from tests.test_gpu_pagerank import parity, say

pytestmark = pytest.mark.gpu


def test_autotune_graph_picks_valid_setting_and_stays_exact(tmp_path):
    g = nwgr_amd.rmat(15, 16)
    deg = nwgr_amd.build_degrees(g)
    pg = PageRankGraph(g, deg)
    cache = str(tmp_path / "tune.json")
    res = autotune(pg, in_degrees=g.in_degrees(), sweeps=4, reps=1, rounds=1, cache_path=cache, log=say)
    t = res["tuning"]
    say(f"autotune pick {t} {res['us_per_sweep']:.1f} us/sweep, {res['trials']} trials")
    assert t["engine"] in (0, 1)
    space = SEARCH_SPACE_TILED if t["engine"] == 1 else SEARCH_SPACE_BINNED
    for name, values in space.items():
        assert t[name] in values or t[name] == PageRankGraph(g, deg).tuning[name], (name, t[name])
    assert pg.tuning == t
    parity(g, deg, "uniform", 1e-6, 100, device_graph=pg)
    # the cache hands back the same pick without timing anything
    pg2 = PageRankGraph(g, deg)
    hit = autotune(pg2, in_degrees=g.in_degrees(), cache_path=cache)
    assert hit.get("cached") and hit["tuning"] == t and pg2.tuning == t
    with open(cache) as f:
        assert len(json.load(f)) == 1


def test_autotune_multi_picks_valid_setting_and_stays_exact():
    ns = sizes(48, 64, 40000, 9)
    g, vbase = nwgr_amd.rmat_many(ns, 8)
    deg = nwgr_amd.build_degrees(g)
    with MultiPageRank(g, vbase, deg) as mp:
        r0, _ = mp.run(tol=1e-6)
        res = autotune_multi(mp, sweeps=4, reps=1, rounds=1, log=say)
        t = res["tuning"]
        say(f"autotune_multi pick {t} {res['ms_per_run']:.3f} ms")
        for name, values in SEARCH_SPACE_MULTI.items():
            assert t[name] in values or t[name] == {"lds_max_n": 36864, "heavy_min": 512}[name]
        assert mp.tuning == t
        r1, res1 = mp.run(tol=1e-6)
        assert res1.converged
        if t["heavy_min"] == 512:
            # only the row-length threshold (lanes per row) changes the summation order
            np.testing.assert_array_equal(r0, r1)
        check_graphs(g, vbase, deg, r1, "uniform", range(len(ns)), mp=mp)
