"""One GPU's share of a row-partitioned solve (needs the GPU).

A multi-GPU solve would give each GPU a contiguous slice of the destination
rows (balanced by entry count) and the full gathered vector; each sweep every
GPU computes its slice, then the slices are exchanged.  Here each share runs on
the one GPU as the same n-vertex graph with entries only in its rows: one
sweep of every share from the same input, put together, must equal one sweep
of the whole graph, which the FP64 oracle pins (L1 <= 1e-6 after one sweep).
"""
import numpy as np
import pytest

import nwgr_amd
import oracle
from nwgr_amd import PageRankGraph
from tests.conftest import l1
from tests.test_gpu_engines import ENGINES
from tests.test_gpu_pagerank import say

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("ename", ["binned", "tiled15"])
@pytest.mark.parametrize("parts", [3, 8])
def test_row_shards_compose_one_sweep(ename, parts):
    say(f"shards {parts} {ename}")
    g = nwgr_amd.rmat(16, 16)
    deg = nwgr_amd.build_degrees(g)
    r64, _ = oracle.jacobi(g.offsets, g.indices, deg, 0.85, 0.0, 1, dangling=oracle.UNIFORM)
    full = PageRankGraph(g, deg)
    full.tuning = ENGINES[ename]
    r_full, res = full.run(tol=0.0, max_iters=1, dangling="uniform")
    assert res.iters == 1
    cuts = nwgr_amd.row_cuts(g, parts)
    assert cuts[0] == 0 and cuts[-1] == g.n and np.all(np.diff(cuts) >= 0)
    got = np.empty(g.n, np.float32)
    for k in range(parts):
        lo, hi = int(cuts[k]), int(cuts[k + 1])
        sh = nwgr_amd.row_shard(g, lo, hi)
        with PageRankGraph(sh, deg) as pg:  # the global out-degrees divide the contributions
            pg.tuning = ENGINES[ename]
            r, res = pg.run(tol=0.0, max_iters=1, dangling="uniform")
            assert res.iters == 1
        got[lo:hi] = r[lo:hi]
    assert l1(got, r64) <= 1e-6
    assert l1(got, r_full) <= 1e-6
