"""Many small graphs in one persistent launch against the FP64 oracle, graph
by graph (needs the GPU).

Every graph of a block-diagonal batch is the plain Jacobi solve on its own
vertices: L1 <= 1e-6 against the oracle at the graph's own sweep count, the
sweep count within +-1 of the oracle's, the same top order.
"""
import numpy as np
import pytest

import nwgr_amd
import oracle
from nwgr_amd import MultiPageRank
from nwgr_amd._lib import NWGR_EINVAL, NwgrError
from tests.conftest import assert_same_top_k, l1
from tests.test_gpu_pagerank import MODES, say

pytestmark = pytest.mark.gpu


def sizes(count, lo, hi, seed):
    rng = np.random.default_rng(seed)
    return np.exp(rng.uniform(np.log(lo), np.log(hi), count)).astype(np.uint32)


def subgraph(g, vbase, gi):
    v0, v1 = int(vbase[gi]), int(vbase[gi + 1])
    offs = g.offsets[v0:v1 + 1].astype(np.int64)
    idx = g.indices[offs[0]:offs[-1]].astype(np.int64) - v0
    return (offs - offs[0]).astype(np.uint32), idx.astype(np.uint32), v0, v1


def check_graphs(g, vbase, deg, r, mode, which, tol=1e-6, max_iters=100, mp=None, teleport=None):
    iters, errs = mp.graph_results()
    for gi in which:
        offs, idx, v0, v1 = subgraph(g, vbase, gi)
        d = deg[v0:v1]
        t = None if teleport is None else teleport[v0:v1].astype(np.float64)
        r64, _ = oracle.jacobi(offs, idx, d, 0.85, 0.0, int(iters[gi]), dangling=MODES[mode], teleport=t)
        _, k_ref = oracle.jacobi(offs, idx, d, 0.85, tol, max_iters, dangling=MODES[mode], teleport=t)
        err = l1(r[v0:v1], r64)
        assert err <= 1e-6, (gi, err)
        assert abs(int(iters[gi]) - int(k_ref)) <= 1, (gi, iters[gi], k_ref)
        # FP64 ranks closer than 1e-6 relative are within the FP32 sums'
        # rounding (a pair 2.2e-7 apart, two FP32 ulps, swaps when the lanes
        # of a row add in another grouping)
        assert_same_top_k(r[v0:v1], r64, min(20, v1 - v0), tie_rel=1e-6)


@pytest.mark.parametrize("mode", ["leak", "uniform", "teleport"])
@pytest.mark.parametrize("tuning", [dict(lds_max_n=36864, heavy_min=512), dict(lds_max_n=0, heavy_min=16),
                                    dict(lds_max_n=2048, heavy_min=100000)])
def test_multi_graphs_match_oracle(mode, tuning):
    say(f"multi {mode} {tuning}")
    ns = sizes(60, 16, 6000, 3)
    ns[0] = 65536  # the largest allowed graph (global-memory gathers unless lds_max_n allows)
    ns[1] = 2      # the smallest graph the generator makes
    g, vbase = nwgr_amd.rmat_many(ns, 8)
    deg = nwgr_amd.build_degrees(g)
    rng = np.random.default_rng(1)
    tel = None
    if mode == "teleport":
        tel = np.empty(g.n, np.float32)
        for gi in range(len(ns)):
            v0, v1 = vbase[gi], vbase[gi + 1]
            tel[v0:v1] = rng.dirichlet(np.ones(v1 - v0))
    with MultiPageRank(g, vbase, deg) as mp:
        mp.tuning = tuning
        r, res = mp.run(damping=0.85, tol=1e-6, max_iters=100, dangling=mode, teleport=tel)
        assert res.converged
        check_graphs(g, vbase, deg, r, mode, range(len(ns)), mp=mp, teleport=tel)


def test_multi_deterministic_and_tuning_invariant():
    ns = sizes(40, 100, 20000, 4)
    g, vbase = nwgr_amd.rmat_many(ns, 8)
    with MultiPageRank(g, vbase) as mp:
        r1, _ = mp.run(tol=1e-7)
        r2, _ = mp.run(tol=1e-7)
        np.testing.assert_array_equal(r1, r2)
        mp.tuning = {"lds_max_n": 0}
        r3, _ = mp.run(tol=1e-7)
        np.testing.assert_array_equal(r1, r3)  # same sums, LDS or global gathers


def test_multi_rejects_bad_layouts():
    g, vbase = nwgr_amd.rmat_many(np.array([10, 20], np.uint32), 4)
    with pytest.raises(NwgrError) as e:
        MultiPageRank(g, np.array([0, 15, 30], np.uint32))  # edges cross the cut
    assert e.value.code == NWGR_EINVAL
    big = nwgr_amd.build_adjacency([0], [1], n=70000)
    with pytest.raises(NwgrError):
        MultiPageRank(big, np.array([0, 70000], np.uint32))
    with pytest.raises(NwgrError):
        MultiPageRank(g, np.array([0, 30, 30], np.uint32))
    with MultiPageRank(g, vbase) as mp:
        with pytest.raises(NwgrError):
            mp.tuning = {"lds_max_n": 40000}


@pytest.mark.slow
@pytest.mark.timeout(900)
def test_config4b_2048_rmat_graphs():
    """2,048 R-MAT graphs, n log-uniform in [1k, 64k], edge factor 8, one
    launch; every graph sums to 1 (UNIFORM) and every graph against the
    oracle (L1, sweep count, top order)."""
    say("multi 2048 graphs")
    ns = sizes(2048, 1024, 65536, 27491095)
    g, vbase = nwgr_amd.rmat_many(ns, 8)
    deg = nwgr_amd.build_degrees(g)
    with MultiPageRank(g, vbase, deg) as mp:
        r, res = mp.run(damping=0.85, tol=1e-6, max_iters=100, dangling="uniform")
        assert res.converged
        sums = np.add.reduceat(r.astype(np.float64), vbase[:-1].astype(np.int64))
        assert np.abs(sums - 1.0).max() < 1e-4
        iters, _ = mp.graph_results()
        say(f"  {g.m} edges, {res.ms:.1f} ms, sweeps {iters.min()}..{iters.max()}")
        check_graphs(g, vbase, deg, r, "uniform", range(len(ns)), mp=mp)
        say(f"  all {len(ns)} graphs match the oracle")
