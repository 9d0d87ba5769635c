"""Both sweep engines (degree-binned and tiled, at several tile widths) against
the FP64 oracle, and the tuning interface that selects them (needs the GPU).

The gates are those of test_gpu_pagerank: L1 <= 1e-6 against the oracle at
the same sweep count, sweep count within +-1, identical top-100 order.
"""
import numpy as np
import pytest
import torch

import nwgr_amd
import oracle
from nwgr_amd import PageRankGraph
from nwgr_amd._lib import NWGR_EINVAL, NwgrError
from tests.conftest import assert_same_top_k, l1
from tests.test_gpu_pagerank import parity, say

pytestmark = pytest.mark.gpu

ENGINES = {
    "binned": dict(engine=0),
    "tiled15": dict(engine=1, tile_bits=15),
    "tiled14": dict(engine=1, tile_bits=14),
    "tiled13": dict(engine=1, tile_bits=13, tile_fill_cost=0),
}


def graphs():
    rng = np.random.default_rng(11)
    n = 150_000
    src = np.concatenate([np.arange(1, 120_001), rng.integers(0, n, 500_000)])
    dst = np.concatenate([np.zeros(120_000, np.int64), rng.integers(0, n, 500_000)])
    return {
        "rmat16": lambda: nwgr_amd.rmat(16, 16),
        "er17": lambda: nwgr_amd.erdos_renyi(1 << 17, 8 << 17),
        "hub": lambda: nwgr_amd.build_adjacency(src, dst, n=n, dedup=True, drop_self_loops=True),
    }


@pytest.mark.parametrize("gname", ["rmat16", "er17", "hub"])
@pytest.mark.parametrize("ename", list(ENGINES))
def test_engine_parity(gname, ename):
    say(f"engine {ename} on {gname}")
    g = graphs()[gname]()
    deg = nwgr_amd.build_degrees(g)
    pg = PageRankGraph(g, deg)
    pg.tuning = ENGINES[ename]
    for mode in ("uniform", "leak"):
        parity(g, deg, mode, 1e-6, 100, device_graph=pg)
        assert pg.last_engine() == "binned" if ename == "binned" else pg.last_engine().startswith("tiled")


@pytest.mark.parametrize("n,bits", [(32767, 15), (32768, 15), (32769, 15), (65537, 15), (8193, 13), (16385, 14)])
def test_tiled_vertex_counts_at_tile_boundaries(n, bits):
    """Vertex counts one below, at and one above a tile width: the last tile
    (partly padding, or one vertex) and the quiet rows behind it."""
    say(f"tiled bits {bits} n {n}")
    rng = np.random.default_rng(n)
    m = 12 * n
    src = rng.integers(0, n, m)
    dst = np.minimum(rng.geometric(8.0 / n, m) - 1, n - 1)  # skewed rows: about 40 % without entries
    g = nwgr_amd.build_adjacency(src, dst, n=n, dedup=True, drop_self_loops=True)
    deg = nwgr_amd.build_degrees(g)
    pg = PageRankGraph(g, deg)
    pg.tuning = dict(engine=1, tile_bits=bits, tile_fill_cost=0)
    for mode in ("uniform", "leak"):
        parity(g, deg, mode, 1e-6, 100, device_graph=pg)
    assert pg.last_engine().startswith("tiled")


def test_engines_agree_and_switch_back_and_forth():
    """Switching engines on one handle re-plans without stale state."""
    g = nwgr_amd.rmat(15, 16)
    deg = nwgr_amd.build_degrees(g)
    pg = PageRankGraph(g, deg)
    outs = {}
    for ename in ["tiled15", "binned", "tiled13", "tiled15", "binned"]:
        pg.tuning = ENGINES[ename]
        r, res = pg.run(tol=1e-7, max_iters=100, dangling="uniform")
        if ename in outs:
            assert np.array_equal(outs[ename][0], r), ename  # same engine, same bits
        outs[ename] = (r.copy(), res.iters)
    assert l1(outs["binned"][0], outs["tiled15"][0]) < 1e-6
    assert abs(outs["binned"][1] - outs["tiled15"][1]) <= 1


def test_auto_engine_choice():
    """Auto takes the tiled engine for R-MAT (long runs per tile) and the
    binned engine for a uniform graph whose entries rarely share a tile."""
    g = nwgr_amd.rmat(14, 16)
    pg = PageRankGraph(g, nwgr_amd.build_degrees(g))
    assert pg.tuning["engine"] == 2
    pg.run(tol=1e-6, max_iters=100, dangling="uniform")
    assert pg.last_engine() == "tiled"
    # 2^18 vertices in 8 tiles of 2^15 and ~4 entries per row: runs of ~1
    n = 1 << 18
    rng = np.random.default_rng(5)
    src = rng.integers(0, n, 4 * n)
    dst = rng.integers(0, n, 4 * n)
    g2 = nwgr_amd.build_adjacency(src, dst, n=n, dedup=True, drop_self_loops=True)
    deg2 = nwgr_amd.build_degrees(g2)
    pg2 = PageRankGraph(g2, deg2)
    pg2.tuning = {"tile_bits": 13}  # 32 tiles: ~1 entry per (row, tile) run
    parity(g2, deg2, "uniform", 1e-6, 100, device_graph=pg2)
    assert pg2.last_engine() == "binned"


@pytest.mark.parametrize("field,value", [("engine", 3), ("tile_bits", 12), ("tile_bits", 16),
                                         ("tile_fill_cost", 1 << 20)])
def test_bad_tuning_rejected_and_rolled_back(karate, field, value):
    d, g, deg = karate
    pg = PageRankGraph(g, deg)
    before = pg.tuning
    with pytest.raises(NwgrError) as e:
        pg.tuning = {field: value}
    assert e.value.code == NWGR_EINVAL
    assert pg.tuning == before
    r, _ = pg.run(tol=1e-7, max_iters=100)
    r64, _ = oracle.jacobi(g.offsets, g.indices, deg, 0.85, 1e-7, 100)
    assert l1(r, r64) < 1e-6


@pytest.mark.parametrize("ename", ["binned", "tiled15", "tiled13"])
def test_weighted_adjoint_per_engine(ename):
    """Weighted forward and adjoint gradients per engine against the FP64
    forward and adjoint of the oracle (relative L2 <= 1e-4)."""
    from nwgr_amd.autograd import weighted_pagerank
    say(f"weighted + adjoint, engine {ename}")
    g = nwgr_amd.rmat(13, 8)
    n, m = g.n, g.m
    rng = np.random.default_rng(7)
    w = (1.0 - rng.random(m)).astype(np.float32)
    t = rng.dirichlet(np.ones(n)).astype(np.float32)
    wl = rng.standard_normal(n).astype(np.float32)
    pg = PageRankGraph(g)
    pg.tuning = ENGINES[ename]
    dev = torch.device("cuda:0")
    wd = torch.from_numpy(w).to(dev).requires_grad_(True)
    td = torch.from_numpy(t).to(dev).requires_grad_(True)
    r = weighted_pagerank(pg, wd, td, tol=1e-7, max_iters=300, dangling="teleport", adj_tol=1e-7)
    assert pg.last_engine() == "binned" if ename == "binned" else pg.last_engine().startswith("tiled")
    (r * torch.from_numpy(wl).to(dev)).sum().backward()
    w64, t64, wl64 = w.astype(np.float64), t.astype(np.float64), wl.astype(np.float64)
    r64, _ = oracle.wpr_forward(g.offsets, g.indices, w64, t64, 0.85, 1e-15, 5000, oracle.TELEPORT, nthreads=8)
    _, dw64, dt64, _ = oracle.wpr_adjoint(g.offsets, g.indices, w64, t64, r64, wl64, 0.85, 1e-13, 5000,
                                          oracle.TELEPORT, nthreads=8)
    assert l1(r.detach().cpu().numpy(), r64) < 1e-5

    def rel2(a, b):
        return float(np.linalg.norm(np.asarray(a, np.float64) - b) / np.linalg.norm(b))

    assert rel2(wd.grad.cpu().numpy(), dw64) <= 1e-4
    assert rel2(td.grad.cpu().numpy(), dt64) <= 1e-4


def test_tile_pass_cut_leaves_results_bit_identical():
    """The tile pass's workgroup ranges (cut at setup from measured workgroup
    times, at this size) decide only which workgroup adds which chunk: every
    partial sum goes to the same slot.  Cuts made with other fill costs give
    bit-identical ranks, sweep counts and errors, and match the FP64 oracle."""
    say("tile pass cuts, R-MAT s21")
    g = nwgr_amd.rmat(21, 16)
    deg = nwgr_amd.build_degrees(g)
    pg = PageRankGraph(g, deg)
    assert pg.tuning["tile_fill_cost"] == 40
    base = {}
    for mode in ("uniform", "leak"):
        r, res = pg.run(tol=1e-6, max_iters=100, dangling=mode)
        assert pg.last_engine() == "tiled"
        base[mode] = (r, res.iters, res.error)
    for fill in (0, 1, 400, 40):
        pg.tuning = {"tile_fill_cost": fill}
        for mode in ("uniform", "leak"):
            r, res = pg.run(tol=1e-6, max_iters=100, dangling=mode)
            r0, it0, e0 = base[mode]
            assert res.iters == it0 and res.error == e0, (fill, mode)
            assert np.array_equal(r, r0), (fill, mode)
    r64, k64 = oracle.jacobi(g.offsets, g.indices, deg, 0.85, 1e-6, 100, dangling=oracle.UNIFORM, nthreads=4)
    assert abs(base["uniform"][1] - k64) <= 1
    assert l1(base["uniform"][0], r64) <= 1e-6
