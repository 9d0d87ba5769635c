"""The accumulate form of the tiled engine's second pass (needs the GPU).

Rows are cut into blocks; a workgroup streams a block's partial sums tile by
tile into per-row double accumulators in LDS.  Gates as for the other
engines: FP32 ranks within L1 <= 1e-6 of the FP64 oracle at the same sweep
count, sweep count within +-1, identical top-100 order; the residual solve,
the weighted forward and the gradients through it; and reproducibility where
the double sums are exact.
"""
import numpy as np
import pytest

import nwgr_amd
import oracle
from nwgr_amd import PageRankGraph
from tests.conftest import l1
from tests.test_gpu_delta import delta_parity
from tests.test_gpu_engines import graphs
from tests.test_gpu_pagerank import parity, say

pytestmark = pytest.mark.gpu

ACC = {
    "acc15": dict(engine=1, tile_bits=15, tile_combine=1),
    "acc13": dict(engine=1, tile_bits=13, tile_fill_cost=0, tile_combine=1),
}


@pytest.mark.parametrize("gname", ["rmat16", "er17", "hub"])
@pytest.mark.parametrize("ename", list(ACC))
def test_accumulate_parity(gname, ename):
    say(f"accumulate {ename} on {gname}")
    g = graphs()[gname]()
    deg = nwgr_amd.build_degrees(g)
    pg = PageRankGraph(g, deg)
    pg.tuning = ACC[ename]
    for mode in ("uniform", "leak"):
        parity(g, deg, mode, 1e-6, 100, device_graph=pg)
        assert pg.last_engine() == "tiled-accumulate"


def test_accumulate_teleport_and_init():
    g = nwgr_amd.rmat(15, 16)
    deg = nwgr_amd.build_degrees(g)
    pg = PageRankGraph(g, deg)
    pg.tuning = ACC["acc15"]
    rng = np.random.default_rng(5)
    tel = rng.dirichlet(np.ones(g.n)).astype(np.float32)
    parity(g, deg, "teleport", 1e-6, 100, teleport=tel, device_graph=pg)
    init = rng.dirichlet(np.ones(g.n)).astype(np.float32)
    parity(g, deg, "uniform", 1e-6, 100, init=init, device_graph=pg)


def test_accumulate_low_locality_reproducible_and_agrees_with_gather():
    """On an Erdos-Renyi graph a row's partials are within a small factor of
    each other, so their double sum is exact: repeated solves are bitwise
    equal, and the ranks agree with the gathering pass to float rounding."""
    g = nwgr_amd.erdos_renyi(1 << 18, 16 << 18)
    deg = nwgr_amd.build_degrees(g)
    pg = PageRankGraph(g, deg)
    pg.tuning = dict(engine=1, tile_bits=13, tile_combine=1)
    r1, res1 = pg.run(tol=1e-6, max_iters=100, dangling="uniform")
    r2, res2 = pg.run(tol=1e-6, max_iters=100, dangling="uniform")
    assert res1.iters == res2.iters
    np.testing.assert_array_equal(r1, r2)
    pg.tuning = dict(tile_combine=0)
    r3, res3 = pg.run(tol=1e-6, max_iters=100, dangling="uniform")
    assert pg.last_engine() == "tiled"
    assert abs(res3.iters - res1.iters) <= 1
    assert l1(r1, r3) <= 1e-6


def test_auto_picks_accumulate_for_large_low_locality_graph():
    """Erdos-Renyi with n = 2^22 (past the auto tiled size, one entry per
    (row, tile) run on average): the auto choice accumulates."""
    g = nwgr_amd.erdos_renyi(1 << 22, 16 << 22)
    deg = nwgr_amd.build_degrees(g)
    pg = PageRankGraph(g, deg)
    r, res = pg.run(tol=1e-6, max_iters=100, dangling="uniform")
    assert pg.last_engine() == "tiled-accumulate"
    r64, _ = oracle.jacobi(g.offsets, g.indices, deg, 0.85, 0.0, res.iters, dangling=oracle.UNIFORM, nthreads=16)
    assert l1(r, r64) <= 1e-6
    say(f"  er 2^22 auto: sweeps={res.iters} {res.ms:.2f} ms")


@pytest.mark.parametrize("thr", [1e-4, 1e-6])
def test_accumulate_delta_parity(thr):
    g = nwgr_amd.rmat(14, 16)
    deg = nwgr_amd.build_degrees(g)
    pg = PageRankGraph(g, deg)
    pg.tuning = ACC["acc15"]
    delta_parity(g, deg, thr, 100, pg)
    assert pg.last_engine() == "tiled-accumulate"


def test_accumulate_weighted_gradients():
    """Weighted forward and adjoint (on the transpose) with the accumulate pass:
    gradients against the FP64 adjoint."""
    torch = pytest.importorskip("torch")
    from nwgr_amd.autograd import weighted_pagerank
    from tests.test_gpu_weighted import case, dev, rel2
    g, w, t, wl = case(13, seed=4)
    pg = PageRankGraph(g)
    pg.tuning = ACC["acc13"]
    wd = dev(w).requires_grad_(True)
    td = dev(t).requires_grad_(True)
    r = weighted_pagerank(pg, wd, td, tol=1e-7, max_iters=300, dangling="uniform", adj_tol=1e-7)
    assert pg.last_engine() == "tiled-accumulate"
    (r * dev(wl)).sum().backward()
    w64, t64 = w.astype(np.float64), t.astype(np.float64)
    r64, _ = oracle.wpr_forward(g.offsets, g.indices, w64, t64, 0.85, 1e-15, 5000, oracle.UNIFORM, nthreads=8)
    assert l1(r.detach().cpu().numpy(), r64) <= 1e-6
    _, dw64, dt64, _ = oracle.wpr_adjoint(g.offsets, g.indices, w64, t64, r64, wl.astype(np.float64), 0.85, 1e-13,
                                          5000, oracle.UNIFORM, nthreads=8)
    assert rel2(wd.grad.cpu().numpy(), dw64) <= 1e-4
    assert rel2(td.grad.cpu().numpy(), dt64) <= 1e-4
    torch.cuda.synchronize()
