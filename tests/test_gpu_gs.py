"""Block Gauss-Seidel on the GPU (nwgr_pr_gs_run_f32) against the FP64 block
Gauss-Seidel oracle with the same row order and phase cuts.

The reference's default page_rank sweeps in place (v11,
include/nwgraph/algorithms/page_rank.hpp:124-139); the device mode makes that
deterministic with phases: Jacobi inside a phase, in place between phases.
Gates as for the Jacobi solve: FP32 ranks within L1 <= 1e-6 of the FP64
oracle run for the same number of sweeps, sweep count within +-1 of the
oracle's at the same tolerance, the same top order.
"""
import sys

import numpy as np
import pytest

import nwgr_amd
import oracle
from nwgr_amd import PageRankGraph
from nwgr_amd._lib import NWGR_EINVAL, NwgrError
from tests.conftest import assert_same_top_k, l1

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def gs_parity(pg, g, deg, phases, tol=1e-6, max_iters=200, teleport=None, init=None, topk=100, nthreads=16):
    r, res = pg.run_gs(phases, damping=0.85, tol=tol, max_iters=max_iters, teleport=teleport, init=init)
    order, cuts = pg.gs_phases(phases)
    assert sorted(order.tolist()) == list(range(g.n))
    tel = None if teleport is None else np.asarray(teleport, np.float64)
    ini = None if init is None else np.asarray(init, np.float64)
    r64, _ = oracle.bgs(g.offsets, g.indices, deg, 0.85, 0.0, res.iters, order=order, cuts=cuts, teleport=tel,
                        init=ini, nthreads=nthreads)
    _, k_ref = oracle.bgs(g.offsets, g.indices, deg, 0.85, tol, max_iters, order=order, cuts=cuts, teleport=tel,
                          init=ini, nthreads=nthreads)
    err = l1(r, r64)
    assert err <= 1e-6, f"L1 {err:.3e} at sweep {res.iters}"
    assert abs(int(res.iters) - int(k_ref)) <= 1, (res.iters, k_ref)
    assert_same_top_k(r, r64, topk)
    assert pg.last_engine() == "tiled-gauss-seidel"
    say(f"  phases={phases} sweeps={res.iters} (oracle {k_ref}) L1={err:.2e}")
    return r, res


def test_karate_phases(karate):
    """Karate: the golden vector of test/page_rank_test.cpp:72-75 with one
    phase (the Jacobi sweep), parity with 2, 4 and 8 phases at tol 1e-6 (at
    1e-7, karate's FP32 ranks sit at their rounding floor, where sweep counts
    scatter).  On this graph the phased sweeps take more sweeps than Jacobi
    (the FP64 oracle agrees: karate leaks no mass, so the Jacobi error of the
    uniform start has no component along the slowest mode, which the
    in-place sweeps bring back)."""
    d, g, deg = karate
    pg = PageRankGraph(g, deg)
    r, res = gs_parity(pg, g, deg, 1, tol=1e-7, max_iters=100, topk=34)
    ans = np.array(d["answer"])
    assert np.all(np.abs(r - ans) <= d["rel_tolerance"] * ans)
    _, res_j = pg.run(tol=1e-7, max_iters=100, dangling="leak")
    assert res.iters == res_j.iters
    for phases in (2, 4, 8):
        gs_parity(pg, g, deg, phases, tol=1e-6, max_iters=200, topk=34)


@pytest.mark.parametrize("scale", [10, 14, 16])
def test_rmat_phases(scale):
    say(f"gs rmat s{scale}")
    g = nwgr_amd.rmat(scale, 16)
    deg = nwgr_amd.build_degrees(g)
    pg = PageRankGraph(g, deg)
    counts = [gs_parity(pg, g, deg, P)[1].iters for P in (2, 8, 64)]
    assert counts == sorted(counts, reverse=True) and counts[-1] < counts[0], counts


def test_one_phase_is_the_jacobi_tiled_solve():
    """One phase holds every row: its layout is the Jacobi engine's, so the
    ranks are bit for bit those of the tiled Jacobi solve."""
    g = nwgr_amd.rmat(13, 16)
    deg = nwgr_amd.build_degrees(g)
    pg = PageRankGraph(g, deg)
    pg.tuning = {"engine": 1}
    rj, res_j = pg.run(tol=1e-6, max_iters=100, dangling="leak")
    rg, res_g = pg.run_gs(1, tol=1e-6, max_iters=100)
    assert res_g.iters == res_j.iters
    np.testing.assert_array_equal(rg, rj)


def test_teleport_init_and_cap():
    g = nwgr_amd.rmat(12, 16)
    deg = nwgr_amd.build_degrees(g)
    pg = PageRankGraph(g, deg)
    rng = np.random.default_rng(9)
    tel = rng.dirichlet(np.ones(g.n)).astype(np.float32)
    ini = rng.dirichlet(np.ones(g.n)).astype(np.float32)
    gs_parity(pg, g, deg, 4, teleport=tel, init=ini)
    r, res = pg.run_gs(4, tol=0.0, max_iters=7)
    assert res.iters == 7 and not res.converged
    order, cuts = pg.gs_phases(4)
    r64, _ = oracle.bgs(g.offsets, g.indices, deg, 0.85, 0.0, 7, order=order, cuts=cuts)
    assert l1(r, r64) <= 1e-6
    r0, res0 = pg.run_gs(4, tol=1e-6, max_iters=0)
    assert res0.iters == 0 and np.allclose(r0, 1.0 / g.n)


def test_tile_width_change_rebuilds_the_phases():
    g = nwgr_amd.rmat(14, 16)
    deg = nwgr_amd.build_degrees(g)
    pg = PageRankGraph(g, deg)
    gs_parity(pg, g, deg, 4)
    pg.tuning = {"tile_bits": 13}
    gs_parity(pg, g, deg, 4)
    # and the Jacobi solve after the phased one is unaffected
    r, res = pg.run(tol=1e-6, max_iters=100, dangling="leak")
    r64, _ = oracle.jacobi(g.offsets, g.indices, deg, 0.85, 0.0, res.iters)
    assert l1(r, r64) <= 1e-6


def test_bad_arguments():
    g = nwgr_amd.rmat(9, 8)
    deg = nwgr_amd.build_degrees(g)
    pg = PageRankGraph(g, deg)
    for kw in ({"phases": 0}, {"phases": 65}):
        with pytest.raises(NwgrError) as e:
            pg.run_gs(tol=1e-6, max_iters=20, **kw)
        assert e.value.code == NWGR_EINVAL
    from nwgr_amd._lib import PrParams, PrResult, pagerank_lib
    import ctypes
    p = PrParams()
    pagerank_lib().nwgr_pr_params_default(ctypes.byref(p))
    p.dangling = 1  # UNIFORM: not the reference's in-place sweep
    out = np.empty(g.n, np.float32)
    res = PrResult()
    assert pagerank_lib().nwgr_pr_gs_run_f32(pg._h, ctypes.byref(p), 4, out.ctypes.data,
                                             ctypes.byref(res)) == NWGR_EINVAL


@pytest.mark.slow
@pytest.mark.timeout(900)
def test_rmat_s22_phases_cut_the_sweeps():
    """R-MAT s22, LEAK, tol 1e-6: the phased solve takes clearly fewer sweeps
    than the 49 of Jacobi, with parity against the FP64 block oracle."""
    say("gs rmat s22 (generate)")
    g = nwgr_amd.rmat(22, 16)
    deg = nwgr_amd.build_degrees(g)
    pg = PageRankGraph(g, deg)
    _, res_j = pg.run(tol=1e-6, max_iters=100, dangling="leak")
    for phases in (4, 16):
        r, res = gs_parity(pg, g, deg, phases, tol=1e-6, max_iters=100)
        say(f"  s22 phases={phases}: {res.iters} sweeps against {res_j.iters} Jacobi, {res.ms:.2f} ms")
        assert res.iters < res_j.iters - 4


def test_big_rows_in_phase_layouts():
    """Phase layouts holding rows of more than 2048 partials (items of their
    own in the gather pass)."""
    from tests.test_gpu_pagerank import big_row_graph
    say("gs big rows")
    g = big_row_graph()
    deg = nwgr_amd.build_degrees(g)
    pg = PageRankGraph(g, deg)
    for phases in (2, 5):
        gs_parity(pg, g, deg, phases)
