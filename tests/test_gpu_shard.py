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


def share_reference(g, deg, rows, tol, max_iters):
    """FP64 sweeps of the rows `rows` alone (LEAK), the other rows' ranks held
    at 1/n: r_S' = (1-d)/n + d (A_S c), c = r / deg, stop on the L1 change of
    r_S.  Returns (r, sweeps)."""
    import scipy.sparse as sp
    n = g.n
    A = sp.csr_matrix((np.ones(g.m), g.indices, g.offsets.astype(np.int64)), shape=(n, n))
    As = A[rows]
    r = np.full(n, 1.0 / n)
    with np.errstate(divide="ignore"):
        c = r / deg.astype(np.float64)
    c[~np.isfinite(c)] = 0.0  # never gathered when the divisors are out-degrees
    for it in range(1, max_iters + 1):
        rn = 0.15 / n + 0.85 * (As @ c)
        err = np.abs(rn - r[rows]).sum()
        r[rows] = rn
        with np.errstate(divide="ignore"):
            c[rows] = rn / deg[rows]
        c[~np.isfinite(c)] = 0.0
        if err < tol:
            return r, it
    return r, max_iters


@pytest.mark.parametrize("shares", [3, 8])
def test_each_share_alone_on_its_own_layout(shares):
    """nwgr_pr_share_run_device_f32: a share's rows swept on the layout that
    holds only its entries, against the FP64 restatement of that sweep (which
    the dense solve of the share's fixed point pins at s10)."""
    import torch
    g = nwgr_amd.rmat(14, 16)
    deg = nwgr_amd.build_degrees(g)
    pg = PageRankGraph(g, deg)
    order, cuts = pg.gs_phases(shares)
    # the shares are the row partition a multi-GPU solve takes: about equal
    # sweep cost (8 per entry, 2.9 per fragment -- about min(entries, tiles
    # holding sources) + entries // 256 of them -- 80 per row with entries, 1
    # per row without that is a source, 0 per isolated vertex)
    counts = np.diff(g.offsets.astype(np.int64))[order]
    source = (deg[order] > 0)
    n_act = int(np.count_nonzero((counts > 0) | source))
    assert np.all(((counts > 0) | source)[:n_act])  # the isolated vertices are last
    tiles = -(-max(n_act, 1) // 2 ** 15)
    frags = np.minimum(counts, tiles) + counts // 256
    cost = np.where(counts > 0, 8.0 * counts + 2.9 * frags + 80.0, np.where(source, 1.0, 0.0))
    per = np.add.reduceat(cost, cuts[:-1].astype(np.int64))
    assert per.max() <= cost.sum() / shares + cost.max(), (per, cost.sum() / shares)
    out = torch.empty(g.n, dtype=torch.float32, device="cuda:0")
    for q in range(shares):
        rows = order[cuts[q]:cuts[q + 1]].astype(np.int64)
        res = pg.run_share_device(out.data_ptr(), shares, q, tol=1e-6, max_iters=200)
        r = out.cpu().numpy()
        r_ref, k_ref = share_reference(g, deg, rows, 1e-6, 200)
        r_m, _ = share_reference(g, deg, rows, 0.0, res.iters)
        assert abs(res.iters - k_ref) <= 1, (q, res.iters, k_ref)
        assert l1(r[rows], r_m[rows]) <= 1e-6, q
        others = np.setdiff1d(np.arange(g.n), rows)
        np.testing.assert_allclose(r[others], 1.0 / g.n, rtol=1e-6)


def test_share_reference_fixed_point_dense():
    """The restatement above converges to the dense solution of the share's
    fixed point (I - d A_SS D_S^-1) r_S = (1-d)/n + d A_{S,rest} c_rest."""
    g = nwgr_amd.rmat(10, 8)
    deg = nwgr_amd.build_degrees(g).astype(np.float64)
    n = g.n
    A = np.zeros((n, n))
    for i in range(n):
        for k in range(g.offsets[i], g.offsets[i + 1]):
            A[i, g.indices[k]] += 1.0
    rows = np.arange(100, 700)
    rest = np.setdiff1d(np.arange(n), rows)
    dinv = np.where(deg > 0, 1.0 / np.maximum(deg, 1), 0.0)
    M = np.eye(len(rows)) - 0.85 * A[np.ix_(rows, rows)] * dinv[rows]
    rhs = 0.15 / n + 0.85 * A[np.ix_(rows, rest)] @ (dinv[rest] / n)
    want = np.linalg.solve(M, rhs)
    r, _ = share_reference(g, deg.astype(np.uint32), rows, 1e-15, 5000)
    assert np.abs(r[rows] - want).sum() < 1e-12


def test_balanced_shares_keep_parity():
    """nwgr_pr_share_balance: the rows re-cut by measured share times.  The
    cuts stay a partition of the relabelled rows; every share swept alone
    matches the FP64 restatement of its rows' sweeps, and the block
    Gauss-Seidel solve with phases = shares follows the new cuts (the FP64
    block Gauss-Seidel restatement on those cuts pins it)."""
    import torch
    from tests.test_gpu_gs import gs_parity
    say("balanced shares, rmat s15")
    g = nwgr_amd.rmat(15, 16)
    deg = nwgr_amd.build_degrees(g)
    pg = PageRankGraph(g, deg)
    _, model_cuts = pg.gs_phases(4)
    us = pg.share_balance(4, rounds=2)
    assert us > 0
    order, cuts = pg.gs_phases(4)
    assert cuts[0] == 0 and cuts[-1] == g.n and np.all(np.diff(cuts.astype(np.int64)) >= 0)
    say(f"  model cuts {model_cuts.tolist()} balanced {cuts.tolist()} slowest {us:.1f} us")
    out = torch.empty(g.n, dtype=torch.float32, device="cuda:0")
    for q in range(4):
        rows = order[cuts[q]:cuts[q + 1]].astype(np.int64)
        res = pg.run_share_device(out.data_ptr(), 4, q, tol=1e-6, max_iters=200)
        r = out.cpu().numpy()
        r_m, _ = share_reference(g, deg, rows, 0.0, res.iters)
        assert l1(r[rows], r_m[rows]) <= 1e-6, q
    gs_parity(pg, g, deg, 4)


def test_balance_on_a_tiny_graph(karate):
    """The measured balance on karate (34 rows, 8 shares: some shares hold
    a row or two): the cuts stay a partition, zero rounds keeps the model's
    cuts, and the phased solve on the balanced cuts keeps parity with the
    block Gauss-Seidel restatement on those cuts."""
    from tests.test_gpu_gs import gs_parity
    _, g, deg = karate
    pg = PageRankGraph(g, deg)
    _, model_cuts = pg.gs_phases(8)
    assert pg.share_balance(8, rounds=0) > 0
    np.testing.assert_array_equal(pg.gs_phases(8)[1], model_cuts)
    pg.share_balance(8, rounds=1)
    _, cuts = pg.gs_phases(8)
    assert cuts[0] == 0 and cuts[-1] == g.n and np.all(np.diff(cuts.astype(np.int64)) >= 0)
    gs_parity(pg, g, deg, 8, topk=34)
    with pytest.raises(nwgr_amd.NwgrError):
        pg.share_balance(8, rounds=9)
