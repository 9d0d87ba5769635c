"""Batched personalized PageRank (k one-hot teleport columns in one solve) on
the GPU against the FP64 oracle (needs the GPU).

Per column j the batched solve is the plain Jacobi iteration with teleport =
e_{seed_j}; the gates are those of the plain path: L1 <= 1e-6 per column
against the oracle run for the same number of sweeps, sweep count within +-1
of the oracle's (the batch stops when its slowest column converges), and the
same top order per column.  At the full size (R-MAT s20, 256 columns) the
oracle is too slow for every column, so the test uses a size-independent
property instead: the iteration is linear in the teleport vector from a
fixed start, so after any number of sweeps a weighted combination of the
columns (weights summing to 1) equals the plain solve with the combined
teleport (LEAK and UNIFORM dangling), and 16 columns are checked against the
oracle.
"""
import numpy as np
import pytest

import nwgr_amd
import oracle
from nwgr_amd import PageRankGraph
from nwgr_amd._lib import NWGR_EINVAL, NwgrError
from tests.conftest import assert_same_top_k, l1
from tests.test_gpu_pagerank import MODES, say

pytestmark = pytest.mark.gpu


def onehot(n, v):
    t = np.zeros(n, np.float64)
    t[v] = 1.0
    return t


@pytest.mark.parametrize("mode", ["leak", "uniform", "teleport"])
@pytest.mark.parametrize("scale,k", [(11, 13), (13, 4)])
def test_ppr_columns_match_oracle(scale, k, mode):
    say(f"ppr s{scale} k={k} {mode}")
    g = nwgr_amd.rmat(scale, 16)
    deg = nwgr_amd.build_degrees(g)
    rng = np.random.default_rng(scale * 7 + k)
    seeds = rng.integers(0, g.n, k).astype(np.uint32)
    seeds[-1] = seeds[0]  # a repeated seed gives two equal columns
    pg = PageRankGraph(g, deg)
    R, res = pg.ppr_batch(seeds, damping=0.85, tol=1e-6, max_iters=200, dangling=mode)
    assert R.shape == (g.n, k)
    np.testing.assert_array_equal(R[:, 0], R[:, -1])
    k_ref = 0
    for j, s in enumerate(seeds):
        t = onehot(g.n, s)
        r64, _ = oracle.jacobi(g.offsets, g.indices, deg, 0.85, 0.0, res.iters, dangling=MODES[mode], teleport=t,
                               nthreads=8)
        _, kj = oracle.jacobi(g.offsets, g.indices, deg, 0.85, 1e-6, 200, dangling=MODES[mode], teleport=t,
                              nthreads=8)
        k_ref = max(k_ref, kj)
        err = l1(R[:, j], r64)
        assert err <= 1e-6, (j, err)
        assert_same_top_k(R[:, j], r64, 10)
    assert abs(res.iters - k_ref) <= 1, (res.iters, k_ref)
    cols = pg.ppr_column_errors()
    assert len(cols) == k and cols.max() == pytest.approx(res.error) and res.converged


def test_ppr_fixed_sweeps_and_single_column():
    g = nwgr_amd.rmat(10, 8)
    deg = nwgr_amd.build_degrees(g)
    pg = PageRankGraph(g, deg)
    R, res = pg.ppr_batch([5], damping=0.85, tol=0.0, max_iters=7, dangling="uniform")
    assert res.iters == 7 and not res.converged
    r64, _ = oracle.jacobi(g.offsets, g.indices, deg, 0.85, 0.0, 7, dangling=oracle.UNIFORM,
                           teleport=onehot(g.n, 5))
    assert l1(R[:, 0], r64) <= 1e-6
    # max_iters = 0 leaves the start vector
    R0, res0 = pg.ppr_batch([1, 2], max_iters=0)
    assert res0.iters == 0
    np.testing.assert_allclose(R0, 1.0 / g.n, rtol=1e-6)


def test_ppr_bad_arguments(karate):
    d, g, deg = karate
    pg = PageRankGraph(g, deg)
    for seeds in ([], list(range(257)), [g.n], [0, 1 << 31]):
        with pytest.raises(NwgrError) as e:
            pg.ppr_batch(np.asarray(seeds, np.int64).astype(np.uint32))
        assert e.value.code == NWGR_EINVAL
    R, res = pg.ppr_batch([0, 33], tol=1e-7, max_iters=100, dangling="leak")
    assert res.converged


def test_ppr_plain_solver_still_correct_after_batch(karate):
    """The batched solve shares the handle's state with the plain solve."""
    d, g, deg = karate
    pg = PageRankGraph(g, deg)
    pg.ppr_batch([3, 4, 5], tol=1e-6)
    r, _ = pg.run(tol=1e-7, max_iters=100, dangling="leak")
    r64, _ = oracle.jacobi(g.offsets, g.indices, deg, 0.85, 1e-7, 100)
    assert l1(r, r64) < 1e-6


@pytest.mark.slow
@pytest.mark.parametrize("mode", ["leak", "uniform"])
def test_config4a_rmat_s20_256_columns_linearity(mode):
    """R-MAT s20 ef16, 256 one-hot seeds, 30 sweeps: the mean of the columns
    and four random weighted combinations of them equal the plain solve with
    the combined teleport (which checks every column and its position); 16
    columns against the FP64 oracle; each column's mass under UNIFORM."""
    import torch
    say(f"ppr s20 k=256 {mode}")
    g = nwgr_amd.rmat(20, 16)
    deg = nwgr_amd.build_degrees(g)
    pg = PageRankGraph(g, deg)
    rng = np.random.default_rng(20)
    seeds = rng.integers(0, g.n, 256).astype(np.uint32)
    dev = torch.device("cuda:0")
    sd = torch.from_numpy(seeds.astype(np.int32)).to(dev)
    R = torch.empty((g.n, 256), dtype=torch.float32, device=dev)
    sweeps = 30
    res = pg.ppr_batch_device(sd.data_ptr(), 256, R.data_ptr(), tol=0.0, max_iters=sweeps, dangling=mode)
    torch.cuda.synchronize()
    assert res.iters == sweeps
    mean_cols = R.double().mean(dim=1).cpu().numpy()
    t = np.bincount(seeds, minlength=g.n).astype(np.float32) / 256.0
    r_mix, res_p = pg.run(damping=0.85, tol=0.0, max_iters=sweeps, dangling=mode, teleport=t)
    assert res_p.iters == sweeps
    err = l1(mean_cols, r_mix)
    say(f"  L1(mean of columns, mixed teleport) = {err:.2e}")
    assert err <= 1e-6
    # Every column's identity: for random positive weights a (summing to 1,
    # so the common start 1/n mixes to itself), sum_j a_j R[:, j] is the
    # plain solve with teleport sum_j a_j e_{seed_j}.  A permutation of the
    # columns, or one column wrong, changes the left side only.
    for trial in range(4):
        a = rng.dirichlet(np.ones(256))
        mixed = (R.double() @ torch.from_numpy(a).to(dev)).cpu().numpy()
        t_a = np.zeros(g.n, np.float64)
        np.add.at(t_a, seeds, a)
        r_a, res_a = pg.run(damping=0.85, tol=0.0, max_iters=sweeps, dangling=mode, teleport=t_a.astype(np.float32))
        assert res_a.iters == sweeps
        err = l1(mixed, r_a)
        say(f"  combination {trial}: L1 {err:.2e}")
        assert err <= 1e-6, (trial, err)
    Rh = R.cpu().numpy()
    if mode == "uniform":  # each column is a distribution
        mass = Rh.sum(axis=0, dtype=np.float64)
        assert np.abs(mass - 1.0).max() < 1e-5, np.abs(mass - 1.0).max()
    # columns at both ends, at the float4 lane boundaries and at random
    cols = sorted({0, 1, 2, 3, 4, 127, 128, 253, 254, 255} | set(rng.choice(256, 6, replace=False).tolist()))
    for j in cols:
        r64, _ = oracle.jacobi(g.offsets, g.indices, deg, 0.85, 0.0, sweeps, dangling=MODES[mode],
                               teleport=onehot(g.n, seeds[j]), nthreads=16)
        assert l1(Rh[:, j], r64) <= 1e-6, j
    # converged solve of the full batch
    res = pg.ppr_batch_device(sd.data_ptr(), 256, R.data_ptr(), tol=1e-6, max_iters=200, dangling=mode)
    assert res.converged and pg.ppr_column_errors().max() < 1e-6
    say(f"  converged in {res.iters} sweeps, {res.ms:.1f} ms")
