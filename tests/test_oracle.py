"""Pins the CPU oracle (oracle/) before anything is checked against it.

Every check here is independent of the oracle's own formula: the reference's
known-answer vector, closed forms, dense linear solves, invariants of the
iteration, and finite differences.
"""
import numpy as np
import pytest

import oracle
from tests.conftest import l1

D = 0.85


def csr_from_edges(n, src, dst):
    import nwgr_amd
    return nwgr_amd.build_adjacency(np.asarray(src, np.uint32), np.asarray(dst, np.uint32), n=n)


def random_graph(n, m, seed, dangling_frac=0.2):
    rng = np.random.default_rng(seed)
    src = rng.integers(0, n, m)
    dst = rng.integers(0, n, m)
    keep = src != dst
    # make some vertices dangling (no out-edges)
    dang = rng.random(n) < dangling_frac
    keep &= ~dang[src]
    import nwgr_amd
    return nwgr_amd.build_adjacency(src[keep], dst[keep], n=n, dedup=True)


def dense_fixed_point(g, deg, mode, t=None, d=D):
    """Solve r = (1-d) t + d (P^T r + (1_dangling . r) u) exactly."""
    n = g.n
    A = np.zeros((n, n))
    for i in range(n):
        for j in g[i]:
            A[i, j] += 1.0 / deg[j]
    t = np.full(n, 1.0 / n) if t is None else np.asarray(t, np.float64)
    if mode == oracle.UNIFORM:
        u = np.full(n, 1.0 / n)
    elif mode == oracle.TELEPORT:
        u = t
    else:
        u = np.zeros(n)
    A += np.outer(u, (deg == 0).astype(np.float64))
    return np.linalg.solve(np.eye(n) - d * A, (1 - d) * t)


# --- reference known-answer test ------------------------------------------------

def test_karate_golden(karate):
    """page_rank_v1 on karate (test/page_rank_test.cpp:93-101): 0.5 % relative."""
    d, g, deg = karate
    r, it = oracle.jacobi(g.offsets, g.indices, deg, d["damping"], d["threshold"], d["max_iters"])
    ans = np.array(d["answer"])
    assert np.all(np.abs(r - ans) <= d["rel_tolerance"] * ans)
    assert it < d["max_iters"]


def test_karate_gs_same_fixed_point(karate):
    d, g, deg = karate
    rj, _ = oracle.jacobi(g.offsets, g.indices, deg, D, 1e-13, 1000)
    rg, _ = oracle.gs(g.offsets, g.indices, deg, D, 1e-13, 1000)
    assert l1(rj, rg) < 1e-11


def test_karate_iteration_counts_tol_1e4(karate):
    """Jacobi (v1/v10) and Gauss-Seidel (v11) reach tol 1e-4 in different numbers of sweeps."""
    _, g, deg = karate
    _, itj = oracle.jacobi(g.offsets, g.indices, deg, D, 1e-4, 100)
    _, itg = oracle.gs(g.offsets, g.indices, deg, D, 1e-4, 100)
    assert itg < itj


# --- closed forms ---------------------------------------------------------------

def test_edgeless_leak_and_uniform():
    n = 5
    g = csr_from_edges(n, [], [])
    deg = np.zeros(n, np.uint32)
    r, it, h = oracle.jacobi(g.offsets, g.indices, deg, D, 1e-12, 10, dangling=oracle.LEAK, return_hist=True)
    np.testing.assert_allclose(r, (1 - D) / n, rtol=1e-15)
    assert it == 2 and abs(h[0] - D) < 1e-15 and h[1] == 0.0
    r, it = oracle.jacobi(g.offsets, g.indices, deg, D, 1e-12, 10, dangling=oracle.UNIFORM)
    np.testing.assert_allclose(r, 1.0 / n, rtol=1e-15)
    assert it == 1


@pytest.mark.parametrize("kind", ["cycle", "complete"])
def test_regular_graphs_stay_uniform(kind):
    n = 7
    if kind == "cycle":
        src, dst = np.arange(n), (np.arange(n) + 1) % n
    else:
        src, dst = np.nonzero(~np.eye(n, dtype=bool))
    g = csr_from_edges(n, src, dst)
    deg = oracle.out_degree(n, g.indices)
    r, it = oracle.jacobi(g.offsets, g.indices, deg, D, 1e-12, 10)
    np.testing.assert_allclose(r, 1.0 / n, rtol=1e-14)
    assert it == 1
    rg, itg = oracle.gs(g.offsets, g.indices, deg, D, 1e-12, 10)
    np.testing.assert_allclose(rg, 1.0 / n, rtol=1e-14)


def test_star_leaves_to_hub():
    n = 9
    g = csr_from_edges(n, np.arange(1, n), np.zeros(n - 1))
    deg = oracle.out_degree(n, g.indices)
    r1, _ = oracle.jacobi(g.offsets, g.indices, deg, D, 0.0, 1)
    assert abs(r1[0] - ((1 - D) / n + D * (n - 1) / n)) < 1e-15
    np.testing.assert_allclose(r1[1:], (1 - D) / n, rtol=1e-15)
    r, _ = oracle.jacobi(g.offsets, g.indices, deg, D, 1e-15, 50)
    assert abs(r[0] - ((1 - D) / n + D * (n - 1) * (1 - D) / n)) < 1e-15


# --- dense solves, invariants ---------------------------------------------------

@pytest.mark.parametrize("mode", [oracle.LEAK, oracle.UNIFORM, oracle.TELEPORT])
def test_fixed_point_matches_dense_solve(mode):
    g = random_graph(60, 240, seed=5 + mode)
    deg = oracle.out_degree(g.n, g.indices)
    assert (deg == 0).any()
    t = np.random.default_rng(1).dirichlet(np.ones(g.n)) if mode == oracle.TELEPORT else None
    r, it = oracle.jacobi(g.offsets, g.indices, deg, D, 1e-15, 2000, dangling=mode, teleport=t)
    exact = dense_fixed_point(g, deg, mode, t)
    assert l1(r, exact) < 1e-12
    if mode != oracle.LEAK:
        assert abs(r.sum() - 1.0) < 1e-12


def test_planted_mistake_is_caught_by_dense_solve():
    """A dropped dangling term (UNIFORM solved as LEAK) must fail the check."""
    g = random_graph(60, 240, seed=9)
    deg = oracle.out_degree(g.n, g.indices)
    r_leak, _ = oracle.jacobi(g.offsets, g.indices, deg, D, 1e-15, 2000, dangling=oracle.LEAK)
    assert l1(r_leak, dense_fixed_point(g, deg, oracle.UNIFORM)) > 1e-3


@pytest.mark.parametrize("mode", [oracle.LEAK, oracle.UNIFORM])
def test_l1_contraction_and_mass(mode):
    g = random_graph(300, 2000, seed=11)
    deg = oracle.out_degree(g.n, g.indices)
    r, it, h = oracle.jacobi(g.offsets, g.indices, deg, D, 0.0, 40, dangling=mode, return_hist=True)
    big = h[:-1] > 1e-12  # below that the error is rounding noise
    assert big.sum() > 10
    assert np.all(h[1:][big] <= D * h[:-1][big] * (1 + 1e-9))


def test_parallel_jacobi_equals_serial():
    g = random_graph(5000, 40000, seed=3)
    deg = oracle.out_degree(g.n, g.indices)
    r1, it1 = oracle.jacobi(g.offsets, g.indices, deg, D, 1e-10, 100, dangling=oracle.UNIFORM)
    r8, it8 = oracle.jacobi(g.offsets, g.indices, deg, D, 1e-10, 100, dangling=oracle.UNIFORM, nthreads=8)
    assert it1 == it8
    assert l1(r1, r8) < 1e-13


def test_matched_iterate_semantics():
    """tol = 0 runs exactly max_iters sweeps; sweeps k and k+1 differ by err_hist[k]."""
    g = random_graph(200, 1500, seed=4)
    deg = oracle.out_degree(g.n, g.indices)
    r5, it5 = oracle.jacobi(g.offsets, g.indices, deg, D, 0.0, 5)
    r6, it6, h = oracle.jacobi(g.offsets, g.indices, deg, D, 0.0, 6, return_hist=True)
    assert it5 == 5 and it6 == 6
    assert abs(l1(r5, r6) - h[5]) < 1e-15


# --- the Gauss-Seidel restatements (CPU baseline) --------------------------------

@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_gs_parallel_one_thread_equals_serial(dtype):
    g = random_graph(3000, 20000, seed=21)
    deg = oracle.out_degree(g.n, g.indices)
    rs, its = oracle.gs(g.offsets, g.indices, deg, D, 1e-6, 100, dtype=dtype)
    rp, itp = oracle.gs_par(g.offsets, g.indices, deg, D, 1e-6, 100, nthreads=1, dtype=dtype, fast=False)
    assert its == itp
    np.testing.assert_array_equal(rs, rp)


def test_gs_parallel_many_threads_same_fixed_point():
    g = random_graph(20000, 200000, seed=22)
    deg = oracle.out_degree(g.n, g.indices)
    rj, _ = oracle.jacobi(g.offsets, g.indices, deg, D, 1e-13, 1000)
    rp, _ = oracle.gs_par(g.offsets, g.indices, deg, D, 1e-12, 1000, nthreads=8, dtype=np.float64)
    assert l1(rj, rp) < 1e-10
    r32, _ = oracle.jacobi_par(g.offsets, g.indices, deg, D, 1e-7, 1000, nthreads=8)
    assert l1(rj, r32) < 1e-5


# --- weighted PageRank and its adjoint ---------------------------------------------

def _weighted_case(seed, mode):
    g = random_graph(40, 200, seed=seed)
    rng = np.random.default_rng(seed)
    w = 1.0 - rng.random(g.m)                       # (0, 1]
    t = rng.dirichlet(np.ones(g.n))
    wl = rng.standard_normal(g.n)
    return g, w, t, wl


def test_weighted_with_unit_weights_equals_plain():
    g = random_graph(100, 700, seed=30)
    deg = oracle.out_degree(g.n, g.indices)
    t = np.full(g.n, 1.0 / g.n)
    rw, _ = oracle.wpr_forward(g.offsets, g.indices, np.ones(g.m), t, D, 1e-14, 2000, dangling=oracle.UNIFORM)
    rp, _ = oracle.jacobi(g.offsets, g.indices, deg, D, 1e-14, 2000, dangling=oracle.UNIFORM)
    assert l1(rw, rp) < 1e-13


@pytest.mark.parametrize("mode", [oracle.LEAK, oracle.UNIFORM, oracle.TELEPORT])
def test_adjoint_matches_central_differences(mode):
    g, w, t, wl = _weighted_case(31 + mode, mode)

    def loss(w_, t_):
        r, _ = oracle.wpr_forward(g.offsets, g.indices, w_, t_, D, 1e-15, 5000, dangling=mode)
        return float(wl @ r)

    r, _ = oracle.wpr_forward(g.offsets, g.indices, w, t, D, 1e-15, 5000, dangling=mode)
    lam, dw, dt, _ = oracle.wpr_adjoint(g.offsets, g.indices, w, t, r, wl, D, 1e-15, 5000, dangling=mode)
    h = 1e-6
    rng = np.random.default_rng(0)
    for k in rng.choice(g.m, 12, replace=False):
        wp, wm = w.copy(), w.copy()
        wp[k] += h
        wm[k] -= h
        fd = (loss(wp, t) - loss(wm, t)) / (2 * h)
        assert abs(fd - dw[k]) <= 1e-6 * max(1e-3, abs(fd)), (k, fd, dw[k])
    for v in rng.choice(g.n, 8, replace=False):
        tp, tm = t.copy(), t.copy()
        tp[v] += h
        tm[v] -= h
        fd = (loss(w, tp) - loss(w, tm)) / (2 * h)
        assert abs(fd - dt[v]) <= 1e-6 * max(1e-3, abs(fd)), (v, fd, dt[v])


def test_top_k_tie_break():
    r = np.array([0.1, 0.3, 0.3, 0.2, 0.3])
    assert list(oracle.top_k(r, 4)) == [1, 2, 4, 3]


@pytest.mark.parametrize("mode", [oracle.LEAK, oracle.UNIFORM, oracle.TELEPORT])
def test_parallel_weighted_equals_serial(mode):
    g, w, t, wl = _weighted_case(40 + mode, mode)
    r1, i1 = oracle.wpr_forward(g.offsets, g.indices, w, t, D, 1e-14, 3000, dangling=mode)
    r8, i8 = oracle.wpr_forward(g.offsets, g.indices, w, t, D, 1e-14, 3000, dangling=mode, nthreads=8)
    assert i1 == i8 and l1(r1, r8) < 1e-15
    a = oracle.wpr_adjoint(g.offsets, g.indices, w, t, r1, wl, D, 1e-13, 3000, dangling=mode)
    b = oracle.wpr_adjoint(g.offsets, g.indices, w, t, r1, wl, D, 1e-13, 3000, dangling=mode, nthreads=8)
    assert a[3] == b[3]
    np.testing.assert_allclose(a[1], b[1], rtol=0, atol=1e-15)
    np.testing.assert_allclose(a[2], b[2], rtol=0, atol=1e-15)


# --- residual ("delta") PageRank, page_rank_v13 / page_rank_v3 -------------------

def _path(L):
    """0 -> 1 -> ... -> L-1 (row i lists its in-neighbours)."""
    return csr_from_edges(L, np.arange(L - 1), np.arange(1, L))


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_delta_path_closed_form(dtype):
    """On a path every residual is (1-d) d^t after t hops and is pushed once:
    rank_j = (1-d) sum_{i<=j} d^i, iteration t sends from L - t vertices."""
    L = 7
    g = _path(L)
    deg = oracle.out_degree(L, g.indices)
    r, it, pushes = oracle.delta(g.offsets, g.indices, deg, D, 0.0, 50, dtype)
    want = np.array([(1 - D) * sum(D ** i for i in range(j + 1)) for j in range(L)])
    np.testing.assert_allclose(r, want, rtol=1e-6 if dtype == np.float32 else 1e-14)
    assert it == L
    np.testing.assert_array_equal(pushes, np.arange(L - 1, -1, -1))


def test_delta_threshold_and_iteration_cap():
    # 0 -> 1: the delta 0.1275 reaching vertex 1 is below the threshold 0.13,
    # so iteration 2 pushes nothing and the residual is dropped
    g = csr_from_edges(2, [0], [1])
    deg = oracle.out_degree(2, g.indices)
    r, it, pushes = oracle.delta(g.offsets, g.indices, deg, D, 0.13, 20)
    np.testing.assert_allclose(r, [0.15, 0.15], rtol=1e-15)
    assert it == 2 and list(pushes) == [1, 0]
    # threshold at 1 - d: nothing ever pushes
    r, it, pushes = oracle.delta(g.offsets, g.indices, deg, D, 1 - D, 20)
    assert it == 1 and list(pushes) == [0] and not r.any()
    # caps: one iteration pushes the initial residuals, none leaves zeros
    path = _path(5)
    pdeg = oracle.out_degree(5, path.indices)
    r, it, _ = oracle.delta(path.offsets, path.indices, pdeg, D, 0.0, 1)
    assert it == 1
    np.testing.assert_allclose(r, np.full(5, 1 - D), rtol=1e-15)
    r, it, pushes = oracle.delta(path.offsets, path.indices, pdeg, D, 0.0, 0)
    assert it == 0 and len(pushes) == 0 and not r.any()


def test_delta_converges_to_leak_fixed_point():
    """With a vanishing threshold the pushed mass sums the Neumann series:
    rank / n = (1-d)/n (I - d P)^-1 1, the LEAK fixed point."""
    g = random_graph(80, 400, seed=21)
    deg = oracle.out_degree(g.n, g.indices)
    r, it, pushes = oracle.delta(g.offsets, g.indices, deg, D, 1e-16, 1000)
    assert pushes[-1] == 0 and it < 1000
    want = dense_fixed_point(g, deg, oracle.LEAK)
    assert l1(r / g.n, want) < 1e-12
    # a plausible mistake, dangling mass spread uniformly, is far off
    assert l1(want, dense_fixed_point(g, deg, oracle.UNIFORM)) > 1e-3


@pytest.mark.parametrize("seed", [3, 4])
def test_delta_v3_keeps_nothing_that_pushes(seed):
    """v3 keeps a residual at or below the threshold until a delta replaces it;
    such a residual never pushes, so v3 and v13 give the same ranks."""
    g = random_graph(300, 1500, seed=seed)
    deg = oracle.out_degree(g.n, g.indices)
    for thr in (1e-4, 1e-3, 0.05):
        a = oracle.delta(g.offsets, g.indices, deg, D, thr, 60)
        b = oracle.delta(g.offsets, g.indices, deg, D, thr, 60, keep=True)
        np.testing.assert_array_equal(a[0], b[0])
        assert a[1] == b[1]
        np.testing.assert_array_equal(a[2], b[2])


def test_delta_float_restatement_close_to_double():
    g = random_graph(500, 4000, seed=5)
    deg = oracle.out_degree(g.n, g.indices)
    r64, it64, _ = oracle.delta(g.offsets, g.indices, deg, D, 1e-7, 100)
    r32, it32, _ = oracle.delta(g.offsets, g.indices, deg, D, 1e-7, 100, np.float32)
    assert abs(it64 - it32) <= 1
    assert l1(r32.astype(np.float64) / g.n, r64 / g.n) < 1e-6


# ---- block Gauss-Seidel (the pin of the device's phased in-place mode)


def _rmat_case(scale=10):
    import nwgr_amd
    g = nwgr_amd.rmat(scale, 8)
    return g, nwgr_amd.build_degrees(g)


def test_bgs_one_phase_is_jacobi():
    g, deg = _rmat_case()
    for tol, its in ((1e-6, 200), (0.0, 7)):
        r1, k1, h1 = oracle.bgs(g.offsets, g.indices, deg, 0.85, tol, its, return_hist=True)
        r2, k2, h2 = oracle.jacobi(g.offsets, g.indices, deg, 0.85, tol, its, return_hist=True)
        assert k1 == k2
        np.testing.assert_array_equal(r1, r2)
        np.testing.assert_array_equal(h1, h2)


def test_bgs_one_row_per_phase_is_serial_gauss_seidel():
    """Phases of one row each in the identity order: the serial in-place
    sweep of v11 (page_rank.hpp:124-139), same sweep count, ranks to double
    rounding ((1-d) t with t = 1/n against (1-d)/n)."""
    g, deg = _rmat_case()
    cuts = np.arange(g.n + 1, dtype=np.uint32)
    r1, k1, h1 = oracle.bgs(g.offsets, g.indices, deg, 0.85, 1e-9, 300, cuts=cuts, return_hist=True)
    r2, k2, h2 = oracle.gs(g.offsets, g.indices, deg, 0.85, 1e-9, 300, return_hist=True)
    assert k1 == k2
    np.testing.assert_allclose(r1, r2, rtol=1e-12, atol=1e-18)
    np.testing.assert_allclose(h1, h2, rtol=1e-9, atol=1e-18)


def test_bgs_fixed_point_and_fewer_sweeps_than_jacobi():
    """Any order and phase cut reaches the LEAK fixed point
    (I - d P^T) r = (1-d)/n (dense solve, which a dropped damping factor or
    a wrong gather index would miss); on R-MAT, more phases take no more
    sweeps to a tolerance, and 16 take fewer than Jacobi."""
    g, deg = _rmat_case(9)
    n = g.n
    P = np.zeros((n, n))
    for i in range(n):
        for k in range(g.offsets[i], g.offsets[i + 1]):
            j = g.indices[k]
            P[i, j] += 1.0 / deg[j]
    want = np.linalg.solve(np.eye(n) - 0.85 * P, np.full(n, 0.15 / n))
    rng = np.random.default_rng(1)
    order = rng.permutation(n).astype(np.uint32)
    counts = []
    for phases in (1, 2, 4, 16):
        cuts = np.linspace(0, n, phases + 1).round().astype(np.uint32)
        r, k = oracle.bgs(g.offsets, g.indices, deg, 0.85, 1e-13, 1000, order=order, cuts=cuts)
        assert np.abs(r - want).sum() < 1e-11, phases
        _, k6 = oracle.bgs(g.offsets, g.indices, deg, 0.85, 1e-6, 1000, order=order, cuts=cuts)
        counts.append(k6)
    assert counts[-1] < counts[0] and counts == sorted(counts, reverse=True), counts


def test_bgs_teleport_and_init():
    """A teleport vector and an initial iterate: the fixed point solves
    (I - d P^T) r = (1-d) t whatever the start."""
    g, deg = _rmat_case(8)
    n = g.n
    rng = np.random.default_rng(4)
    t = rng.dirichlet(np.ones(n))
    r0 = rng.dirichlet(np.ones(n))
    cuts = np.array([0, n // 3, n], np.uint32)
    r, k = oracle.bgs(g.offsets, g.indices, deg, 0.85, 1e-14, 2000, cuts=cuts, teleport=t, init=r0)
    rj, kj = oracle.jacobi(g.offsets, g.indices, deg, 0.85, 1e-14, 2000, teleport=t, init=r0)
    assert np.abs(r - rj).sum() < 1e-12 and k < kj


def test_bgs_threads_give_the_same_ranks():
    g, deg = _rmat_case(14)
    cuts = np.array([0, 1000, 5000, g.n], np.uint32)
    r1, k1 = oracle.bgs(g.offsets, g.indices, deg, 0.85, 1e-8, 200, cuts=cuts)
    r8, k8 = oracle.bgs(g.offsets, g.indices, deg, 0.85, 1e-8, 200, cuts=cuts, nthreads=8)
    assert k1 == k8
    np.testing.assert_array_equal(r1, r8)
