"""The drop-in calls on the GPU: resident graphs kept between calls, the FP64
solve of Real = double, checks of the caller's vectors, and bounded memory
for any max_iters.

The reference's page_rank runs on a graph the caller built once
(include/nwgraph/algorithms/page_rank.hpp:92-119) and pr.exe repeats the
call over trials (bench/pr.cpp:94-95); its PageRank test computes in double
(test/page_rank_test.cpp:66,93)."""
import ctypes
import os
import shutil
import subprocess
import sys

import numpy as np
import pytest

import nwgr_amd
import oracle
from nwgr_amd import PageRankGraph
from nwgr_amd._lib import NWGR_EINVAL, NwgrError
from tests.conftest import ROOT, assert_same_top_k, l1

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture
def fresh_cache():
    """An empty cache of capacity 2; returns a function giving (hits, misses,
    entries) counted from the start of the test (the library's counters are
    cumulative)."""
    nwgr_amd.cache_clear()
    nwgr_amd.cache_set_capacity(2)
    base = nwgr_amd.cache_stats()

    def stats():
        s = nwgr_amd.cache_stats()
        return s["hits"] - base["hits"], s["misses"] - base["misses"], s["entries"]

    yield stats
    nwgr_amd.cache_clear()
    nwgr_amd.cache_set_capacity(2)


def test_repeated_calls_reuse_the_resident_graph(fresh_cache):
    say("dropin: repeated calls on rmat s14")
    g = nwgr_amd.rmat(14, 16)
    deg = nwgr_amd.build_degrees(g)
    r1 = np.zeros(g.n, np.float32)
    r2 = np.zeros(g.n, np.float32)
    nwgr_amd.page_rank(g, deg, r1, 0.85, 1e-6, 100)
    assert fresh_cache() == (0, 1, 1)
    pg, hit = PageRankGraph.cached(g, deg)
    assert hit
    setup = pg.setup_ms
    nwgr_amd.page_rank(g, deg, r2, 0.85, 1e-6, 100)
    assert fresh_cache() == (2, 1, 1)
    np.testing.assert_array_equal(r1, r2)  # bitwise: same layout, same sums
    pg2, hit = PageRankGraph.cached(g, deg)
    assert hit and pg2._h.value == pg._h.value and pg2.setup_ms == setup  # no second build
    r64, _ = oracle.jacobi(g.offsets, g.indices, deg, 0.85, 0.0, pg2.run(tol=1e-6, max_iters=100)[1].iters)
    assert l1(r1, r64) <= 1e-6


def test_changed_degrees_or_indices_invalidate(fresh_cache):
    say("dropin: in-place edits invalidate")
    g = nwgr_amd.rmat(12, 16)
    deg = nwgr_amd.build_degrees(g)
    r = np.zeros(g.n, np.float32)
    nwgr_amd.page_rank(g, deg, r, 0.85, 1e-6, 100)
    # the same degree array edited in place (one entry): a miss, and the
    # ranks are those of the new divisors
    v = int(np.argmax(deg))
    deg[v] += 1
    nwgr_amd.page_rank(g, deg, r, 0.85, 1e-6, 100)
    assert fresh_cache() == (0, 2, 1)
    fresh, _ = PageRankGraph(g, deg).run(tol=1e-6, max_iters=100, dangling="leak")
    np.testing.assert_array_equal(r, fresh)
    # an index edited at a compared position (the first entry): a miss
    u = int(g.indices[0])
    g.indices[0] = (u + 1) % g.n
    r_edit = np.zeros(g.n, np.float32)
    nwgr_amd.page_rank(g, deg, r_edit, 0.85, 1e-6, 100)
    assert fresh_cache()[1] == 3
    g.indices[0] = u
    # a different degree array (new address): a miss
    deg2 = deg.copy()
    nwgr_amd.page_rank(g, deg2, r, 0.85, 1e-6, 100)
    assert fresh_cache()[1:] == (4, 2)  # the capacity


def test_capacity_and_destroy(fresh_cache):
    graphs = [nwgr_amd.rmat(s, 8) for s in (9, 10, 11)]
    degs = [nwgr_amd.build_degrees(x) for x in graphs]
    for x, d in zip(graphs, degs):
        nwgr_amd.page_rank(x, d, np.zeros(x.n, np.float32), 0.85, 1e-4, 20)
    assert nwgr_amd.cache_stats()["entries"] == 2
    nwgr_amd.cache_set_capacity(1)
    assert nwgr_amd.cache_stats()["entries"] == 1
    # destroying a cached handle drops it from the cache
    pg, hit = PageRankGraph.cached(graphs[2], degs[2])
    assert hit
    from nwgr_amd._lib import pagerank_lib
    pagerank_lib().nwgr_pr_graph_destroy(pg._h)
    pg._h = ctypes.c_void_p()
    assert nwgr_amd.cache_stats()["entries"] == 0


def test_fp64_karate_golden_and_oracle(karate):
    """page_rank in double on karate, as the reference's test runs it
    (threshold 1e-7, max 100, 0.5 % of test/page_rank_test.cpp:72-75), and
    against the FP64 oracle at the same sweep."""
    d, g, deg = karate
    ranks = np.zeros(g.n, np.float64)
    nwgr_amd.page_rank(g, deg, ranks, d["damping"], d["threshold"], d["max_iters"])
    ans = np.array(d["answer"])
    assert np.all(np.abs(ranks - ans) <= d["rel_tolerance"] * ans)
    pg = PageRankGraph(g, deg)
    r, res = pg.run64(damping=0.85, tol=1e-7, max_iters=100, dangling="leak")
    np.testing.assert_array_equal(r, ranks)
    r64, k = oracle.jacobi(g.offsets, g.indices, deg, 0.85, 1e-7, 100)
    assert res.converged and res.iters == k
    assert l1(r, r64) <= 1e-13
    r_hist = pg.last_errors()
    _, _, hist = oracle.jacobi(g.offsets, g.indices, deg, 0.85, 1e-7, 100, return_hist=True)
    np.testing.assert_allclose(r_hist, hist, rtol=1e-9)


@pytest.mark.parametrize("mode", ["leak", "uniform", "teleport"])
def test_fp64_rmat_below_the_fp32_floor(mode):
    """tol 1e-12: below what FP32 ranks resolve at this size; the FP64 solve
    converges and matches the oracle's sweep count and ranks to 1e-12, where
    the FP32 solve of the same problem stays at its rounding noise."""
    say(f"fp64 rmat s14 {mode}")
    g = nwgr_amd.rmat(14, 16)
    deg = nwgr_amd.build_degrees(g)
    tel = np.random.default_rng(5).dirichlet(np.ones(g.n)).astype(np.float32) if mode == "teleport" else None
    pg = PageRankGraph(g, deg)
    r, res = pg.run64(tol=1e-12, max_iters=300, dangling=mode, teleport=tel)
    tel64 = None if tel is None else tel.astype(np.float64)
    mode_id = {"leak": oracle.LEAK, "uniform": oracle.UNIFORM, "teleport": oracle.TELEPORT}[mode]
    r64, k = oracle.jacobi(g.offsets, g.indices, deg, 0.85, 1e-12, 300, dangling=mode_id, teleport=tel64)
    assert res.converged and abs(res.iters - k) <= 1, (res.iters, k)
    r64m, _ = oracle.jacobi(g.offsets, g.indices, deg, 0.85, 0.0, res.iters, dangling=mode_id, teleport=tel64)
    assert l1(r, r64m) <= 1e-12
    assert_same_top_k(r, r64m, 100)
    # the FP32 iterates stop changing at a float fixed point (their error
    # can reach 0) whose distance to the FP64 ranks is rounding noise
    r32, _ = pg.run(tol=1e-12, max_iters=300, dangling=mode, teleport=tel)
    assert l1(r32, r64m) > 1e3 * max(l1(r, r64m), 1e-16)


def test_fp64_hub_rows_and_zero_sweeps():
    """Rows longer than the hub threshold are summed in chunks (last chunk
    adds them in order): a star of 20,000 leaves into one hub, plus a cycle."""
    n_leaf = 20000
    src = np.concatenate([np.arange(1, n_leaf + 1), np.arange(1, n_leaf + 1)]).astype(np.uint32)
    dst = np.concatenate([np.zeros(n_leaf), np.r_[np.arange(2, n_leaf + 1), 1]]).astype(np.uint32)
    g = nwgr_amd.build_adjacency(src, dst, n=n_leaf + 1)
    deg = nwgr_amd.build_degrees(g)
    assert int(np.diff(g.offsets).max()) == n_leaf
    pg = PageRankGraph(g, deg)
    r, res = pg.run64(tol=1e-13, max_iters=200, dangling="uniform")
    r64, _ = oracle.jacobi(g.offsets, g.indices, deg, 0.85, 0.0, res.iters, dangling=oracle.UNIFORM)
    # the hub's 20,000 terms are added in another order than the oracle's:
    # double rounding, about 1e-17 per vertex and sweep
    assert l1(r, r64) <= 1e-11
    r0, res0 = pg.run64(tol=1e-13, max_iters=0)
    assert res0.iters == 0 and np.all(r0 == 1.0 / g.n)


def test_cpp_dropin_double_and_trials(karate, tmp_path):
    """page_rank<csr_view, double> and pr.exe-style trials of the float call
    (bench/pr.cpp:92-97): every trial gives the same ranks."""
    d, g, deg = karate
    gxx = shutil.which("g++")
    if not gxx:
        pytest.skip("g++ missing")
    from nwgr_amd import _lib
    exe = tmp_path / "dropin_pr"
    subprocess.run([gxx, "-std=c++20", "-O1", "-I", os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "cpp", "dropin_pr.cpp"), "-o", str(exe),
                    "-L", os.path.dirname(_lib.PAGERANK_SO), "-lnwgr_pagerank",
                    "-Wl,-rpath," + os.path.dirname(_lib.PAGERANK_SO), "-Wl,--allow-shlib-undefined"],
                   check=True)
    gf = tmp_path / "karate.txt"
    with open(gf, "w") as f:
        f.write(f"{g.n} {g.m}\n" + " ".join(map(str, g.offsets)) + "\n" + " ".join(map(str, g.indices)) + "\n"
                + " ".join(map(str, deg)) + "\n")
    out = subprocess.run([str(exe), str(gf), "1e-7", "100", "0", "11", "f64"], capture_output=True, text=True,
                         timeout=120)
    assert out.returncode == 0, out.stderr
    r = np.array([float(x) for x in out.stdout.split()])
    r64, _ = oracle.jacobi(g.offsets, g.indices, deg, 0.85, 1e-7, 100)
    assert l1(r, r64) <= 1e-13
    ans = np.array(d["answer"])
    assert np.all(np.abs(r - ans) <= 0.005 * ans)
    out = subprocess.run([str(exe), str(gf), "1e-4", "20", "0", "11", "f32", "5"], capture_output=True, text=True,
                         timeout=120)
    assert out.returncode == 0, out.stderr
    assert out.stderr.count("trial ") == 5


@pytest.mark.parametrize("field", ["teleport", "init"])
@pytest.mark.parametrize("bad", [float("nan"), -0.5, float("inf")])
def test_bad_vectors_are_rejected(field, bad):
    import torch
    g = nwgr_amd.rmat(10, 8)
    deg = nwgr_amd.build_degrees(g)
    pg = PageRankGraph(g, deg)
    v = np.full(g.n, 1.0 / g.n, np.float32)
    v[123] = bad
    with pytest.raises(NwgrError) as e:
        pg.run(tol=1e-6, max_iters=50, dangling="uniform", **{field: v})
    assert e.value.code == NWGR_EINVAL and field in str(e.value)
    dv = torch.from_numpy(v).to("cuda:0")
    out = torch.empty(g.n, dtype=torch.float32, device="cuda:0")
    with pytest.raises(NwgrError) as e:
        pg.run_device(out.data_ptr(), tol=1e-6, max_iters=50, dangling="uniform", **{field + "_ptr": dv.data_ptr()})
    assert e.value.code == NWGR_EINVAL
    # the handle still solves afterwards
    v[123] = 1.0 / g.n
    r, res = pg.run(tol=1e-6, max_iters=50, dangling="uniform", **{field: v})
    assert res.converged


def test_unbounded_max_iters_needs_no_more_memory():
    """max_iters = 2^32 - 1 (the drop-in's clamp of SIZE_MAX) keeps O(1)
    device memory: the per-sweep history is bounded."""
    g = nwgr_amd.rmat(12, 16)
    deg = nwgr_amd.build_degrees(g)
    pg = PageRankGraph(g, deg)
    before = pg.device_bytes
    r, res = pg.run(tol=1e-6, max_iters=2**32 - 1, dangling="uniform")
    assert res.converged
    assert pg.device_bytes - before <= 2 * 4 * g.n + 4096  # the output staging only
    r2, res2 = pg.run(tol=1e-6, max_iters=100, dangling="uniform")
    np.testing.assert_array_equal(r, r2)
    assert len(pg.last_errors()) == res.iters
    # a solve longer than the history: chunks keep following the decay
    r3, res3 = pg.run(tol=0.0, max_iters=5000, dangling="uniform")
    assert res3.iters == 5000 and len(pg.last_errors()) == 4096


def test_empty_graph_on_the_device():
    g = nwgr_amd.CSRGraph(np.zeros(1, np.uint32), np.zeros(0, np.uint32))
    pg = PageRankGraph(g, np.zeros(0, np.uint32))
    r, res = pg.run(tol=1e-4, max_iters=20)
    assert len(r) == 0 and res.iters == 1
    r, res = pg.run64(tol=1e-4, max_iters=20)
    assert len(r) == 0 and res.iters == 1
