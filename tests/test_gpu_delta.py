"""Residual ("delta") PageRank on the GPU (page_rank_v13,
experimental/algorithms/page_rank.hpp:632-698) against the oracle's
restatement (needs the GPU).

Gates: the iteration count within +-1 of the FP64 oracle's at the same
threshold; ranks / n within L1 <= 1e-6 of the FP64 oracle run for the same
number of iterations; the pushes of every iteration within 1% (+2) of the
FP32 restatement's (the reference's operation order in float; only
residuals within rounding of the threshold can decide differently);
identical top-100 order.  Known answers (a path, the threshold and the
iteration cap) are checked exactly.
"""
import numpy as np
import pytest

import nwgr_amd
import oracle
from nwgr_amd import PageRankGraph
from nwgr_amd._lib import NwgrError
from tests.conftest import assert_same_top_k, l1
from tests.test_gpu_engines import ENGINES
from tests.test_gpu_pagerank import say

pytestmark = pytest.mark.gpu

D = 0.85


def delta_parity(g, deg, thr, max_iters, pg):
    r, res = pg.run_delta(damping=D, threshold=thr, max_iters=max_iters)
    pushes = pg.last_errors()
    assert len(pushes) == res.iters
    r64, k64, _ = oracle.delta(g.offsets, g.indices, deg, D, thr, max_iters)
    if k64 != res.iters:
        r64, _, _ = oracle.delta(g.offsets, g.indices, deg, D, thr, res.iters)
    _, k32, p32 = oracle.delta(g.offsets, g.indices, deg, D, thr, max_iters, np.float32)
    assert abs(int(res.iters) - int(k64)) <= 1, (res.iters, k64)
    n = g.n
    err = l1(r.astype(np.float64) / n, r64 / n)
    assert err <= 1e-6, f"L1 {err:.3e} after {res.iters} iterations"
    k = min(len(pushes), len(p32))
    assert np.all(np.abs(pushes[:k] - p32[:k]) <= 0.01 * p32[:k] + 2), (pushes[:k], p32[:k])
    assert res.converged == (pushes[-1] == 0)
    gap = assert_same_top_k(r, r64, 100)
    say(f"  iterations={res.iters} (oracle {k64}, f32 {k32}) L1={err:.2e} pushes={pushes[:6].astype(int).tolist()}"
        f" top-100 gap={gap:.2e}")
    return r, res


@pytest.mark.parametrize("ename", ["binned", "tiled15", "tiled13"])
@pytest.mark.parametrize("thr", [1e-4, 1e-6])
def test_delta_rmat_parity(ename, thr):
    say(f"delta rmat14 {ename} thr {thr}")
    g = nwgr_amd.rmat(14, 16)
    deg = nwgr_amd.build_degrees(g)
    pg = PageRankGraph(g, deg)
    pg.tuning = ENGINES[ename]
    delta_parity(g, deg, thr, 100, pg)
    assert pg.last_engine() == "binned" if ename == "binned" else pg.last_engine().startswith("tiled")


@pytest.mark.parametrize("ename", ["binned", "tiled15"])
def test_delta_path_known_answer(ename):
    L = 40
    g = nwgr_amd.build_adjacency(np.arange(L - 1, dtype=np.uint32), np.arange(1, L, dtype=np.uint32), n=L)
    deg = nwgr_amd.build_degrees(g)
    pg = PageRankGraph(g, deg)
    pg.tuning = ENGINES[ename]
    r, res = pg.run_delta(damping=D, threshold=0.0, max_iters=100)
    want = np.array([(1 - D) * sum(D ** i for i in range(j + 1)) for j in range(L)])
    np.testing.assert_allclose(r, want, rtol=2e-6)
    assert res.iters == L and res.converged
    np.testing.assert_array_equal(pg.last_errors(), np.arange(L - 1, -1, -1))
    # the float restatement does the same operations on a path: bit for bit
    r32, _, _ = oracle.delta(g.offsets, g.indices, deg, D, 0.0, 100, np.float32)
    np.testing.assert_array_equal(r, r32)


def test_delta_threshold_and_caps(karate):
    _, g, deg = karate
    pg = PageRankGraph(g, deg)
    r, res = pg.run_delta(threshold=1e-4, max_iters=0)
    assert res.iters == 0 and not r.any()
    r, res = pg.run_delta(threshold=1e-4, max_iters=1)
    assert res.iters == 1 and not res.converged
    np.testing.assert_array_equal(r, np.full(g.n, np.float32(1.0 - np.float64(np.float32(D))), np.float32))
    r, res = pg.run_delta(threshold=0.15, max_iters=20)  # above 1 - d: nothing pushes
    assert res.iters == 1 and res.converged and not r.any()
    for cap in (2, 3, 7):
        r, res = pg.run_delta(threshold=1e-9, max_iters=cap)
        r64, k64, _ = oracle.delta(g.offsets, g.indices, deg, D, 1e-9, cap)
        assert res.iters == k64 == cap
        assert l1(r / g.n, r64 / g.n) < 1e-7
    # the plain solve on the same handle is unaffected
    parity_r, pres = pg.run(tol=1e-7, max_iters=100)
    r64, _ = oracle.jacobi(g.offsets, g.indices, deg, D, 0.0, pres.iters)
    assert l1(parity_r, r64) < 1e-6


def test_delta_approaches_leak_pagerank():
    """A small threshold pushes nearly all mass: ranks / n -> the LEAK fixed point."""
    g = nwgr_amd.rmat(13, 8)
    deg = nwgr_amd.build_degrees(g)
    pg = PageRankGraph(g, deg)
    r, res = pg.run_delta(threshold=1e-12, max_iters=400)
    r_fix, _ = oracle.jacobi(g.offsets, g.indices, deg, D, 1e-15, 2000)
    assert l1(r.astype(np.float64) / g.n, r_fix) < 1e-6, res


def test_delta_rejects_unsupported_params(karate):
    _, g, deg = karate
    pg = PageRankGraph(g, deg)
    import ctypes
    from nwgr_amd._lib import PrResult, check, pagerank_lib
    from nwgr_amd.page_rank import _params
    p = _params(D, 1e-4, 20, "uniform", 0, False)  # only LEAK is a residual solve
    out = np.empty(g.n, np.float32)
    with pytest.raises(NwgrError):
        check(pagerank_lib().nwgr_pr_delta_run_f32(pg._h, ctypes.byref(p), out.ctypes.data, ctypes.byref(PrResult())))


def test_delta_dropin_function(karate):
    _, g, deg = karate
    ranks = np.zeros(g.size(), np.float32)
    assert nwgr_amd.page_rank_delta(g, deg, ranks, 0.85, 1e-5, 50, 4) is None
    r32, _, _ = oracle.delta(g.offsets, g.indices, deg, D, 1e-5, 50, np.float32)
    np.testing.assert_allclose(ranks, r32, rtol=1e-5)


@pytest.mark.parametrize("ename", ["binned", "tiled15"])
def test_delta_rmat_s20(ename):
    """R-MAT s20 ef16, the threshold of pr.exe (1e-4) and a tight one."""
    say(f"delta rmat20 {ename}")
    g = nwgr_amd.rmat(20, 16)
    deg = nwgr_amd.build_degrees(g)
    pg = PageRankGraph(g, deg)
    pg.tuning = ENGINES[ename]
    for thr in (1e-4, 1e-7):
        delta_parity(g, deg, thr, 100, pg)


def test_cpp_dropin_v13_and_v3(karate, tmp_path):
    """The C++ page_rank_v13 / page_rank_v3 templates (include/nwgr/page_rank.hpp)
    as pr.exe versions 13 and 3 call them (bench/pr.cpp:110,146)."""
    import os
    import shutil
    import subprocess
    from nwgr_amd import _lib
    from tests.conftest import ROOT
    _, g, deg = karate
    gxx = shutil.which("g++")
    if not gxx:
        pytest.skip("g++ missing")
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
    r32, _, _ = oracle.delta(g.offsets, g.indices, deg, D, 1e-5, 50, np.float32)
    outs = []
    for version in ("13", "3"):
        out = subprocess.run([str(exe), str(gf), "1e-5", "50", "0", version], capture_output=True, text=True,
                             timeout=120)
        assert out.returncode == 0, out.stderr
        r = np.array([float(x) for x in out.stdout.split()], np.float32)
        np.testing.assert_allclose(r, r32, rtol=1e-5)
        outs.append(r)
    np.testing.assert_array_equal(outs[0], outs[1])
