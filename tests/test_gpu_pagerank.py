"""Parity of the HIP PageRank path against the FP64 oracle (needs the GPU).

Gates (per workload): FP32 device ranks within L1 <= 1e-6 of the FP64 Jacobi
oracle run for the SAME number of sweeps; sweep count within +-1 of the
oracle's at the same tolerance; identical top-100 order (vertex-id
tie-break).
"""
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

MODES = {"leak": oracle.LEAK, "uniform": oracle.UNIFORM, "teleport": oracle.TELEPORT}
L1_GATE = 1e-6


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def parity(g, deg, mode="uniform", tol=1e-6, max_iters=100, teleport=None, init=None, topk=100,
           nthreads=16, device_graph=None):
    pg = device_graph or PageRankGraph(g, deg)
    r, res = pg.run(damping=0.85, tol=tol, max_iters=max_iters, dangling=mode, teleport=teleport, init=init)
    tel64 = None if teleport is None else np.asarray(teleport, np.float64)
    ini64 = None if init is None else np.asarray(init, np.float64)
    r64, k64 = oracle.jacobi(g.offsets, g.indices, deg, 0.85, 0.0, res.iters, dangling=MODES[mode],
                             teleport=tel64, init=ini64, nthreads=nthreads)
    _, k_ref = oracle.jacobi(g.offsets, g.indices, deg, 0.85, tol, max_iters, dangling=MODES[mode],
                             teleport=tel64, init=ini64, nthreads=nthreads)
    err = l1(r, r64)
    assert err <= L1_GATE, f"L1 {err:.3e} at sweep {res.iters}"
    assert abs(int(res.iters) - int(k_ref)) <= 1, (res.iters, k_ref)
    gap = assert_same_top_k(r, r64, topk)
    print(f"  sweeps={res.iters} (oracle {k_ref}) L1={err:.2e} min top-{min(topk, g.n)} gap={gap:.2e}", flush=True)
    return r, res, r64, err


def test_karate_golden_and_oracle(karate):
    d, g, deg = karate
    say("karate")
    pg = PageRankGraph(g, deg)
    r, res = pg.run(damping=d["damping"], tol=d["threshold"], max_iters=d["max_iters"], dangling="leak")
    ans = np.array(d["answer"])
    assert np.all(np.abs(r - ans) <= d["rel_tolerance"] * ans)
    parity(g, deg, "leak", d["threshold"], d["max_iters"], topk=34)


def test_dropin_function_fills_caller_array(karate):
    d, g, deg = karate
    ranks = np.zeros(g.size(), np.float32)
    assert nwgr_amd.page_rank(g, deg, ranks, 0.85, 1e-7, 100, 8) is None
    r64, _ = oracle.jacobi(g.offsets, g.indices, deg, 0.85, 1e-7, 100)
    assert l1(ranks, r64) < 1e-6


@pytest.mark.parametrize("mode", ["leak", "uniform", "teleport"])
@pytest.mark.parametrize("scale", [10, 14])
def test_rmat_small_parity(scale, mode):
    say(f"rmat s{scale} {mode}")
    g = nwgr_amd.rmat(scale, 16)
    deg = nwgr_amd.build_degrees(g)
    tel = np.random.default_rng(scale).dirichlet(np.ones(g.n)).astype(np.float32) if mode == "teleport" else None
    parity(g, deg, mode, 1e-6, 100, teleport=tel)


def test_config1_rmat_s16_tol_1e4():
    """R-MAT s16 ef16, uniform teleport, dangling UNIFORM, tol 1e-4, max 20."""
    say("rmat s16 config1")
    g = nwgr_amd.rmat(16, 16)
    deg = nwgr_amd.build_degrees(g)
    parity(g, deg, "uniform", 1e-4, 20)


def test_er_parity():
    say("er 2^16")
    g = nwgr_amd.erdos_renyi(1 << 16, 16 << 16)
    deg = nwgr_amd.build_degrees(g)
    parity(g, deg, "uniform", 1e-6, 100)


def test_hub_rows_split_in_chunks():
    """In-degree far above the hub threshold: chunked hub rows combine exactly."""
    say("hubs")
    n = 200_000
    rng = np.random.default_rng(3)
    # two hubs with 150k and 60k in-edges plus a random background
    src = np.concatenate([np.arange(1, 150_001), np.arange(100, 60_100), rng.integers(0, n, 400_000)])
    dst = np.concatenate([np.zeros(150_000, np.int64), np.ones(60_000, np.int64), rng.integers(0, n, 400_000)])
    g = nwgr_amd.build_adjacency(src, dst, n=n, dedup=True, drop_self_loops=True)
    deg = nwgr_amd.build_degrees(g)
    assert np.diff(g.offsets).max() > 100_000
    parity(g, deg, "uniform", 1e-6, 100)
    parity(g, deg, "leak", 1e-6, 100)


def big_row_graph():
    """Two rows whose partials (one per 256-entry run inside a source tile)
    pass 2048 at any tile width: 650,000 and 560,000 in-neighbours, plus a
    random background."""
    n = 700_000
    rng = np.random.default_rng(5)
    src = np.concatenate([np.arange(1, 650_001), np.arange(50_000, 610_000), rng.integers(0, n, 1_000_000)])
    dst = np.concatenate([np.zeros(650_000, np.int64), np.ones(560_000, np.int64), rng.integers(0, n, 1_000_000)])
    g = nwgr_amd.build_adjacency(src, dst, n=n, dedup=True, drop_self_loops=True)
    assert np.sort(np.diff(g.offsets))[-2] > 2048 * 256
    return g


@pytest.mark.parametrize("tuning", [{}, dict(engine=1, tile_bits=13, tile_combine=0)])
def test_big_rows_added_by_the_whole_workgroup(tuning):
    """Rows of more than 2048 partials are items of their own in the gather
    pass (every thread adds a strided share, wavefronts in order)."""
    say(f"big rows {tuning}")
    g = big_row_graph()
    deg = nwgr_amd.build_degrees(g)
    pg = PageRankGraph(g, deg)
    pg.tuning = tuning
    parity(g, deg, "uniform", 1e-6, 100, device_graph=pg)
    parity(g, deg, "leak", 1e-6, 100, device_graph=pg)
    tel = np.random.default_rng(6).random(g.n).astype(np.float32)
    tel /= tel.sum()
    parity(g, deg, "teleport", 1e-6, 100, teleport=tel, device_graph=pg)
    assert pg.last_engine().startswith("tiled")


@pytest.mark.parametrize("tuning", [{}, dict(engine=0), dict(engine=1, tile_combine=0),
                                    dict(engine=1, tile_combine=1)])
@pytest.mark.parametrize("n", [1, 2, 5, 70000])
def test_edgeless_graphs(n, tuning):
    g = nwgr_amd.build_adjacency([], [], n=n)
    deg = np.zeros(n, np.uint32)
    for mode, want, sweeps in (("leak", 0.15 / n, 2), ("uniform", 1.0 / n, 1)):
        pg = PageRankGraph(g, deg)
        pg.tuning = tuning
        r, res = pg.run(tol=1e-9, max_iters=10, dangling=mode)
        np.testing.assert_allclose(r, want, rtol=1e-6)
        assert res.iters == sweeps
    # the residual solve: every vertex pushes its 1 - d once, nobody sends
    pg = PageRankGraph(g, deg)
    pg.tuning = tuning
    r, res = pg.run_delta(threshold=1e-4, max_iters=10)
    assert res.iters == 1 and res.converged
    np.testing.assert_array_equal(r, np.float32(1.0 - np.float64(np.float32(0.85))))


def test_iteration_cap_and_zero_sweeps():
    g = nwgr_amd.rmat(10, 8)
    deg = nwgr_amd.build_degrees(g)
    pg = PageRankGraph(g, deg)
    r, res = pg.run(tol=0.0, max_iters=3, dangling="uniform")
    assert res.iters == 3 and not res.converged
    r64, _ = oracle.jacobi(g.offsets, g.indices, deg, 0.85, 0.0, 3, dangling=oracle.UNIFORM)
    assert l1(r, r64) < 1e-7
    r0, res0 = pg.run(tol=1e-6, max_iters=0)
    assert res0.iters == 0
    np.testing.assert_allclose(r0, 1.0 / g.n, rtol=1e-7)
    for cap in (1, 2, 15, 16, 17, 33):  # around the captured chunk length
        _, rc = pg.run(tol=0.0, max_iters=cap, chunk=16)
        assert rc.iters == cap


def test_per_sweep_errors_match_oracle():
    g = nwgr_amd.rmat(12, 16)
    deg = nwgr_amd.build_degrees(g)
    pg = PageRankGraph(g, deg)
    _, res = pg.run(tol=1e-6, max_iters=60, dangling="uniform")
    errs = pg.last_errors()
    _, k, h = oracle.jacobi(g.offsets, g.indices, deg, 0.85, 1e-6, 60, dangling=oracle.UNIFORM, return_hist=True)
    assert len(errs) == res.iters and abs(len(h) - res.iters) <= 1
    big = h[: min(len(errs), len(h))] > 1e-5   # well above the FP32 noise floor of the L1 change
    np.testing.assert_allclose(errs[: len(big)][big], h[: len(big)][big], rtol=1e-3)


def test_init_vector_and_warm_start():
    g = nwgr_amd.rmat(12, 16)
    deg = nwgr_amd.build_degrees(g)
    pg = PageRankGraph(g, deg)
    r_conv, res1 = pg.run(tol=1e-7, max_iters=100, dangling="uniform")
    init = np.random.default_rng(0).dirichlet(np.ones(g.n)).astype(np.float32)
    parity(g, deg, "uniform", 1e-6, 100, init=init, device_graph=pg)
    _, res2 = pg.run(tol=1e-6, max_iters=100, dangling="uniform", init=r_conv)
    assert res2.iters <= 2


def test_deterministic_bits():
    g = nwgr_amd.rmat(14, 16)
    deg = nwgr_amd.build_degrees(g)
    pg = PageRankGraph(g, deg)
    a, ra = pg.run(tol=1e-6, max_iters=100, dangling="uniform")
    b, rb = pg.run(tol=1e-6, max_iters=100, dangling="uniform")
    np.testing.assert_array_equal(a, b)
    assert ra.iters == rb.iters and ra.error == rb.error


def test_caller_degrees_used_as_given(karate):
    """The reference divides by whatever degrees it is given (page_rank.hpp:116,135)."""
    d, g, deg = karate
    deg2 = deg * 2
    r, res = PageRankGraph(g, deg2).run(tol=1e-9, max_iters=200, dangling="leak")
    r64, _ = oracle.jacobi(g.offsets, g.indices, deg2, 0.85, 0.0, res.iters)
    assert l1(r, r64) < 1e-6


def test_null_degrees_means_out_degree(karate):
    d, g, deg = karate
    a, _ = PageRankGraph(g, None).run(tol=1e-7, max_iters=100)
    b, _ = PageRankGraph(g, deg).run(tol=1e-7, max_iters=100)
    np.testing.assert_array_equal(a, b)


def test_malformed_csr_rejected():
    bad = nwgr_amd.CSRGraph(np.array([0, 2, 1, 3], np.uint32), np.array([0, 1, 2], np.uint32))
    with pytest.raises(NwgrError) as e:
        PageRankGraph(bad)
    assert e.value.code == NWGR_EINVAL
    bad = nwgr_amd.CSRGraph(np.array([0, 1, 2], np.uint32), np.array([0, 7], np.uint32))
    with pytest.raises(NwgrError) as e:
        PageRankGraph(bad)
    assert e.value.code == NWGR_EINVAL


def test_bad_params_rejected(karate):
    _, g, deg = karate
    pg = PageRankGraph(g, deg)
    with pytest.raises(NwgrError):
        pg.run(damping=1.5)
    with pytest.raises(NwgrError):
        pg.run(dangling=2)  # TELEPORT without a teleport vector


def test_device_pointers_from_torch():
    import torch
    g = nwgr_amd.rmat(12, 16)
    deg = nwgr_amd.build_degrees(g)
    dev = torch.device("cuda", 0)
    off = torch.from_numpy(g.offsets.view(np.int32)).to(dev)
    ind = torch.from_numpy(g.indices.view(np.int32)).to(dev)
    dg = torch.from_numpy(deg.view(np.int32)).to(dev)
    stream = torch.cuda.current_stream().cuda_stream
    pg = PageRankGraph.from_device(off.data_ptr(), ind.data_ptr(), g.n, g.m, dg.data_ptr(), 0, stream)
    out = torch.empty(g.n, dtype=torch.float32, device=dev)
    tel = torch.full((g.n,), 1.0 / g.n, dtype=torch.float32, device=dev)
    res = pg.run_device(out.data_ptr(), tol=1e-6, max_iters=100, dangling="teleport",
                        teleport_ptr=tel.data_ptr(), stream=stream)
    r = out.cpu().numpy()
    r_host, res_host = PageRankGraph(g, deg).run(tol=1e-6, max_iters=100, dangling="uniform")
    np.testing.assert_array_equal(r, r_host)  # TELEPORT with t = 1/n is UNIFORM
    assert res.iters == res_host.iters


def test_cpp_dropin_runs(karate, tmp_path):
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
    out = subprocess.run([str(exe), str(gf), "1e-7", "100", "10"], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    head, tail = out.stdout.split("Verifying\n")
    r = np.array([float(x) for x in head.split()])
    ans = np.array(d["answer"])
    assert np.all(np.abs(r - ans) <= 0.005 * ans)
    # print_n_ranks: the 10 best as "id: rank", by rank then id
    lines = tail.strip().splitlines()
    ids = [int(x.split(":")[0]) for x in lines]
    want = sorted(range(g.n), key=lambda v: (-np.float32(r[v]), v))[:10]
    assert ids == want, (ids, want)
    assert all(abs(float(x.split(":")[1]) - r[v]) <= 1e-6 for x, v in zip(lines, ids))


@pytest.fixture(scope="module")
def rmat22():
    say("rmat s22 (generate)")
    g = nwgr_amd.rmat(22, 16)
    return g, nwgr_amd.build_degrees(g)


@pytest.mark.slow
@pytest.mark.parametrize("mode", ["uniform", "leak"])
def test_config2_rmat_s22_full_size(rmat22, mode):
    """R-MAT s22 ef16, FP32, tol 1e-6, max 100, vs FP64 at the matched sweep:
    UNIFORM dangling (the headline solve) and LEAK (the reference's, the
    solve the time-to-tolerance comparison with the host path uses)."""
    g, deg = rmat22
    say(f"rmat s22 {mode} n={g.n} m={g.m} (solve + oracle)")
    r, res, r64, err = parity(g, deg, mode, 1e-6, 100, nthreads=16)
    say(f"rmat s22 {mode} sweeps={res.iters} L1={err:.3e} ms={res.ms:.2f}")
    if mode == "uniform":
        assert abs(float(r.sum(dtype=np.float64)) - 1.0) < 1e-5
    else:  # the leaked mass is the FP64 iterate's
        assert abs(float(r.sum(dtype=np.float64)) - float(r64.sum())) < 1e-5


@pytest.fixture(scope="module")
def er25():
    say("er 2^25 (generate)")
    g = nwgr_amd.erdos_renyi(1 << 25, 16 << 25)
    return g, nwgr_amd.build_degrees(g)


@pytest.mark.slow
@pytest.mark.timeout(900)
@pytest.mark.parametrize("mode", ["uniform", "leak"])
def test_config3_er_2pow25_full_size(er25, mode):
    """Erdos-Renyi G(n = 2^25, m = 16n drawn with replacement, deduplicated,
    self-loops dropped), FP32, tol 1e-6, max 100, vs FP64 at the matched
    sweep; mass conservation under UNIFORM dangling, the FP64 iterate's
    mass under LEAK."""
    g, deg = er25
    say(f"er 2^25 {mode} n={g.n} m={g.m} (solve + oracle)")
    r, res, r64, err = parity(g, deg, mode, 1e-6, 100, nthreads=16)
    say(f"er 2^25 {mode} sweeps={res.iters} L1={err:.3e} ms={res.ms:.2f}")
    if mode == "uniform":
        assert abs(float(r.sum(dtype=np.float64)) - 1.0) < 1e-5
    else:
        assert abs(float(r.sum(dtype=np.float64)) - float(r64.sum())) < 1e-5


def test_trace_table_times_every_sweep(karate, capfd):
    """params.trace prints the reference's table (page_rank.hpp:42-52) with each
    sweep's own device time: one sweep per captured launch, timed by events."""
    _, g, deg = karate
    pg = PageRankGraph(g, deg)
    capfd.readouterr()
    r, res = pg.run(tol=1e-7, max_iters=100, dangling="leak", trace=True)
    out = capfd.readouterr().out.splitlines()
    head = out.index(next(ln for ln in out if ln.split()[:4] == ["iter", "error", "time", "outgoing"]))
    rows = [ln.split() for ln in out[head + 1:head + 1 + res.iters]]
    assert [int(x[0]) for x in rows] == list(range(res.iters))
    errs, times = pg.last_errors(), pg.last_sweep_times()
    assert len(times) == res.iters and np.all(times > 0)
    np.testing.assert_allclose([float(x[1]) for x in rows], errs, rtol=1e-6, atol=1e-10)
    np.testing.assert_allclose([float(x[2]) for x in rows], times / 1e3, rtol=1e-5, atol=1e-10)
    # the traced solve computes what the untraced one does
    r2, res2 = pg.run(tol=1e-7, max_iters=100, dangling="leak")
    assert res2.iters == res.iters
    np.testing.assert_array_equal(r, r2)
    assert len(pg.last_sweep_times()) == 0
