"""Graphs whose entry counts exceed one device sort call.

The reference's CSR holds up to 2^32 - 1 entries in uint32 offsets
(util/defaults.hpp:23-24).  Graph construction here sorts and scans on the
device; hipCUB counts items in int, so counts above the pass size (2^30 by
default, nwgr_pr_set_sort_batch) are sorted and scanned in passes.

* With a pass size of 2^16 the passes run on small graphs: every solve on the
  graph built in passes must equal, bit for bit and sweep for sweep, the solve
  on the graph built by single sort calls (the passes are exact: stable
  placements and carried scans).
* A graph with more than 2^31 entries, built on the device: each row of a
  random graph G repeated 40 times (unsorted: the copies follow one another),
  with every divisor multiplied by 40.  In exact arithmetic its ranks are G's,
  so the FP64 oracle on G is the reference.
"""
import sys

import numpy as np
import pytest

import nwgr_amd
import oracle
from nwgr_amd import CSRGraph, PageRankGraph
from nwgr_amd._lib import check, pagerank_lib
from tests.conftest import l1

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture
def small_pass():
    L = pagerank_lib()
    check(L.nwgr_pr_set_sort_batch(1 << 16))
    try:
        yield 1 << 16
    finally:
        check(L.nwgr_pr_set_sort_batch(1 << 30))


def star_multigraph(seed=3):
    """Row 0 holds every vertex once plus 2^16 repeats (one row longer than
    the pass), the other rows 0-8 entries with repeats; rows unsorted."""
    rng = np.random.default_rng(seed)
    n = 1 << 17
    lens = rng.integers(0, 9, n)
    lens[0] = n + (1 << 16)
    offs = np.zeros(n + 1, np.uint64)
    np.cumsum(lens, out=offs[1:])
    idx = rng.integers(0, n, int(offs[-1])).astype(np.uint32)
    idx[: n] = rng.permutation(n).astype(np.uint32)
    return CSRGraph(offs.astype(np.uint32), idx)


def er_with_quiet_rows(seed=5):
    rng = np.random.default_rng(seed)
    n = 1 << 16
    lens = rng.integers(1, 40, n)
    lens[rng.random(n) < 0.1] = 0
    offs = np.zeros(n + 1, np.uint64)
    np.cumsum(lens, out=offs[1:])
    idx = rng.integers(0, n, int(offs[-1])).astype(np.uint32)
    return CSRGraph(offs.astype(np.uint32), idx)


def solves(g, deg):
    """Every solve of the drop-in path on a freshly built handle."""
    pg = PageRankGraph(g, deg)
    out = {}
    try:
        out["tiled"] = pg.run(tol=1e-6, max_iters=100, dangling="leak")
        out["tiled_uniform"] = pg.run(tol=1e-6, max_iters=100, dangling="uniform")
        out["fp64"] = pg.run64(tol=1e-10, max_iters=200)
        out["gs4"] = pg.run_gs(4, tol=1e-6, max_iters=100)
        out["gs4_order"] = pg.gs_phases(4)
        pg.tuning = {"engine": 0}
        out["binned"] = pg.run(tol=1e-6, max_iters=100, dangling="leak")
    finally:
        pg.close()
    return out


@pytest.mark.parametrize("name", ["rmat15", "rmat18", "star", "er_quiet"])
def test_passes_equal_single_calls(name, small_pass):
    say(f"passes {name}")
    g = {"rmat15": lambda: nwgr_amd.rmat(15, 16), "rmat18": lambda: nwgr_amd.rmat(18, 16),
         "star": star_multigraph, "er_quiet": er_with_quiet_rows}[name]()
    deg = nwgr_amd.build_degrees(g)
    assert g.m > 4 * small_pass
    check(pagerank_lib().nwgr_pr_set_sort_batch(1 << 30))
    one = solves(g, deg)
    check(pagerank_lib().nwgr_pr_set_sort_batch(small_pass))
    many = solves(g, deg)
    for key in ("tiled", "tiled_uniform", "fp64", "gs4", "binned"):
        (ra, resa), (rb, resb) = one[key], many[key]
        assert resa.iters == resb.iters, (key, resa.iters, resb.iters)
        np.testing.assert_array_equal(ra, rb, err_msg=key)
    for a, b in zip(one["gs4_order"], many["gs4_order"]):
        np.testing.assert_array_equal(a, b)
    # and the ranks are the oracle's
    r, res = one["tiled"]
    r64, _ = oracle.jacobi(g.offsets, g.indices, deg, 0.85, 0.0, res.iters, nthreads=16)
    assert l1(r, r64) <= 1e-6
    say(f"  {name}: m={g.m} sweeps={res.iters} L1={l1(r, r64):.2e}")


@pytest.mark.slow
@pytest.mark.timeout(900)
def test_more_than_2_31_entries():
    import torch

    T = 40
    rng = np.random.default_rng(11)
    n = 1 << 20
    lens = rng.integers(32, 97, n)
    lens[rng.random(n) < 0.05] = 0
    offs = np.zeros(n + 1, np.uint64)
    np.cumsum(lens, out=offs[1:])
    m0 = int(offs[-1])
    idx = rng.integers(0, n, m0).astype(np.uint32)
    deg = oracle.out_degree(n, idx).astype(np.uint32)
    m = T * m0
    assert (1 << 31) < m <= (1 << 32) - (1 << 28)
    say(f"2^31+ entries: n={n} m={m} (G: {m0} entries, each row {T} times)")

    dev = torch.device("cuda:0")
    offs_t = torch.from_numpy(offs.astype(np.int64)).to(dev)
    offs2_t = offs_t * T
    len_t = offs_t[1:] - offs_t[:-1]
    idx_t = torch.from_numpy(idx.view(np.int32)).to(dev)
    idx2 = torch.empty(m, dtype=torch.int32, device=dev)
    step = 1 << 27
    for p0 in range(0, m, step):
        p = torch.arange(p0, min(m, p0 + step), device=dev, dtype=torch.int64)
        row = torch.searchsorted(offs2_t, p, right=True) - 1
        q = (p - offs2_t[row]) % len_t[row]
        idx2[p0:p0 + len(p)] = idx_t[offs_t[row] + q]
        del p, row, q
    offs2 = torch.from_numpy((offs * T).astype(np.uint32).view(np.int32)).to(dev)
    deg2 = torch.from_numpy((deg.astype(np.uint64) * T).astype(np.uint32).view(np.int32)).to(dev)
    del offs_t, offs2_t, len_t, idx_t
    torch.cuda.synchronize()
    say("  built on the device; creating the handle")
    pg = PageRankGraph.from_device(offs2.data_ptr(), idx2.data_ptr(), n, m, deg2.data_ptr())
    try:
        assert pg.m == m
        del idx2
        torch.cuda.empty_cache()
        say(f"  setup {pg.setup_ms:.0f} ms, {pg.device_bytes / 2**30:.1f} GiB on the device")
        _, k_ref = oracle.jacobi(offs.astype(np.uint32), idx, deg, 0.85, 1e-6, 100, nthreads=16)
        for engine, label in ((2, "default"), (0, "binned")):
            pg.tuning = {"engine": engine}
            r, res = pg.run(tol=1e-6, max_iters=100, dangling="leak")
            r64, _ = oracle.jacobi(offs.astype(np.uint32), idx, deg, 0.85, 0.0, res.iters, nthreads=16)
            err = l1(r, r64)
            say(f"  {label} ({pg.last_engine()}): sweeps={res.iters} (oracle {k_ref}) L1={err:.2e} "
                f"{res.ms:.1f} ms")
            # about 2500 terms per row summed in FP32
            assert err <= 1e-5
            assert abs(res.iters - k_ref) <= 1
        r, res = pg.run64(tol=1e-10, max_iters=200)
        r64, _ = oracle.jacobi(offs.astype(np.uint32), idx, deg, 0.85, 0.0, res.iters, nthreads=16)
        err = l1(r, r64)
        say(f"  fp64: sweeps={res.iters} L1={err:.2e}")
        assert err <= 1e-12
    finally:
        pg.close()
