"""Device top-k (nwgr_top_k_*) and PageRank on graphs read from the
reference's MatrixMarket data files (needs the GPU)."""
import os

import numpy as np
import pytest

import nwgr_amd
import oracle
from nwgr_amd import PageRankGraph
from nwgr_amd._lib import NwgrError
from tests.conftest import GOLDEN, assert_same_top_k
from tests.test_gpu_pagerank import parity, say

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n,k,levels", [(1, 1, 1), (1000, 10, 7), (100_000, 1000, 50), (3_000_000, 100, 0)])
def test_top_k_matches_numpy_order(n, k, levels):
    """Non-increasing rank, ascending id among equal ranks."""
    rng = np.random.default_rng(n)
    r = (rng.integers(0, levels, n) / max(levels, 1)).astype(np.float32) if levels else \
        rng.random(n).astype(np.float32)
    ids, vals = nwgr_amd.top_k(r, k)
    want = np.lexsort((np.arange(n), -r.astype(np.float64)))[:k]
    np.testing.assert_array_equal(ids, want)
    np.testing.assert_array_equal(vals, r[want])


def test_top_k_device_form_and_errors():
    import torch
    r = torch.rand(5000, device="cuda:0")
    ids = torch.empty(50, dtype=torch.int32, device="cuda:0")
    vals = torch.empty(50, dtype=torch.float32, device="cuda:0")
    nwgr_amd.top_k_device(r.data_ptr(), 5000, 50, ids.data_ptr(), vals.data_ptr(),
                          torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    rh = r.cpu().numpy()
    want = np.lexsort((np.arange(5000), -rh.astype(np.float64)))[:50]
    np.testing.assert_array_equal(ids.cpu().numpy(), want)
    with pytest.raises(ValueError):
        nwgr_amd.top_k(rh, 5001)
    with pytest.raises(NwgrError):
        nwgr_amd.top_k_device(r.data_ptr(), 5000, 6000, ids.data_ptr(), None, None)


@pytest.mark.parametrize("name", ["USAir97.mtx", "Erdos981.mtx", "dwt___66.mtx", "karate.mtx"])
@pytest.mark.parametrize("mode", ["leak", "uniform"])
def test_pagerank_on_matrix_market_inputs(name, mode):
    """load_graph -> build_adjacency<1> -> build_degrees -> page_rank, the
    chain of bench/pr.cpp:73-89, on the reference's data files."""
    say(f"mm {name} {mode}")
    el = nwgr_amd.load_graph(os.path.join(GOLDEN, "mtx", name))
    g = nwgr_amd.adjacency_from_edge_list(el)
    deg = nwgr_amd.build_degrees(g)
    r, res, r64, err = parity(g, deg, mode, 1e-6, 100, topk=min(20, g.n))
    ids, _ = nwgr_amd.top_k(r, min(10, g.n))
    assert_same_top_k(r, r64, min(10, g.n))
    np.testing.assert_array_equal(ids, oracle.top_k(r, min(10, g.n)))
