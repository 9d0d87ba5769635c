"""In-CSR construction on the device against the host builder (needs the GPU).

The host builder (nwgr_amd.build_adjacency, graphgen.cpp) restates the
reference's edge_list -> adjacency<1> build (build.hpp:209-265, uniq
:425-435, remove_self_loops :437-443) with sorted rows; the device build must
produce the identical offsets and indices, and out-degrees equal to
build_degrees (bench/common.hpp:109-120).
"""
import numpy as np
import pytest
import torch

import nwgr_amd
from nwgr_amd import PageRankGraph
from nwgr_amd._lib import NWGR_EINVAL, NwgrError
from nwgr_amd.page_rank import build_csr_device, graph_from_edges_device

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda:0")


def edges(n, m, seed, skew=False):
    rng = np.random.default_rng(seed)
    if skew:  # heavy duplicates and hub rows
        src = (rng.pareto(1.2, m) * 10).astype(np.int64) % n
        dst = (rng.pareto(1.2, m) * 10).astype(np.int64) % n
    else:
        src = rng.integers(0, n, m)
        dst = rng.integers(0, n, m)
    return src.astype(np.uint32), dst.astype(np.uint32)


def to_dev(a):
    return torch.from_numpy(a.astype(np.int32)).to(DEV)


@pytest.mark.parametrize("dedup,drop", [(False, False), (True, False), (False, True), (True, True)])
@pytest.mark.parametrize("n,m,skew", [(1000, 20000, False), (1 << 16, 1 << 20, True), (1, 5, False)])
def test_device_csr_equals_host(n, m, skew, dedup, drop):
    src, dst = edges(n, m, n + m, skew)
    host = nwgr_amd.build_adjacency(src, dst, n=n, dedup=dedup, drop_self_loops=drop)
    offs, idx, deg = build_csr_device(to_dev(src), to_dev(dst), n, dedup=dedup, drop_self_loops=drop)
    np.testing.assert_array_equal(offs.cpu().numpy().view(np.uint32), host.offsets)
    np.testing.assert_array_equal(idx.cpu().numpy().view(np.uint32), host.indices)
    np.testing.assert_array_equal(deg.cpu().numpy().view(np.uint32), nwgr_amd.build_degrees(host))


def test_device_csr_edge_cases():
    e = torch.empty(0, dtype=torch.int32, device=DEV)
    offs, idx, deg = build_csr_device(e, e, 7)
    assert offs.cpu().tolist() == [0] * 8 and idx.numel() == 0 and deg.cpu().tolist() == [0] * 7
    with pytest.raises(NwgrError) as err:
        build_csr_device(to_dev(np.array([0, 9], np.uint32)), to_dev(np.array([1, 2], np.uint32)), 5)
    assert err.value.code == NWGR_EINVAL


def test_graph_from_device_edges_solves_like_host_graph():
    g = nwgr_amd.rmat(14, 16)
    # the CSR's own edge list, shuffled, with duplicates and self-loops added
    rows = np.repeat(np.arange(g.n, dtype=np.uint32), np.diff(g.offsets))
    src, dst = g.indices.copy(), rows
    rng = np.random.default_rng(2)
    extra = rng.integers(0, g.m, 5000)
    src = np.concatenate([src, src[extra], np.arange(100, dtype=np.uint32)])
    dst = np.concatenate([dst, dst[extra], np.arange(100, dtype=np.uint32)])
    perm = rng.permutation(len(src))
    src, dst = src[perm], dst[perm]
    pg_dev = graph_from_edges_device(to_dev(src), to_dev(dst), g.n)
    assert pg_dev.m == g.m
    r_dev, res_dev = pg_dev.run(tol=1e-6, max_iters=100, dangling="uniform")
    r_host, res_host = PageRankGraph(g, nwgr_amd.build_degrees(g)).run(tol=1e-6, max_iters=100, dangling="uniform")
    assert res_dev.iters == res_host.iters
    np.testing.assert_array_equal(r_dev, r_host)
