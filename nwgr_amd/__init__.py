"""nwgr_amd -- NWGraph's PageRank hot path, native to the AMD Instinct MI355X.

The product path is ``libnwgr_pagerank.so`` (hand-written HIP kernels for
gfx950 behind the C ABI of include/nwgr_pagerank.h).  This package holds the
host-side mirror of the reference interface (``page_rank``), the device
handle (``PageRankGraph``), the graph builders / generators that feed it, and
the torch autograd bridge.
"""
from .graph import (CSRGraph, EdgeList, adjacency_from_edge_list, build_adjacency, build_degrees,  # noqa: F401
                    erdos_renyi, load_graph, read_mm, read_nw_adjacency, read_nw_edge_list, rmat, rmat_many,
                    row_cuts, row_shard, transpose, write_nw_adjacency, write_nw_edge_list)
from .page_rank import (PageRankGraph, RunResult, cache_clear, cache_set_capacity, cache_stats,  # noqa: F401
                        device_count, page_rank, page_rank_delta, top_k, top_k_device)
from .multi import MultiPageRank  # noqa: F401
from ._lib import NwgrError  # noqa: F401

__all__ = ["CSRGraph", "EdgeList", "build_adjacency", "build_degrees", "rmat", "erdos_renyi", "rmat_many", "transpose",
           "read_mm", "read_nw_edge_list", "write_nw_edge_list", "read_nw_adjacency", "write_nw_adjacency",
           "load_graph", "adjacency_from_edge_list", "row_cuts", "row_shard",
           "PageRankGraph", "MultiPageRank", "RunResult", "page_rank", "page_rank_delta", "device_count", "NwgrError", "top_k",
           "top_k_device", "cache_clear", "cache_stats", "cache_set_capacity"]
