"""nwgr_amd -- NWGraph's PageRank hot path, native to the AMD Instinct MI355X.

The product path is ``libnwgr_pagerank.so`` (hand-written HIP kernels for
gfx950 behind the C ABI of include/nwgr_pagerank.h).  This package holds the
host-side mirror of the reference interface (``page_rank``), the device
handle (``PageRankGraph``), the graph builders / generators that feed it, and
the torch autograd bridge.
"""
from .graph import (CSRGraph, build_adjacency, build_degrees, erdos_renyi, rmat, rmat_many,  # noqa: F401
                    transpose)
from .page_rank import PageRankGraph, RunResult, device_count, page_rank  # noqa: F401
from .multi import MultiPageRank  # noqa: F401
from ._lib import NwgrError  # noqa: F401

__all__ = ["CSRGraph", "build_adjacency", "build_degrees", "rmat", "erdos_renyi", "rmat_many", "transpose",
           "PageRankGraph", "MultiPageRank", "RunResult", "page_rank", "device_count", "NwgrError"]
