"""Host-side graph types and builders (the data either side of the hot path).

``CSRGraph`` is the reference's ``adjacency<1>`` ("in-CSR"): row ``i`` lists
the in-neighbours ``j`` of vertex ``i`` (edges ``j -> i``); uint32 offsets
(n+1) and indices (m), as ``indexed_struct_of_arrays``
(include/nwgraph/containers/compressed.hpp:60-90, 179-186, 215-218).

* ``build_adjacency``  -- fill_directed<1> (include/nwgraph/build.hpp:209-265),
  with the optional uniq (:425-435) and remove_self_loops (:437-443).
* ``build_degrees``    -- bench/common.hpp:109-120 (out-degree by counting
  every neighbour occurrence).
* ``rmat`` / ``erdos_renyi`` / ``rmat_many`` -- synthetic inputs (no
  reference counterpart; the recipes are specified in include/nwgr_graphgen.h).
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np

from ._lib import graphgen_lib

CSR_DEDUP = 1
CSR_DROP_SELF_LOOPS = 2

RMAT_A, RMAT_B, RMAT_C = 0.57, 0.19, 0.19
DEFAULT_SEED = 27491095  # the reference drivers' default --seed (bench/bfs.cpp:32)


@dataclass
class CSRGraph:
    """In-CSR adjacency: offsets[n+1], indices[m] (uint32)."""

    offsets: np.ndarray
    indices: np.ndarray

    def __post_init__(self):
        self.offsets = np.ascontiguousarray(self.offsets, dtype=np.uint32)
        self.indices = np.ascontiguousarray(self.indices, dtype=np.uint32)

    @property
    def n(self) -> int:
        return len(self.offsets) - 1

    @property
    def m(self) -> int:
        return len(self.indices)

    def size(self) -> int:
        """Number of vertices, as graph.size() (compressed.hpp:218)."""
        return self.n

    def __len__(self) -> int:
        return self.n

    def __getitem__(self, i: int) -> np.ndarray:
        """Row view graph[i] (compressed.hpp:215-216)."""
        return self.indices[self.offsets[i]:self.offsets[i + 1]]

    def in_degrees(self) -> np.ndarray:
        return np.diff(self.offsets).astype(np.uint32)


def _take(h) -> CSRGraph:
    L = graphgen_lib()
    try:
        n = L.nwgr_host_csr_n(h)
        m = L.nwgr_host_csr_m(h)
        off = np.ctypeslib.as_array(L.nwgr_host_csr_offsets(h), (n + 1,)).copy()
        ind = np.ctypeslib.as_array(L.nwgr_host_csr_indices(h), (m,)).copy() if m else np.zeros(0, np.uint32)
        return CSRGraph(off, ind)
    finally:
        L.nwgr_host_csr_free(h)


def _check_gen(rc: int) -> None:
    if rc != 0:
        raise ValueError(graphgen_lib().nwgr_gen_last_error().decode(errors="replace"))


def build_adjacency(src, dst, n: int | None = None, dedup: bool = False, drop_self_loops: bool = False,
                    nthreads: int = 0) -> CSRGraph:
    """In-CSR of the directed edge list src[i] -> dst[i] (rows sorted)."""
    src = np.ascontiguousarray(src, dtype=np.uint32)
    dst = np.ascontiguousarray(dst, dtype=np.uint32)
    if src.shape != dst.shape:
        raise ValueError("src and dst differ in length")
    if n is None:  # N = max id + 1, as edge_list::push_back (edge_list.hpp:143-154)
        n = int(max(src.max(initial=0), dst.max(initial=0))) + 1 if len(src) else 0
    flags = (CSR_DEDUP if dedup else 0) | (CSR_DROP_SELF_LOOPS if drop_self_loops else 0)
    h = ctypes.c_void_p()
    _check_gen(graphgen_lib().nwgr_csr_build(n, len(src), src.ctypes.data, dst.ctypes.data, flags, nthreads,
                                             ctypes.byref(h)))
    return _take(h)


def build_degrees(graph: CSRGraph) -> np.ndarray:
    """Out-degrees: each neighbour j of each row counts once (common.hpp:109-120)."""
    return np.bincount(graph.indices, minlength=graph.n).astype(np.uint32)


def rmat(scale: int, edge_factor: int = 16, seed: int = DEFAULT_SEED, a: float = RMAT_A, b: float = RMAT_B,
         c: float = RMAT_C, nthreads: int = 0) -> CSRGraph:
    """Directed R-MAT, N = 2**scale, self-loops and duplicates removed."""
    h = ctypes.c_void_p()
    _check_gen(graphgen_lib().nwgr_gen_rmat(scale, edge_factor, seed, a, b, c, nthreads, ctypes.byref(h)))
    return _take(h)


def erdos_renyi(n: int, m_draws: int, seed: int = DEFAULT_SEED, nthreads: int = 0) -> CSRGraph:
    """Directed G(n, m) drawn with replacement, self-loops and duplicates removed."""
    h = ctypes.c_void_p()
    _check_gen(graphgen_lib().nwgr_gen_er(n, m_draws, seed, nthreads, ctypes.byref(h)))
    return _take(h)


def rmat_many(ns, edge_factor: int = 8, seed: int = DEFAULT_SEED, a: float = RMAT_A, b: float = RMAT_B,
              c: float = RMAT_C, nthreads: int = 0):
    """Independent R-MAT graphs stored block-diagonally in one CSR.

    Returns (graph, vbase) with vbase[g] the first global vertex of graph g
    (len(ns) + 1 entries)."""
    ns = np.ascontiguousarray(ns, dtype=np.uint32)
    h = ctypes.c_void_p()
    _check_gen(graphgen_lib().nwgr_gen_rmat_many(len(ns), ns.ctypes.data, edge_factor, seed, a, b, c, nthreads,
                                                 ctypes.byref(h)))
    vbase = np.zeros(len(ns) + 1, np.uint32)
    vbase[1:] = np.cumsum(ns, dtype=np.uint64).astype(np.uint32)
    return _take(h), vbase


def transpose(graph: CSRGraph, nthreads: int = 0):
    """Out-CSR of an in-CSR: (offsets, indices, perm) with perm[k] the in-CSR
    position of out-CSR entry k."""
    n, m = graph.n, graph.m
    off = np.zeros(n + 1, np.uint32)
    ind = np.zeros(max(m, 1), np.uint32)
    perm = np.zeros(max(m, 1), np.uint32)
    _check_gen(graphgen_lib().nwgr_csr_transpose(n, m, graph.offsets.ctypes.data, graph.indices.ctypes.data,
                                                 off.ctypes.data, ind.ctypes.data, perm.ctypes.data, nthreads))
    return off, ind[:m], perm[:m]


def row_cuts(graph: CSRGraph, parts: int) -> np.ndarray:
    """Boundaries (parts + 1 row ids) of `parts` contiguous slices of the
    destination rows with about equal entry counts: the 1-D row partition a
    multi-GPU solve would give each GPU."""
    if parts < 1:
        raise ValueError("parts must be >= 1")
    targets = (np.arange(1, parts, dtype=np.int64) * graph.m) // parts
    cuts = np.searchsorted(graph.offsets.astype(np.int64), targets, side="left")
    return np.concatenate([[0], cuts, [graph.n]]).astype(np.int64)


def row_shard(graph: CSRGraph, lo: int, hi: int) -> CSRGraph:
    """One GPU's share of a row-partitioned solve: the same n vertices (sources
    keep their ids, so the gathered vector is the full one) with entries only
    in rows [lo, hi).  A sweep over it computes rows [lo, hi) of the full
    graph's sweep from the same input vector."""
    if not 0 <= lo <= hi <= graph.n:
        raise ValueError("need 0 <= lo <= hi <= n")
    offs = graph.offsets.astype(np.int64)
    new = np.empty(graph.n + 1, np.int64)
    new[:lo + 1] = 0
    new[lo:hi + 1] = offs[lo:hi + 1] - offs[lo]
    new[hi + 1:] = offs[hi] - offs[lo]
    return CSRGraph(new, graph.indices[offs[lo]:offs[hi]])


def splitmix64_at(key: int, k: int) -> int:
    return int(graphgen_lib().nwgr_splitmix64_at(key, k))


# ---------------------------------------------------------------------------
# Ingestion: the inputs the reference's drivers load (bench/common.hpp:82-101)


@dataclass
class EdgeList:
    """Edge list src[i] -> dst[i] over n vertices, with the file's values when
    it has them (a MatrixMarket pattern file has none) and, when asked for or
    stored in an NW file, an int attribute per edge (`ints`)."""

    n: int
    src: np.ndarray
    dst: np.ndarray
    values: np.ndarray | None = None
    directed: bool = True
    ints: np.ndarray | None = None

    @property
    def m(self) -> int:
        return len(self.src)


def _take_edge_list(h) -> EdgeList:
    L = graphgen_lib()
    try:
        n, m = L.nwgr_edge_list_n(h), L.nwgr_edge_list_m(h)
        src = np.ctypeslib.as_array(L.nwgr_edge_list_src(h), (m,)).copy() if m else np.zeros(0, np.uint32)
        dst = np.ctypeslib.as_array(L.nwgr_edge_list_dst(h), (m,)).copy() if m else np.zeros(0, np.uint32)
        vp = L.nwgr_edge_list_values(h)
        vals = np.ctypeslib.as_array(vp, (m,)).copy() if (vp and m) else None
        ip = L.nwgr_edge_list_int_values(h)
        ints = (np.ctypeslib.as_array(ip, (m,)).copy() if m else np.zeros(0, np.int32)) if ip else None
        return EdgeList(n, src, dst, vals, L.nwgr_edge_list_directedness(h) == 1, ints)
    finally:
        L.nwgr_edge_list_free(h)


def read_mm(path: str, nthreads: int = 0, directed: bool = True, int_attribute: bool = False) -> EdgeList:
    """MatrixMarket coordinate file -> edge list, as read_mm<directed>
    (include/nwgraph/io/mmio.hpp:171-217): 1-based entries, a symmetric file's
    off-diagonal entries in both directions, n = max(rows, largest id).
    directed=False reads every entry once (read_mm<undirected>);
    int_attribute=True also keeps the int attribute read_mm<..., int> reads."""
    h = ctypes.c_void_p()
    _check_gen(graphgen_lib().nwgr_read_mm_as(str(path).encode(), nthreads, 1 if directed else 0,
                                              1 if int_attribute else 0, ctypes.byref(h)))
    return _take_edge_list(h)


def read_nw_edge_list(path: str) -> EdgeList:
    """The reference's binary edge list (index_edge_list::deserialize,
    include/nwgraph/edge_list.hpp:215-233)."""
    h = ctypes.c_void_p()
    _check_gen(graphgen_lib().nwgr_read_nw_edge_list(str(path).encode(), ctypes.byref(h)))
    return _take_edge_list(h)


def write_nw_edge_list(path: str, el: EdgeList, with_ints: bool = False) -> None:
    """index_edge_list::serialize (include/nwgraph/edge_list.hpp:197-203);
    with_ints writes el.ints as a third array (edge_list<..., int>)."""
    src = np.ascontiguousarray(el.src, dtype=np.uint32)
    dst = np.ascontiguousarray(el.dst, dtype=np.uint32)
    attr = None
    if with_ints:
        if el.ints is None or len(el.ints) != len(src):
            raise ValueError("with_ints needs one int attribute per edge")
        attr = np.ascontiguousarray(el.ints, dtype=np.int32)
    _check_gen(graphgen_lib().nwgr_write_nw_edge_list_attr(str(path).encode(), el.n, len(src), src.ctypes.data,
                                                            dst.ctypes.data, None if attr is None else attr.ctypes.data,
                                                            1 if el.directed else 0))


def write_nw_adjacency(path: str, graph: CSRGraph) -> None:
    """indexed_struct_of_arrays::serialize (containers/compressed.hpp:303-313)."""
    _check_gen(graphgen_lib().nwgr_write_nw_adjacency(str(path).encode(), graph.n, graph.m,
                                                       graph.offsets.ctypes.data, graph.indices.ctypes.data))


def read_nw_adjacency(path: str) -> CSRGraph:
    """indexed_struct_of_arrays::deserialize (containers/compressed.hpp:320-333)."""
    h = ctypes.c_void_p()
    _check_gen(graphgen_lib().nwgr_read_nw_adjacency(str(path).encode(), ctypes.byref(h)))
    return _take(h)


def load_graph(path: str, nthreads: int = 0) -> EdgeList:
    """load_graph<directed> (bench/common.hpp:82-101): an "NW" binary edge
    list or a MatrixMarket file, told apart by the first word."""
    with open(path, "rb") as f:
        head = f.read(16)
    if head.startswith(b"NW"):
        return read_nw_edge_list(path)
    if head.startswith(b"%%MatrixMarket"):
        return read_mm(path, nthreads)
    raise ValueError(f"did not recognize graph input file {path}")


def adjacency_from_edge_list(el: EdgeList, dedup: bool = False, drop_self_loops: bool = False,
                             nthreads: int = 0) -> CSRGraph:
    """build_adjacency<1> (bench/common.hpp:103-107) of a loaded edge list."""
    return build_adjacency(el.src, el.dst, n=el.n, dedup=dedup, drop_self_loops=drop_self_loops, nthreads=nthreads)
