"""Many small graphs solved in one persistent launch (``nwgr_multi_*`` in
include/nwgr_pagerank.h).

The graphs are stored block-diagonally in one in-CSR: graph g owns the
vertices ``[vbase[g], vbase[g+1])``, as produced by
``nwgr_amd.graph.rmat_many``.  Every graph is solved on its own with the
semantics of the plain solve (``n_g`` in place of ``n``).  No reference
counterpart: the reference solves one graph per call (page_rank.hpp:89-148).
"""
from __future__ import annotations

import ctypes

import numpy as np

from ._lib import PrResult, check, pagerank_lib
from .graph import CSRGraph
from .page_rank import RunResult, _params


class MultiPageRank:
    """Device-resident handle over a block-diagonal batch of graphs."""

    def __init__(self, graph: CSRGraph, vbase, degrees=None, device: int = 0):
        L = pagerank_lib()
        self._h = ctypes.c_void_p()
        vb = np.ascontiguousarray(vbase, dtype=np.uint32)
        deg = None if degrees is None else np.ascontiguousarray(degrees, dtype=np.uint32)
        check(L.nwgr_multi_create(graph.offsets.ctypes.data, graph.indices.ctypes.data if graph.m else None,
                                  graph.n, graph.m, vb.ctypes.data, len(vb) - 1,
                                  None if deg is None else deg.ctypes.data, device, ctypes.byref(self._h)))
        self.n = graph.n
        self.m = graph.m
        self.count = len(vb) - 1
        self.vbase = vb

    def close(self):
        if self._h:
            pagerank_lib().nwgr_multi_destroy(self._h)
            self._h = ctypes.c_void_p()

    __del__ = close

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    @property
    def setup_ms(self) -> float:
        return float(pagerank_lib().nwgr_multi_setup_ms(self._h))

    @property
    def tuning(self) -> dict:
        a, b = ctypes.c_uint32(), ctypes.c_uint32()
        check(pagerank_lib().nwgr_multi_get_tuning(self._h, ctypes.byref(a), ctypes.byref(b)))
        return {"lds_max_n": a.value, "heavy_min": b.value}

    @tuning.setter
    def tuning(self, values: dict) -> None:
        cur = self.tuning
        unknown = set(values) - set(cur)
        if unknown:
            raise KeyError(f"unknown tuning fields {sorted(unknown)}")
        cur.update(values)
        check(pagerank_lib().nwgr_multi_set_tuning(self._h, cur["lds_max_n"], cur["heavy_min"]))

    def run(self, damping=0.85, tol=1e-6, max_iters=100, dangling="uniform", teleport=None, init=None, out=None):
        """Solve every graph; returns (ranks float32[n], RunResult)."""
        p = _params(damping, tol, max_iters, dangling, 0, False)
        keep = []
        for name, vec in (("teleport", teleport), ("init", init)):
            if vec is not None:
                v = np.ascontiguousarray(vec, dtype=np.float32)
                if v.shape != (self.n,):
                    raise ValueError(f"{name} must have shape ({self.n},)")
                keep.append(v)
                setattr(p, name, v.ctypes.data)
        if out is None:
            out = np.empty(self.n, np.float32)
        res = PrResult()
        check(pagerank_lib().nwgr_multi_run_f32(self._h, ctypes.byref(p), out.ctypes.data, ctypes.byref(res)))
        return out, RunResult(int(res.iters), bool(res.converged), float(res.error), float(res.ms))

    def run_device(self, out_ptr: int, damping=0.85, tol=1e-6, max_iters=100, dangling="uniform",
                   teleport_ptr=None, init_ptr=None, stream=None) -> RunResult:
        p = _params(damping, tol, max_iters, dangling, 0, False)
        p.teleport = teleport_ptr
        p.init = init_ptr
        res = PrResult()
        check(pagerank_lib().nwgr_multi_run_device_f32(self._h, ctypes.byref(p), out_ptr, ctypes.byref(res),
                                                       stream))
        return RunResult(int(res.iters), bool(res.converged), float(res.error), float(res.ms))

    def graph_results(self):
        """(sweeps uint32[count], final L1 change float64[count]) of the last run."""
        it = np.zeros(self.count, np.uint32)
        er = np.zeros(self.count, np.float64)
        check(pagerank_lib().nwgr_multi_graph_results(self._h, it.ctypes.data, er.ctypes.data))
        return it, er
