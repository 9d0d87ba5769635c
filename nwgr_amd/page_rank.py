"""PageRank on the MI355X through the C ABI (include/nwgr_pagerank.h).

``page_rank`` mirrors the reference call shape
``nw::graph::page_rank(graph, degrees, page_rank, damping_factor, threshold,
max_iters, num_threads)`` (include/nwgraph/algorithms/page_rank.hpp:89-91):
it writes the ranks into the caller's pre-sized array and returns None; not
converging within ``max_iters`` is not an error (page_rank.hpp:121-147).
Arithmetic is FP32 on the device with Jacobi sweeps
(experimental/algorithms/page_rank.hpp:450-505).

``PageRankGraph`` is the device-resident handle for repeated solves, the
dangling-mass / teleport extensions and the device-memory (torch) path.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np

from . import _lib
from ._lib import DANGLING, PrParams, PrResult, check, pagerank_lib
from .graph import CSRGraph


@dataclass
class RunResult:
    iters: int        # sweeps performed
    converged: bool   # last sweep's L1 change < tol
    error: float      # L1 change of the last sweep
    ms: float         # device time of the call


def _params(damping, tol, max_iters, dangling, chunk, trace) -> PrParams:
    p = PrParams()
    pagerank_lib().nwgr_pr_params_default(ctypes.byref(p))
    p.damping = float(damping)
    p.tol = float(tol)
    p.max_iters = int(max_iters)
    if isinstance(dangling, str):
        if dangling not in DANGLING:
            raise ValueError(f"dangling must be one of {sorted(DANGLING)}")
        dangling = DANGLING[dangling]
    p.dangling = int(dangling)
    p.chunk = int(chunk)
    p.trace = 1 if trace else 0
    return p


def device_count() -> int:
    return int(pagerank_lib().nwgr_device_count())


class PageRankGraph:
    """A graph uploaded to one GPU, ready for PageRank sweeps."""

    def __init__(self, graph: CSRGraph | None = None, degrees=None, device: int = 0, *, _handle=None,
                 _owned: bool = True):
        L = pagerank_lib()
        self._h = ctypes.c_void_p()
        self._owned = _owned
        if _handle is not None:
            self._h = _handle
        else:
            off = np.ascontiguousarray(graph.offsets, dtype=np.uint32)
            ind = np.ascontiguousarray(graph.indices, dtype=np.uint32)
            deg = None if degrees is None else np.ascontiguousarray(degrees, dtype=np.uint32)
            if deg is not None and len(deg) != graph.n:
                raise ValueError("degrees must have one entry per vertex")
            check(L.nwgr_pr_graph_create(off.ctypes.data, ind.ctypes.data if len(ind) else None, graph.n,
                                         graph.m, None if deg is None else deg.ctypes.data, device,
                                         ctypes.byref(self._h)))
        self.n = int(L.nwgr_pr_graph_n(self._h))
        self.m = int(L.nwgr_pr_graph_m(self._h))
        self.device = device

    @classmethod
    def from_device(cls, offsets, indices, n: int, m: int, degrees=None, device: int = 0, stream=None):
        """Create from device arrays given as raw pointers (ints), e.g.
        torch tensors' data_ptr(); `stream` is a hipStream_t as int or None."""
        L = pagerank_lib()
        h = ctypes.c_void_p()
        check(L.nwgr_pr_graph_create_device(offsets, indices, n, m, degrees, device, stream, ctypes.byref(h)))
        return cls(_handle=h, device=device)

    @classmethod
    def cached(cls, graph: CSRGraph, degrees=None, device: int = 0):
        """The resident graph of the library's cache for (graph, degrees,
        device) -- the handle an earlier call with the same arrays made, when
        their contents still match, else a new one (nwgr_pr_graph_cached).
        Returns (PageRankGraph, hit).  The cache owns the handle: closing the
        returned object does not destroy it; keep it only until the next
        cached() call."""
        L = pagerank_lib()
        off = np.ascontiguousarray(graph.offsets, dtype=np.uint32)
        ind = np.ascontiguousarray(graph.indices, dtype=np.uint32)
        if degrees is not None:
            deg = np.ascontiguousarray(degrees, dtype=np.uint32)
            if len(deg) != graph.n:
                raise ValueError("degrees must have one entry per vertex")
        if off is not graph.offsets or ind is not graph.indices or (degrees is not None and deg is not degrees):
            # the cache keys on array addresses: a converted copy would miss every time
            raise ValueError("cached() needs contiguous uint32 offsets, indices and degrees")
        h = ctypes.c_void_p()
        hit = ctypes.c_uint32()
        check(L.nwgr_pr_graph_cached(off.ctypes.data if len(off) else None, ind.ctypes.data if len(ind) else None,
                                     graph.n, graph.m, None if degrees is None else deg.ctypes.data, device,
                                     ctypes.byref(h), ctypes.byref(hit)))
        return cls(_handle=h, device=device, _owned=False), bool(hit.value)

    def close(self) -> None:
        if getattr(self, "_h", None) and self._h.value:
            if self._owned:
                pagerank_lib().nwgr_pr_graph_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    @property
    def device_bytes(self) -> int:
        return int(pagerank_lib().nwgr_pr_graph_device_bytes(self._h))

    @property
    def setup_ms(self) -> float:
        return float(pagerank_lib().nwgr_pr_graph_setup_ms(self._h))

    @property
    def tuning(self) -> dict:
        """Work decomposition of the sweeps (see nwgr_pr_tuning)."""
        t = _lib.PrTuning()
        check(pagerank_lib().nwgr_pr_graph_get_tuning(self._h, ctypes.byref(t)))
        return {name: int(getattr(t, name)) for name, _ in _lib.PrTuning._fields_}

    @tuning.setter
    def tuning(self, values: dict) -> None:
        cur = self.tuning
        unknown = set(values) - set(cur)
        if unknown:
            raise KeyError(f"unknown tuning fields {sorted(unknown)}")
        cur.update(values)
        t = _lib.PrTuning(**cur)
        check(pagerank_lib().nwgr_pr_graph_set_tuning(self._h, ctypes.byref(t)))

    def last_profile(self) -> dict:
        """Sweep-kernel timing of the last solve: summed event time of the
        captured sweep chunks, number of sweep launches and chunks."""
        ms, launches, chunks = ctypes.c_double(), ctypes.c_uint32(), ctypes.c_uint32()
        check(pagerank_lib().nwgr_pr_last_profile(self._h, ctypes.byref(ms), ctypes.byref(launches),
                                                  ctypes.byref(chunks)))
        return {"sweep_ms": ms.value, "launches": launches.value, "chunks": chunks.value}

    def last_engine(self) -> str | None:
        """Engine of the last solve: "binned", "tiled" (second pass gathering
        each row's partials), "tiled-accumulate" (second pass accumulating
        row blocks in LDS), "tiled-gauss-seidel" (block Gauss-Seidel phases),
        or None before any."""
        e = pagerank_lib().nwgr_pr_last_engine(self._h)
        return {0: "binned", 1: "tiled", 2: "tiled-accumulate", 3: "tiled-gauss-seidel"}.get(e)

    def last_sweep_times(self) -> np.ndarray:
        """Device ms of every sweep of the last run if it was traced, else empty."""
        buf = np.zeros(4096, np.float64)
        k = pagerank_lib().nwgr_pr_last_sweep_times(self._h, buf.ctypes.data, len(buf))
        return buf[:k].copy()

    def last_errors(self) -> np.ndarray:
        buf = np.zeros(4096, np.float64)
        k = pagerank_lib().nwgr_pr_last_errors(self._h, buf.ctypes.data, len(buf))
        return buf[:k].copy()

    def run(self, damping=0.85, tol=1e-4, max_iters=20, dangling="leak", teleport=None, init=None,
            chunk=0, trace=False, out=None):
        """Solve; returns (ranks float32[n] in caller order, RunResult)."""
        p = _params(damping, tol, max_iters, dangling, chunk, trace)
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
        elif out.dtype != np.float32 or out.shape != (self.n,) or not out.flags.c_contiguous:
            raise ValueError("out must be a contiguous float32 array of n entries")
        res = PrResult()
        check(pagerank_lib().nwgr_pr_run_f32(self._h, ctypes.byref(p), out.ctypes.data, ctypes.byref(res)))
        return out, RunResult(int(res.iters), bool(res.converged), float(res.error), float(res.ms))

    def run_gs(self, phases=8, damping=0.85, tol=1e-4, max_iters=20, teleport=None, init=None, chunk=0,
               trace=False, out=None):
        """Block Gauss-Seidel solve (nwgr_pr_gs_run_f32): the relabelled rows
        cut into `phases` phases updated in place one after the other, LEAK
        dangling; returns (ranks float32[n] in caller order, RunResult)."""
        p = _params(damping, tol, max_iters, "leak", chunk, trace)
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
        elif out.dtype != np.float32 or out.shape != (self.n,) or not out.flags.c_contiguous:
            raise ValueError("out must be a contiguous float32 array of n entries")
        res = PrResult()
        check(pagerank_lib().nwgr_pr_gs_run_f32(self._h, ctypes.byref(p), int(phases), out.ctypes.data,
                                                ctypes.byref(res)))
        return out, RunResult(int(res.iters), bool(res.converged), float(res.error), float(res.ms))

    def run_gs_device(self, out_ptr: int, phases=8, damping=0.85, tol=1e-4, max_iters=20, teleport_ptr=None,
                      init_ptr=None, chunk=0, stream=None) -> RunResult:
        """Block Gauss-Seidel solve with the ranks written to device memory."""
        p = _params(damping, tol, max_iters, "leak", chunk, False)
        p.teleport = teleport_ptr
        p.init = init_ptr
        res = PrResult()
        check(pagerank_lib().nwgr_pr_gs_run_device_f32(self._h, ctypes.byref(p), int(phases), out_ptr,
                                                       ctypes.byref(res), stream))
        return RunResult(int(res.iters), bool(res.converged), float(res.error), float(res.ms))

    def run_share_device(self, out_ptr: int, shares: int, share: int, damping=0.85, tol=1e-4, max_iters=20,
                         chunk=0, stream=None) -> RunResult:
        """One GPU's row share swept alone (nwgr_pr_share_run_device_f32):
        rows of share `share` of `shares` updated against the others'
        contributions held at their initial values; LEAK dangling."""
        p = _params(damping, tol, max_iters, "leak", chunk, False)
        res = PrResult()
        check(pagerank_lib().nwgr_pr_share_run_device_f32(self._h, ctypes.byref(p), int(shares), int(share), out_ptr,
                                                          ctypes.byref(res), stream))
        return RunResult(int(res.iters), bool(res.converged), float(res.error), float(res.ms))

    def share_balance(self, shares=8, rounds=2, stream=None) -> float:
        """Re-cut the rows of `shares` shares by their measured sweep times
        (nwgr_pr_share_balance); the shares and the block Gauss-Seidel solve
        with phases = shares then use these cuts.  Returns the slowest share's
        device time per sweep (us)."""
        us = ctypes.c_double()
        check(pagerank_lib().nwgr_pr_share_balance(self._h, int(shares), int(rounds), ctypes.byref(us), stream))
        return float(us.value)

    def gs_phases(self, phases=8):
        """(order uint32[n]: the rows in the order the block Gauss-Seidel solve
        visits them, caller ids; cuts uint32[phases + 1]: where each phase
        begins in that order)."""
        order = np.empty(self.n, np.uint32)
        cuts = np.empty(int(phases) + 1, np.uint32)
        check(pagerank_lib().nwgr_pr_gs_phases(self._h, int(phases), order.ctypes.data, cuts.ctypes.data))
        return order, cuts

    def run64(self, damping=0.85, tol=1e-4, max_iters=20, dangling="leak", teleport=None, init=None, chunk=0,
              out=None):
        """FP64 solve (nwgr_pr_run_f64): the same iteration in double, the
        reference's Real = double; returns (ranks float64[n] in caller order,
        RunResult).  Teleport and initial vectors are taken as float32."""
        p = _params(damping, tol, max_iters, dangling, chunk, False)
        keep = []
        for name, vec in (("teleport", teleport), ("init", init)):
            if vec is not None:
                v = np.ascontiguousarray(vec, dtype=np.float32)
                if v.shape != (self.n,):
                    raise ValueError(f"{name} must have shape ({self.n},)")
                keep.append(v)
                setattr(p, name, v.ctypes.data)
        if out is None:
            out = np.empty(self.n, np.float64)
        elif out.dtype != np.float64 or out.shape != (self.n,) or not out.flags.c_contiguous:
            raise ValueError("out must be a contiguous float64 array of n entries")
        res = PrResult()
        check(pagerank_lib().nwgr_pr_run_f64(self._h, ctypes.byref(p), out.ctypes.data, ctypes.byref(res)))
        return out, RunResult(int(res.iters), bool(res.converged), float(res.error), float(res.ms))

    def run_device64(self, out_ptr: int, damping=0.85, tol=1e-4, max_iters=20, dangling="leak",
                     teleport_ptr=None, init_ptr=None, chunk=0, stream=None) -> RunResult:
        """FP64 solve with the ranks (n doubles) written to device memory."""
        p = _params(damping, tol, max_iters, dangling, chunk, False)
        p.teleport = teleport_ptr
        p.init = init_ptr
        res = PrResult()
        check(pagerank_lib().nwgr_pr_run_device_f64(self._h, ctypes.byref(p), out_ptr, ctypes.byref(res), stream))
        return RunResult(int(res.iters), bool(res.converged), float(res.error), float(res.ms))

    def run_delta(self, damping=0.85, threshold=1e-4, max_iters=20, chunk=0, out=None):
        """Residual ("delta") solve as page_rank_v13
        (experimental/algorithms/page_rank.hpp:632-698): returns (ranks
        float32[n] in caller order, unnormalised, RunResult with iters =
        iterations); last_errors() then holds the pushes of every iteration."""
        p = _params(damping, threshold, max_iters, "leak", chunk, False)
        if out is None:
            out = np.empty(self.n, np.float32)
        elif out.dtype != np.float32 or out.shape != (self.n,) or not out.flags.c_contiguous:
            raise ValueError("out must be a contiguous float32 array of n entries")
        res = PrResult()
        check(pagerank_lib().nwgr_pr_delta_run_f32(self._h, ctypes.byref(p), out.ctypes.data, ctypes.byref(res)))
        return out, RunResult(int(res.iters), bool(res.converged), float(res.error), float(res.ms))

    def run_delta_device(self, out_ptr: int, damping=0.85, threshold=1e-4, max_iters=20, chunk=0,
                         stream=None) -> RunResult:
        """Residual solve with the ranks written to device memory (raw pointer)."""
        p = _params(damping, threshold, max_iters, "leak", chunk, False)
        res = PrResult()
        check(pagerank_lib().nwgr_pr_delta_run_device_f32(self._h, ctypes.byref(p), out_ptr, ctypes.byref(res),
                                                          stream))
        return RunResult(int(res.iters), bool(res.converged), float(res.error), float(res.ms))

    def run_device(self, out_ptr: int, damping=0.85, tol=1e-4, max_iters=20, dangling="leak",
                   teleport_ptr=None, init_ptr=None, chunk=0, stream=None) -> RunResult:
        """Solve with device-resident vectors given as raw pointers."""
        p = _params(damping, tol, max_iters, dangling, chunk, False)
        p.teleport = teleport_ptr
        p.init = init_ptr
        res = PrResult()
        check(pagerank_lib().nwgr_pr_run_device_f32(self._h, ctypes.byref(p), out_ptr, ctypes.byref(res), stream))
        return RunResult(int(res.iters), bool(res.converged), float(res.error), float(res.ms))

    def wpr_forward(self, w_ptr: int, out_ptr: int, damping=0.85, tol=1e-6, max_iters=100, dangling="leak",
                    teleport_ptr=None, chunk=0, stream=None) -> RunResult:
        """Weighted forward on device arrays (raw pointers, caller order)."""
        p = _params(damping, tol, max_iters, dangling, chunk, False)
        p.teleport = teleport_ptr
        res = PrResult()
        check(pagerank_lib().nwgr_wpr_forward_device_f32(self._h, ctypes.byref(p), w_ptr, out_ptr,
                                                         ctypes.byref(res), stream))
        return RunResult(int(res.iters), bool(res.converged), float(res.error), float(res.ms))

    def wpr_adjoint(self, w_ptr: int, r_ptr: int, grad_r_ptr: int, grad_w_ptr: int, grad_t_ptr: int, damping=0.85,
                    tol=1e-6, max_iters=1000, dangling="leak", teleport_ptr=None, chunk=0,
                    stream=None) -> RunResult:
        """Adjoint on device arrays; `tol` is an absolute L1 threshold on lambda."""
        p = _params(damping, tol, max_iters, dangling, chunk, False)
        p.teleport = teleport_ptr
        res = PrResult()
        check(pagerank_lib().nwgr_wpr_adjoint_device_f32(self._h, ctypes.byref(p), w_ptr, r_ptr, grad_r_ptr,
                                                         grad_w_ptr, grad_t_ptr, ctypes.byref(res), stream))
        return RunResult(int(res.iters), bool(res.converged), float(res.error), float(res.ms))


    def ppr_batch(self, seeds, damping=0.85, tol=1e-6, max_iters=100, dangling="leak", out=None):
        """Personalized PageRank for each seed vertex (one-hot teleports), all
        columns solved together; returns (R float32[n, k], RunResult)."""
        sd = np.ascontiguousarray(seeds, dtype=np.uint32)
        k = int(sd.size)
        if out is None:
            out = np.empty((self.n, k), np.float32)
        elif out.dtype != np.float32 or out.shape != (self.n, k) or not out.flags.c_contiguous:
            raise ValueError("out must be a contiguous float32 array of shape (n, k)")
        p = _params(damping, tol, max_iters, dangling, 0, False)
        res = PrResult()
        check(pagerank_lib().nwgr_ppr_batch_f32(self._h, ctypes.byref(p), k, sd.ctypes.data, out.ctypes.data,
                                                ctypes.byref(res)))
        return out, RunResult(int(res.iters), bool(res.converged), float(res.error), float(res.ms))

    def ppr_batch_device(self, seeds_ptr: int, k: int, out_ptr: int, damping=0.85, tol=1e-6, max_iters=100,
                         dangling="leak", stream=None) -> RunResult:
        """Batched personalized solve on device arrays: seeds u32[k], out f32[n, k]."""
        p = _params(damping, tol, max_iters, dangling, 0, False)
        res = PrResult()
        check(pagerank_lib().nwgr_ppr_batch_device_f32(self._h, ctypes.byref(p), k, seeds_ptr, out_ptr,
                                                       ctypes.byref(res), stream))
        return RunResult(int(res.iters), bool(res.converged), float(res.error), float(res.ms))

    def ppr_column_errors(self) -> np.ndarray:
        """Per-column L1 change of the last sweep of the last batched solve."""
        buf = np.zeros(256, np.float64)
        k = pagerank_lib().nwgr_ppr_last_column_errors(self._h, buf.ctypes.data, len(buf))
        return buf[:k].copy()


def _resident(graph: CSRGraph, degrees, device: int) -> PageRankGraph:
    """The device graph of a drop-in call: the cache's resident one when the
    degrees are a contiguous uint32 array (as build_degrees returns them;
    the cache keys on array addresses), else a handle for this call only."""
    if degrees is not None:
        deg = np.ascontiguousarray(degrees, dtype=np.uint32)
        if deg is not degrees:  # a converted copy would never be found again
            return PageRankGraph(graph, deg, device)
    g, _ = PageRankGraph.cached(graph, degrees, device)
    return g


_SOLVE_F32, _SOLVE_F64, _SOLVE_DELTA_F32 = 0, 1, 2


def _cached_run(graph: CSRGraph, degrees, device: int, solver: int, p: PrParams, out: np.ndarray) -> bool:
    """One drop-in call on the cache's resident graph (nwgr_pr_cached_run:
    the arrays' contents are confirmed while the solve runs).  False when
    the arrays are not the contiguous uint32 ones the cache keys on."""
    off = np.ascontiguousarray(graph.offsets, dtype=np.uint32)
    ind = np.ascontiguousarray(graph.indices, dtype=np.uint32)
    deg = None if degrees is None else np.ascontiguousarray(degrees, dtype=np.uint32)
    if off is not graph.offsets or ind is not graph.indices or (degrees is not None and deg is not degrees):
        return False
    if deg is not None and len(deg) != graph.n:
        raise ValueError("degrees must have one entry per vertex")
    res = PrResult()
    check(pagerank_lib().nwgr_pr_cached_run(off.ctypes.data if len(off) else None,
                                            ind.ctypes.data if len(ind) else None, graph.n, graph.m,
                                            None if deg is None else deg.ctypes.data, device, solver,
                                            ctypes.byref(p), out.ctypes.data, ctypes.byref(res), None))
    return True


def _output(page_rank, n: int, dtype=np.float32):
    """(array of `dtype` the solve writes, whether it is page_rank itself)."""
    if isinstance(page_rank, np.ndarray) and page_rank.dtype == dtype and page_rank.flags.c_contiguous \
            and page_rank.shape == (n,):
        return page_rank, True
    return np.empty(n, dtype), False


def page_rank(graph: CSRGraph, degrees, page_rank, damping_factor=0.85, threshold=1e-4, max_iters=20,
              num_threads=None, *, device: int = 0) -> None:
    """Drop-in for nw::graph::page_rank (page_rank.hpp:89-91).

    ``page_rank`` must be a pre-sized array of graph.size() entries (the
    reference writes without resizing, page_rank.hpp:102); it is filled in
    place.  Its dtype plays the role of the reference's Real: a float64
    array gets the FP64 solve, anything else the FP32 one.
    ``num_threads`` is accepted and ignored, as in the reference.  Dangling
    vertices leak their mass, as in the reference.  The device graph is kept
    between calls on the same arrays (nwgr_pr_graph_cached), so a repeated
    call pays for the solve and the copy of the ranks only; an empty graph
    returns at once."""
    if len(page_rank) != graph.size():
        raise ValueError("page_rank must be pre-sized to graph.size()")
    if graph.n == 0:
        return
    fp64 = getattr(page_rank, "dtype", None) == np.float64
    out, direct = _output(page_rank, graph.n, np.float64 if fp64 else np.float32)
    p = _params(damping_factor, threshold, max_iters, "leak", 0, False)
    if not _cached_run(graph, degrees, device, _SOLVE_F64 if fp64 else _SOLVE_F32, p, out):
        with _resident(graph, degrees, device) as g:
            solve = g.run64 if fp64 else g.run
            solve(damping=damping_factor, tol=threshold, max_iters=max_iters, dangling="leak", out=out)
    if not direct:
        page_rank[:] = out


def page_rank_delta(graph: CSRGraph, degrees, page_rank, damping_factor=0.85, threshold=1e-4, max_iters=20,
                    num_threads=None, *, device: int = 0) -> None:
    """Drop-in for page_rank_v13 (experimental/algorithms/page_rank.hpp:
    632-634), the residual-push PageRank of pr.exe version 13; also returns
    what the serial page_rank_v3 (:566-568) returns.

    ``page_rank`` is a pre-sized array filled in place with the unnormalised
    ranks; ``num_threads`` is accepted and ignored.  The device graph is kept
    between calls as for page_rank."""
    if len(page_rank) != graph.size():
        raise ValueError("page_rank must be pre-sized to graph.size()")
    if graph.n == 0:
        return
    out, direct = _output(page_rank, graph.n)
    p = _params(damping_factor, threshold, max_iters, "leak", 0, False)  # the threshold as tol (run_delta)
    if not _cached_run(graph, degrees, device, _SOLVE_DELTA_F32, p, out):
        with _resident(graph, degrees, device) as g:
            g.run_delta(damping=damping_factor, threshold=threshold, max_iters=max_iters, out=out)
    if not direct:
        page_rank[:] = out


def cache_clear() -> None:
    """Destroy the resident graphs the drop-in calls keep (nwgr_pr_cache_clear):
    needed after a graph's arrays are edited in place."""
    pagerank_lib().nwgr_pr_cache_clear()


def cache_stats() -> dict:
    """Hits and misses of the resident-graph cache, and handles held."""
    h, m, e = ctypes.c_uint64(), ctypes.c_uint64(), ctypes.c_uint32()
    pagerank_lib().nwgr_pr_cache_stats(ctypes.byref(h), ctypes.byref(m), ctypes.byref(e))
    return {"hits": h.value, "misses": m.value, "entries": e.value}


def cache_set_capacity(entries: int) -> None:
    """Most resident graphs kept (1 .. 64)."""
    check(pagerank_lib().nwgr_pr_cache_set_capacity(int(entries)))


def build_csr_device(src, dst, n: int, dedup: bool = True, drop_self_loops: bool = True, stream=None):
    """In-CSR of the device edge list src[i] -> dst[i] (torch int32/uint32
    tensors on the GPU), built on the device (``nwgr_csr_build_device``).

    Returns (offsets, indices, out_degree) as torch int32 tensors holding the
    uint32 values; indices has exactly the kept entries."""
    import torch
    if src.shape != dst.shape or src.dim() != 1:
        raise ValueError("src and dst must be 1-D and of equal length")
    m = int(src.numel())
    dev = src.device
    src = src.contiguous().to(torch.int32)
    dst = dst.contiguous().to(torch.int32)
    offs = torch.empty(n + 1, dtype=torch.int32, device=dev)
    idx = torch.empty(max(m, 1), dtype=torch.int32, device=dev)
    deg = torch.empty(max(n, 1), dtype=torch.int32, device=dev)
    m_out = ctypes.c_uint64()
    flags = (1 if dedup else 0) | (2 if drop_self_loops else 0)
    check(pagerank_lib().nwgr_csr_build_device(src.data_ptr() if m else None, dst.data_ptr() if m else None, m, n,
                                               flags, offs.data_ptr(), idx.data_ptr(), deg.data_ptr(),
                                               ctypes.byref(m_out), stream))
    return offs, idx[:m_out.value], deg[:n]


def graph_from_edges_device(src, dst, n: int, dedup: bool = True, drop_self_loops: bool = True, device: int = 0,
                            stream=None) -> "PageRankGraph":
    """PageRankGraph built on the device straight from a device edge list."""
    h = ctypes.c_void_p()
    m = int(src.numel())
    flags = (1 if dedup else 0) | (2 if drop_self_loops else 0)
    check(pagerank_lib().nwgr_pr_graph_create_from_edges_device(
        src.data_ptr() if m else None, dst.data_ptr() if m else None, m, n, flags, device, stream, ctypes.byref(h)))
    return PageRankGraph(_handle=h)


def top_k(ranks, k: int, device: int = 0):
    """The k best-ranked vertices, by non-increasing rank and ascending id
    among ties (nwgr_top_k_f32; the verification of bench/pr.cpp:47-52):
    returns (ids uint32[k], ranks float32[k])."""
    r = np.ascontiguousarray(ranks, dtype=np.float32)
    if not 0 <= k <= len(r):
        raise ValueError("k must be in [0, len(ranks)]")
    ids = np.empty(k, np.uint32)
    vals = np.empty(k, np.float32)
    check(pagerank_lib().nwgr_top_k_f32(r.ctypes.data, len(r), k, ids.ctypes.data, vals.ctypes.data, device))
    return ids, vals


def top_k_device(ranks_ptr: int, n: int, k: int, ids_ptr: int, vals_ptr: int | None = None, stream=None) -> None:
    """Device form of top_k on caller-owned device arrays (e.g. torch tensors)."""
    check(pagerank_lib().nwgr_top_k_device_f32(ranks_ptr, n, k, ids_ptr, vals_ptr, stream))
