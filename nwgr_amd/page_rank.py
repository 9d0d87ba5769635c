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

    def __init__(self, graph: CSRGraph | None = None, degrees=None, device: int = 0, *, _handle=None):
        L = pagerank_lib()
        self._h = ctypes.c_void_p()
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

    def close(self) -> None:
        if getattr(self, "_h", None) and self._h.value:
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
        row blocks in LDS), or None before any."""
        e = pagerank_lib().nwgr_pr_last_engine(self._h)
        return {0: "binned", 1: "tiled", 2: "tiled-accumulate"}.get(e)

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


def page_rank(graph: CSRGraph, degrees, page_rank, damping_factor=0.85, threshold=1e-4, max_iters=20,
              num_threads=None, *, device: int = 0) -> None:
    """Drop-in for nw::graph::page_rank (page_rank.hpp:89-91).

    ``page_rank`` must be a pre-sized float array of graph.size() entries
    (the reference writes without resizing, page_rank.hpp:102); it is filled
    in place.  ``num_threads`` is accepted and ignored, as in the reference.
    Dangling vertices leak their mass, as in the reference."""
    if len(page_rank) != graph.size():
        raise ValueError("page_rank must be pre-sized to graph.size()")
    with PageRankGraph(graph, degrees, device) as g:
        r, _ = g.run(damping=damping_factor, tol=threshold, max_iters=max_iters, dangling="leak")
    page_rank[:] = r


def page_rank_delta(graph: CSRGraph, degrees, page_rank, damping_factor=0.85, threshold=1e-4, max_iters=20,
                    num_threads=None, *, device: int = 0) -> None:
    """Drop-in for page_rank_v13 (experimental/algorithms/page_rank.hpp:
    632-634), the residual-push PageRank of pr.exe version 13; also returns
    what the serial page_rank_v3 (:566-568) returns.

    ``page_rank`` is a pre-sized float array filled in place with the
    unnormalised ranks; ``num_threads`` is accepted and ignored."""
    if len(page_rank) != graph.size():
        raise ValueError("page_rank must be pre-sized to graph.size()")
    with PageRankGraph(graph, degrees, device) as g:
        r, _ = g.run_delta(damping=damping_factor, threshold=threshold, max_iters=max_iters)
    page_rank[:] = r


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
