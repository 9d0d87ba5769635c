"""Weighted PageRank as a ``torch.autograd.Function`` over the C ABI.

r = weighted_pagerank(graph, w, t) solves
    r_v = (1-d) t_v + d (sum_{u->v} (w_uv / s_u) r_u + D q_v)
on the GPU (nwgr_wpr_forward_device_f32); backward solves the adjoint fixed
point on the device-built transpose and returns dL/dw and dL/dt
(nwgr_wpr_adjoint_device_f32).  Only the converged ranks are saved for
backward -- no per-sweep tape.  ``w`` is indexed like the graph's in-CSR
entries (edge k = indices[k] -> row of k), ``t`` like the vertices.
"""
from __future__ import annotations

import torch

from .page_rank import PageRankGraph


class _WeightedPageRank(torch.autograd.Function):
    @staticmethod
    def forward(ctx, w, t, graph, damping, tol, max_iters, dangling, adj_tol, adj_max_iters):
        if w.dtype != torch.float32 or not w.is_cuda:
            raise TypeError("w must be a float32 CUDA tensor")
        w = w.contiguous()
        tc = None if t is None else t.contiguous()
        if tc is not None and (tc.dtype != torch.float32 or tc.device != w.device):
            raise TypeError("t must be a float32 tensor on w's device")
        if w.numel() != graph.m or (tc is not None and tc.numel() != graph.n):
            raise ValueError("w needs one entry per edge and t one per vertex")
        r = torch.empty(graph.n, dtype=torch.float32, device=w.device)
        stream = torch.cuda.current_stream(w.device).cuda_stream
        res = graph.wpr_forward(w.data_ptr(), r.data_ptr(), damping, tol, max_iters, dangling,
                                None if tc is None else tc.data_ptr(), stream=stream)
        ctx.graph, ctx.damping, ctx.dangling = graph, damping, dangling
        ctx.adj_tol, ctx.adj_max_iters, ctx.has_t = adj_tol, adj_max_iters, t is not None
        ctx.forward_result = res
        graph.last_forward = res
        ctx.save_for_backward(w, tc if tc is not None else w.new_empty(0), r)
        return r

    @staticmethod
    def backward(ctx, grad_r):
        w, tc, r = ctx.saved_tensors
        g = grad_r.contiguous().to(torch.float32)
        graph = ctx.graph
        dw = torch.empty_like(w)
        dt = torch.empty(graph.n, dtype=torch.float32, device=w.device)
        # The adjoint is linear in dL/dr: solving it for g / ||g||_1 with the
        # absolute threshold adj_tol is the relative stop rule, and scaling the
        # gradients back by ||g||_1 gives dL/dw, dL/dt.  The norm stays on the
        # device (no host synchronisation before the adjoint is enqueued).
        scale = g.abs().sum().clamp_min(1e-30)
        gn = g / scale
        stream = torch.cuda.current_stream(w.device).cuda_stream
        res = graph.wpr_adjoint(w.data_ptr(), r.data_ptr(), gn.data_ptr(), dw.data_ptr(), dt.data_ptr(),
                                ctx.damping, ctx.adj_tol, ctx.adj_max_iters, ctx.dangling,
                                tc.data_ptr() if ctx.has_t else None, stream=stream)
        ctx.graph.last_adjoint = res
        dw.mul_(scale)
        dt.mul_(scale)
        return dw, (dt if ctx.has_t else None), None, None, None, None, None, None, None


def weighted_pagerank(graph: PageRankGraph, w: torch.Tensor, t: torch.Tensor | None = None, damping: float = 0.85,
                      tol: float = 1e-6, max_iters: int = 100, dangling: str = "leak", adj_tol: float = 1e-7,
                      adj_max_iters: int = 1000) -> torch.Tensor:
    """Differentiable weighted PageRank (see the module docstring).

    ``adj_tol`` is relative: the adjoint stops when the L1 change of lambda
    falls below adj_tol * ||dL/dr||_1."""
    return _WeightedPageRank.apply(w, t, graph, damping, tol, max_iters, dangling, adj_tol, adj_max_iters)
