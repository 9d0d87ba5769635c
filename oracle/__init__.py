"""CPU oracle for the PageRank hot path -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import this package.  It is the checker, never the thing
measured or shipped: the product (``nwgr_amd``) has no path into it.

The arithmetic lives in ``pr_oracle.c`` (plain C, built by ``oracle/Makefile``
into ``liborc.so`` and ``liborc_fast.so``); this module wraps it with numpy.
Reference semantics followed (file:line in NWGraph):

* ``jacobi``   -- ``page_rank_v1``/``page_rank_v10``
  (include/nwgraph/experimental/algorithms/page_rank.hpp:137-174, :450-505),
  plus teleport / dangling-mode extensions (LEAK == the reference).
* ``gs``       -- ``page_rank`` v11 on one thread
  (include/nwgraph/algorithms/page_rank.hpp:89-148).
* ``bgs``      -- block Gauss-Seidel: rows in a given order cut into phases,
  Jacobi inside a phase and in place between phases (one phase: ``jacobi``;
  one row per phase: v11's serial in-place sweep, page_rank.hpp:124-139).
* ``gs_par`` / ``jacobi_par`` -- v11 / v10 with OpenMP in place of TBB; the
  timed CPU baseline.
* ``wpr_forward`` / ``wpr_adjoint`` -- weighted PageRank and its adjoint in
  double (no reference counterpart; same iteration with edge weights).

Pinning (tests/test_oracle.py): karate golden vector of
test/page_rank_test.cpp:72-75, closed forms (edgeless graph, cycle, star),
dense linear solves, GS/Jacobi common fixed point, L1 contraction, central
finite differences of the adjoint.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))

LEAK, UNIFORM, TELEPORT = 0, 1, 2

_u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_vp = ctypes.c_void_p

_libs: dict = {}


def build() -> None:
    """Compile the oracle (gcc, oracle/Makefile)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def _lib(fast: bool = False):
    name = "liborc_fast.so" if fast else "liborc.so"
    if name in _libs:
        return _libs[name]
    path = os.path.join(_HERE, name)
    if not os.path.exists(path):
        build()
    L = ctypes.CDLL(path)
    u32, f64, f32, i32 = ctypes.c_uint32, ctypes.c_double, ctypes.c_float, ctypes.c_int
    L.orc_pr_jacobi_f64.restype = i32
    L.orc_pr_jacobi_f64.argtypes = [u32, _u32p, _u32p, _u32p, f64, f64, u32, i32, _vp, _vp, _f64p, _vp]
    L.orc_pr_jacobi_f64_par.restype = i32
    L.orc_pr_jacobi_f64_par.argtypes = [u32, _u32p, _u32p, _u32p, f64, f64, u32, i32, _vp, _vp, _f64p, _vp,
                                        i32]
    L.orc_pr_bgs_f64.restype = i32
    L.orc_pr_bgs_f64.argtypes = [u32, _u32p, _u32p, _u32p, f64, f64, u32, _u32p, u32, _u32p, _vp, _vp, _f64p, _vp,
                                 i32]
    for fn, real, realp in (("orc_pr_gs_f64", f64, _f64p), ("orc_pr_gs_f32", f32, _f32p)):
        getattr(L, fn).restype = i32
        getattr(L, fn).argtypes = [u32, _u32p, _u32p, _u32p, real, real, u32, realp, _vp]
    for fn, real, realp in (("orc_pr_gs_par_f64", f64, _f64p), ("orc_pr_gs_par_f32", f32, _f32p),
                            ("orc_pr_jacobi_par_f32", f32, _f32p)):
        getattr(L, fn).restype = i32
        getattr(L, fn).argtypes = [u32, _u32p, _u32p, _u32p, real, real, u32, i32, realp, _vp]
    for fn, realp in (("orc_pr_delta_f64", _f64p), ("orc_pr_delta_f32", _f32p)):
        getattr(L, fn).restype = i32
        getattr(L, fn).argtypes = [u32, _u32p, _u32p, _u32p, f64, f64, u32, i32, realp, _vp]
    L.orc_wpr_forward_f64.restype = i32
    L.orc_wpr_forward_f64.argtypes = [u32, _u32p, _u32p, _f64p, _f64p, f64, f64, u32, i32, _f64p, _vp]
    L.orc_wpr_forward_f64_par.restype = i32
    L.orc_wpr_forward_f64_par.argtypes = [u32, _u32p, _u32p, _f64p, _f64p, f64, f64, u32, i32, _vp, _f64p, i32]
    L.orc_wpr_adjoint_f64_par.restype = i32
    L.orc_wpr_adjoint_f64_par.argtypes = [u32, _u32p, _u32p, _f64p, _f64p, _f64p, _f64p, f64, f64, u32, i32,
                                          _f64p, _f64p, _f64p, i32]
    L.orc_wpr_adjoint_f64.restype = i32
    L.orc_wpr_adjoint_f64.argtypes = [u32, _u32p, _u32p, _f64p, _f64p, _f64p, _f64p, f64, f64, u32, i32,
                                      _f64p, _f64p, _f64p]
    _libs[name] = L
    return L


def _u32(a):
    return np.ascontiguousarray(a, dtype=np.uint32)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _ptr(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def jacobi(offsets, indices, deg, damping=0.85, tol=1e-4, max_iters=20, dangling=LEAK,
           teleport=None, init=None, return_hist=False, nthreads=1):
    """FP64 Jacobi PageRank; returns (r, sweeps[, err_hist]).

    nthreads > 1 splits rows over OpenMP threads (identical per-row sums)."""
    offsets, indices, deg = _u32(offsets), _u32(indices), _u32(deg)
    n = len(offsets) - 1
    r = np.empty(n, np.float64)
    tel = None if teleport is None else _f64(teleport)
    ini = None if init is None else _f64(init)
    hist = np.zeros(max(max_iters, 1), np.float64)
    if nthreads > 1:
        it = _lib().orc_pr_jacobi_f64_par(n, offsets, indices, deg, damping, tol, max_iters, dangling,
                                          _ptr(tel), _ptr(ini), r, _ptr(hist), nthreads)
    else:
        it = _lib().orc_pr_jacobi_f64(n, offsets, indices, deg, damping, tol, max_iters, dangling,
                                      _ptr(tel), _ptr(ini), r, _ptr(hist))
    return (r, it, hist[:it]) if return_hist else (r, it)


def bgs(offsets, indices, deg, damping=0.85, tol=1e-4, max_iters=20, order=None, cuts=None, teleport=None,
        init=None, return_hist=False, nthreads=1):
    """FP64 block Gauss-Seidel (orc_pr_bgs_f64), LEAK dangling: rows visited
    in `order` (default 0 .. n-1) cut into phases at positions `cuts`
    (default one phase); returns (r, sweeps[, err_hist]).  nthreads > 1
    splits each phase's rows over OpenMP threads (the same ranks)."""
    offsets, indices, deg = _u32(offsets), _u32(indices), _u32(deg)
    n = len(offsets) - 1
    order = np.arange(n, dtype=np.uint32) if order is None else _u32(order)
    cuts = np.array([0, n], np.uint32) if cuts is None else _u32(cuts)
    if len(order) != n or cuts[0] != 0 or cuts[-1] != n or np.any(np.diff(cuts.astype(np.int64)) < 0):
        raise ValueError("order must list the n rows and cuts run from 0 to n")
    r = np.empty(n, np.float64)
    tel = None if teleport is None else _f64(teleport)
    ini = None if init is None else _f64(init)
    hist = np.zeros(max(max_iters, 1), np.float64)
    it = _lib().orc_pr_bgs_f64(n, offsets, indices, deg, damping, tol, max_iters, order, len(cuts) - 1, cuts,
                               _ptr(tel), _ptr(ini), r, _ptr(hist), nthreads)
    return (r, it, hist[:it]) if return_hist else (r, it)


def gs(offsets, indices, deg, damping=0.85, tol=1e-4, max_iters=20, dtype=np.float64,
       return_hist=False):
    """Serial v11 Gauss-Seidel restatement (LEAK, uniform teleport)."""
    offsets, indices, deg = _u32(offsets), _u32(indices), _u32(deg)
    n = len(offsets) - 1
    r = np.empty(n, dtype)
    hist = np.zeros(max(max_iters, 1), np.float64)
    fn = _lib().orc_pr_gs_f64 if dtype == np.float64 else _lib().orc_pr_gs_f32
    it = fn(n, offsets, indices, deg, damping, tol, max_iters, r, _ptr(hist))
    return (r, it, hist[:it]) if return_hist else (r, it)


def gs_par(offsets, indices, deg, damping=0.85, tol=1e-4, max_iters=20, nthreads=1,
           dtype=np.float32, fast=True):
    """OpenMP v11 restatement (the CPU baseline). Returns (r, sweeps)."""
    offsets, indices, deg = _u32(offsets), _u32(indices), _u32(deg)
    n = len(offsets) - 1
    r = np.empty(n, dtype)
    L = _lib(fast)
    fn = L.orc_pr_gs_par_f32 if dtype == np.float32 else L.orc_pr_gs_par_f64
    it = fn(n, offsets, indices, deg, damping, tol, max_iters, nthreads, r, None)
    return r, it


def jacobi_par(offsets, indices, deg, damping=0.85, tol=1e-4, max_iters=20, nthreads=1, fast=True):
    """OpenMP v10 restatement in float. Returns (r, sweeps)."""
    offsets, indices, deg = _u32(offsets), _u32(indices), _u32(deg)
    n = len(offsets) - 1
    r = np.empty(n, np.float32)
    it = _lib(fast).orc_pr_jacobi_par_f32(n, offsets, indices, deg, damping, tol, max_iters, nthreads,
                                          r, None)
    return r, it


def delta(offsets, indices, deg, damping=0.85, threshold=1e-4, max_iters=20, dtype=np.float64, keep=False):
    """Residual ("delta") PageRank as page_rank_v13
    (experimental/algorithms/page_rank.hpp:632-698), or as page_rank_v3
    (:566-630, sub-threshold residuals kept until a delta arrives) with
    keep=True.  Returns
    (ranks, iterations, pushes) with pushes[i] the number of pushing vertices
    of non-zero degree in iteration i + 1.  dtype float32 runs the float
    restatement (the reference's operation order in float)."""
    offsets, indices, deg = _u32(offsets), _u32(indices), _u32(deg)
    n = len(offsets) - 1
    r = np.empty(n, dtype)
    pushes = np.zeros(max(max_iters, 1), np.int32)
    fn = _lib().orc_pr_delta_f64 if dtype == np.float64 else _lib().orc_pr_delta_f32
    if dtype == np.float32:
        damping = float(np.float32(damping))
        threshold = float(np.float32(threshold))
    it = fn(n, offsets, indices, deg, damping, threshold, max_iters, int(keep), r,
            pushes.ctypes.data_as(ctypes.c_void_p))
    if it < 0:
        raise MemoryError("oracle delta: allocation failed")
    return r, it, pushes[:it]


def wpr_forward(offsets, indices, w, t, damping=0.85, tol=1e-10, max_iters=1000, dangling=LEAK,
                nthreads=1, init=None):
    """FP64 weighted PageRank; returns (r, sweeps).  nthreads > 1 uses the
    OpenMP variant (same per-row sums); init warm-starts it."""
    offsets, indices = _u32(offsets), _u32(indices)
    n = len(offsets) - 1
    r = np.empty(n, np.float64)
    if nthreads > 1 or init is not None:
        ini = None if init is None else _f64(init)
        it = _lib().orc_wpr_forward_f64_par(n, offsets, indices, _f64(w), _f64(t), damping, tol, max_iters,
                                            dangling, _ptr(ini), r, max(nthreads, 1))
    else:
        it = _lib().orc_wpr_forward_f64(n, offsets, indices, _f64(w), _f64(t), damping, tol, max_iters,
                                        dangling, r, None)
    return r, it


def wpr_adjoint(offsets, indices, w, t, r, g, damping=0.85, tol=1e-12, max_iters=2000, dangling=LEAK,
                nthreads=1):
    """FP64 adjoint; returns (lambda, dL/dw, dL/dt, sweeps)."""
    offsets, indices = _u32(offsets), _u32(indices)
    n = len(offsets) - 1
    m = len(indices)
    lam = np.empty(n, np.float64)
    dw = np.empty(max(m, 1), np.float64)
    dt = np.empty(n, np.float64)
    fn = _lib().orc_wpr_adjoint_f64_par if nthreads > 1 else None
    if fn is not None:
        it = fn(n, offsets, indices, _f64(w), _f64(t), _f64(r), _f64(g), damping, tol, max_iters, dangling,
                lam, dw, dt, nthreads)
    else:
        it = _lib().orc_wpr_adjoint_f64(n, offsets, indices, _f64(w), _f64(t), _f64(r), _f64(g), damping,
                                        tol, max_iters, dangling, lam, dw, dt)
    return lam, dw[:m], dt, it


def out_degree(n, indices):
    """Out-degree as build_degrees counts it (bench/common.hpp:109-120)."""
    return np.bincount(_u32(indices), minlength=n).astype(np.uint32)


def top_k(r, k):
    """Indices of the k largest entries, ties broken by ascending vertex id
    (the reference sorts with an unstable std::sort, util/proxysort.hpp:52; the
    id tie-break makes the order well defined)."""
    r = np.asarray(r)
    order = np.lexsort((np.arange(len(r)), -r.astype(np.float64)))
    return order[:k]
