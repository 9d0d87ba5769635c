"""ctypes bindings of the two native libraries.

``libnwgr_pagerank.so`` (HIP, include/nwgr_pagerank.h) is the product path;
there is no CPU fallback: when the library is missing or cannot be loaded,
``pagerank_lib()`` raises.  ``libnwgr_graphgen.so`` (host C++,
include/nwgr_graphgen.h) provides generators and the host CSR builder.
"""
from __future__ import annotations

import ctypes
import os
import re

_PKG = os.path.dirname(os.path.abspath(__file__))
_ROOT = os.path.dirname(_PKG)
PAGERANK_SO = os.path.join(_PKG, "libnwgr_pagerank.so")
GRAPHGEN_SO = os.path.join(_PKG, "libnwgr_graphgen.so")
INCLUDE_DIR = os.path.join(_ROOT, "include")

NWGR_OK, NWGR_EINVAL, NWGR_ENOMEM, NWGR_EDEVICE, NWGR_ENOTSUP = 0, 1, 2, 3, 4
STATUS_NAMES = {0: "NWGR_OK", 1: "NWGR_EINVAL", 2: "NWGR_ENOMEM", 3: "NWGR_EDEVICE", 4: "NWGR_ENOTSUP"}

DANGLING = {"leak": 0, "uniform": 1, "teleport": 2}


class NwgrError(RuntimeError):
    """A failed call into libnwgr_pagerank (carries the nwgr_status code)."""

    def __init__(self, code: int, msg: str):
        super().__init__(f"{STATUS_NAMES.get(code, code)}: {msg}")
        self.code = code


class PrParams(ctypes.Structure):
    _fields_ = [
        ("damping", ctypes.c_double),
        ("tol", ctypes.c_double),
        ("max_iters", ctypes.c_uint32),
        ("dangling", ctypes.c_uint32),
        ("teleport", ctypes.c_void_p),
        ("init", ctypes.c_void_p),
        ("chunk", ctypes.c_uint32),
        ("trace", ctypes.c_uint32),
    ]


class PrResult(ctypes.Structure):
    _fields_ = [
        ("iters", ctypes.c_uint32),
        ("converged", ctypes.c_uint32),
        ("error", ctypes.c_double),
        ("ms", ctypes.c_double),
    ]


class PrTuning(ctypes.Structure):
    _fields_ = [
        ("lds_cache", ctypes.c_uint32),
        ("hub_threshold", ctypes.c_uint32),
        ("hub_chunk", ctypes.c_uint32),
        ("unit_edges", ctypes.c_uint32),
        ("zero_rows_per_unit", ctypes.c_uint32),
        ("blocks_per_cu", ctypes.c_uint32),
        ("chunk", ctypes.c_uint32),
        ("engine", ctypes.c_uint32),
        ("tile_bits", ctypes.c_uint32),
        ("tile_fill_cost", ctypes.c_uint32),
        ("tile_combine", ctypes.c_uint32),
    ]


ENGINE_BINNED, ENGINE_TILED, ENGINE_AUTO = 0, 1, 2
COMBINE_GATHER, COMBINE_ACCUMULATE, COMBINE_AUTO = 0, 1, 2


_pr = None
_gen = None


def _declare_pagerank(L):
    vp, u32, u64, i32, f64 = ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_int, ctypes.c_double
    P = ctypes.POINTER
    L.nwgr_pr_params_default.restype = None
    L.nwgr_pr_params_default.argtypes = [P(PrParams)]
    L.nwgr_pr_graph_create.restype = i32
    L.nwgr_pr_graph_create.argtypes = [vp, vp, u32, u64, vp, i32, P(vp)]
    L.nwgr_pr_graph_create_device.restype = i32
    L.nwgr_pr_graph_create_device.argtypes = [vp, vp, u32, u64, vp, i32, vp, P(vp)]
    L.nwgr_pr_graph_destroy.restype = None
    L.nwgr_pr_graph_destroy.argtypes = [vp]
    L.nwgr_pr_set_sort_batch.restype = i32
    L.nwgr_pr_set_sort_batch.argtypes = [u64]
    L.nwgr_pr_graph_cached.restype = i32
    L.nwgr_pr_graph_cached.argtypes = [vp, vp, u32, u64, vp, i32, P(vp), P(u32)]
    L.nwgr_pr_share_balance.restype = i32
    L.nwgr_pr_share_balance.argtypes = [vp, u32, u32, P(ctypes.c_double), vp]
    L.nwgr_pr_cached_run.restype = i32
    L.nwgr_pr_cached_run.argtypes = [vp, vp, u32, u64, vp, i32, u32, P(PrParams), vp, P(PrResult), P(u32)]
    L.nwgr_pr_cache_clear.restype = None
    L.nwgr_pr_cache_clear.argtypes = []
    L.nwgr_pr_cache_set_capacity.restype = i32
    L.nwgr_pr_cache_set_capacity.argtypes = [u32]
    L.nwgr_pr_cache_stats.restype = None
    L.nwgr_pr_cache_stats.argtypes = [P(u64), P(u64), P(u32)]
    L.nwgr_pr_params_check.restype = i32
    L.nwgr_pr_params_check.argtypes = [P(PrParams), u32]
    L.nwgr_pr_graph_n.restype = u32
    L.nwgr_pr_graph_n.argtypes = [vp]
    L.nwgr_pr_graph_m.restype = u64
    L.nwgr_pr_graph_m.argtypes = [vp]
    L.nwgr_pr_graph_device_bytes.restype = u64
    L.nwgr_pr_graph_device_bytes.argtypes = [vp]
    L.nwgr_pr_graph_setup_ms.restype = f64
    L.nwgr_pr_graph_setup_ms.argtypes = [vp]
    L.nwgr_pr_graph_get_tuning.restype = i32
    L.nwgr_pr_graph_get_tuning.argtypes = [vp, P(PrTuning)]
    L.nwgr_pr_graph_set_tuning.restype = i32
    L.nwgr_pr_graph_set_tuning.argtypes = [vp, P(PrTuning)]
    L.nwgr_pr_run_f32.restype = i32
    L.nwgr_pr_run_f32.argtypes = [vp, P(PrParams), vp, P(PrResult)]
    L.nwgr_pr_run_device_f32.restype = i32
    L.nwgr_pr_run_device_f32.argtypes = [vp, P(PrParams), vp, P(PrResult), vp]
    L.nwgr_pr_gs_run_f32.restype = i32
    L.nwgr_pr_gs_run_f32.argtypes = [vp, P(PrParams), u32, vp, P(PrResult)]
    L.nwgr_pr_gs_run_device_f32.restype = i32
    L.nwgr_pr_gs_run_device_f32.argtypes = [vp, P(PrParams), u32, vp, P(PrResult), vp]
    L.nwgr_pr_share_run_device_f32.restype = i32
    L.nwgr_pr_share_run_device_f32.argtypes = [vp, P(PrParams), u32, u32, vp, P(PrResult), vp]
    L.nwgr_pr_gs_phases.restype = i32
    L.nwgr_pr_gs_phases.argtypes = [vp, u32, vp, vp]
    L.nwgr_pr_run_f64.restype = i32
    L.nwgr_pr_run_f64.argtypes = [vp, P(PrParams), vp, P(PrResult)]
    L.nwgr_pr_run_device_f64.restype = i32
    L.nwgr_pr_run_device_f64.argtypes = [vp, P(PrParams), vp, P(PrResult), vp]
    L.nwgr_pr_delta_run_f32.restype = i32
    L.nwgr_pr_delta_run_f32.argtypes = [vp, P(PrParams), vp, P(PrResult)]
    L.nwgr_pr_delta_run_device_f32.restype = i32
    L.nwgr_pr_delta_run_device_f32.argtypes = [vp, P(PrParams), vp, P(PrResult), vp]
    L.nwgr_pr_last_errors.restype = u32
    L.nwgr_pr_last_errors.argtypes = [vp, vp, u32]
    L.nwgr_pr_last_sweep_times.restype = u32
    L.nwgr_pr_last_sweep_times.argtypes = [vp, vp, u32]
    L.nwgr_pr_last_profile.restype = i32
    L.nwgr_pr_last_profile.argtypes = [vp, P(f64), P(u32), P(u32)]
    L.nwgr_pr_last_engine.restype = i32
    L.nwgr_pr_last_engine.argtypes = [vp]
    L.nwgr_top_k_device_f32.restype = i32
    L.nwgr_top_k_device_f32.argtypes = [vp, u32, u32, vp, vp, vp]
    L.nwgr_top_k_device_f32_noalloc.restype = i32
    L.nwgr_top_k_device_f32_noalloc.argtypes = [vp, u32, u32, vp, vp, vp]
    L.nwgr_top_k_f32.restype = i32
    L.nwgr_top_k_f32.argtypes = [vp, u32, u32, vp, vp, i32]
    L.nwgr_ppr_batch_f32.restype = i32
    L.nwgr_ppr_batch_f32.argtypes = [vp, P(PrParams), u32, vp, vp, P(PrResult)]
    L.nwgr_ppr_batch_device_f32.restype = i32
    L.nwgr_ppr_batch_device_f32.argtypes = [vp, P(PrParams), u32, vp, vp, P(PrResult), vp]
    L.nwgr_ppr_last_column_errors.restype = u32
    L.nwgr_ppr_last_column_errors.argtypes = [vp, vp, u32]
    L.nwgr_csr_build_device.restype = i32
    L.nwgr_csr_build_device.argtypes = [vp, vp, ctypes.c_uint64, u32, u32, vp, vp, vp, P(ctypes.c_uint64), vp]
    L.nwgr_pr_graph_create_from_edges_device.restype = i32
    L.nwgr_pr_graph_create_from_edges_device.argtypes = [vp, vp, ctypes.c_uint64, u32, u32, i32, vp, P(vp)]
    L.nwgr_multi_create.restype = i32
    L.nwgr_multi_create.argtypes = [vp, vp, u32, ctypes.c_uint64, vp, u32, vp, i32, P(vp)]
    L.nwgr_multi_destroy.restype = None
    L.nwgr_multi_destroy.argtypes = [vp]
    L.nwgr_multi_run_f32.restype = i32
    L.nwgr_multi_run_f32.argtypes = [vp, P(PrParams), vp, P(PrResult)]
    L.nwgr_multi_run_device_f32.restype = i32
    L.nwgr_multi_run_device_f32.argtypes = [vp, P(PrParams), vp, P(PrResult), vp]
    L.nwgr_multi_graph_results.restype = i32
    L.nwgr_multi_graph_results.argtypes = [vp, vp, vp]
    L.nwgr_multi_get_tuning.restype = i32
    L.nwgr_multi_get_tuning.argtypes = [vp, P(u32), P(u32)]
    L.nwgr_multi_set_tuning.restype = i32
    L.nwgr_multi_set_tuning.argtypes = [vp, u32, u32]
    L.nwgr_multi_n.restype = u32
    L.nwgr_multi_n.argtypes = [vp]
    L.nwgr_multi_count.restype = u32
    L.nwgr_multi_count.argtypes = [vp]
    L.nwgr_multi_setup_ms.restype = ctypes.c_double
    L.nwgr_multi_setup_ms.argtypes = [vp]
    L.nwgr_wpr_forward_device_f32.restype = i32
    L.nwgr_wpr_forward_device_f32.argtypes = [vp, P(PrParams), vp, vp, P(PrResult), vp]
    L.nwgr_wpr_adjoint_device_f32.restype = i32
    L.nwgr_wpr_adjoint_device_f32.argtypes = [vp, P(PrParams), vp, vp, vp, vp, vp, P(PrResult), vp]
    L.nwgr_last_error.restype = ctypes.c_char_p
    L.nwgr_last_error.argtypes = []
    L.nwgr_device_count.restype = i32
    L.nwgr_device_count.argtypes = []


def pagerank_lib():
    """The HIP library; raises if it is not built (no fallback exists)."""
    global _pr
    if _pr is None:
        if not os.path.exists(PAGERANK_SO):
            raise ImportError(
                f"{PAGERANK_SO} is missing: build it with `python -m nwgr_amd.build` "
                "(the PageRank path has no CPU fallback)")
        L = ctypes.CDLL(PAGERANK_SO, mode=ctypes.RTLD_GLOBAL)
        _declare_pagerank(L)
        _pr = L
    return _pr


def _declare_graphgen(L):
    vp, u32, u64, i32, f64 = ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_int, ctypes.c_double
    P = ctypes.POINTER
    L.nwgr_csr_build.restype = i32
    L.nwgr_csr_build.argtypes = [u32, u64, vp, vp, u32, i32, P(vp)]
    L.nwgr_gen_rmat.restype = i32
    L.nwgr_gen_rmat.argtypes = [u32, u32, u64, f64, f64, f64, i32, P(vp)]
    L.nwgr_gen_er.restype = i32
    L.nwgr_gen_er.argtypes = [u32, u64, u64, i32, P(vp)]
    L.nwgr_gen_rmat_n.restype = i32
    L.nwgr_gen_rmat_n.argtypes = [u32, u64, u64, f64, f64, f64, P(vp)]
    L.nwgr_gen_rmat_many.restype = i32
    L.nwgr_gen_rmat_many.argtypes = [u32, vp, u32, u64, f64, f64, f64, i32, P(vp)]
    L.nwgr_host_csr_n.restype = u32
    L.nwgr_host_csr_n.argtypes = [vp]
    L.nwgr_host_csr_m.restype = u64
    L.nwgr_host_csr_m.argtypes = [vp]
    for f in ("offsets", "indices", "outdeg"):
        fn = getattr(L, "nwgr_host_csr_" + f)
        fn.restype = ctypes.POINTER(ctypes.c_uint32)
        fn.argtypes = [vp]
    L.nwgr_host_csr_free.restype = None
    L.nwgr_host_csr_free.argtypes = [vp]
    L.nwgr_csr_transpose.restype = i32
    L.nwgr_csr_transpose.argtypes = [u32, u64, vp, vp, vp, vp, vp, i32]
    L.nwgr_splitmix64_at.restype = u64
    L.nwgr_splitmix64_at.argtypes = [u64, u64]
    L.nwgr_gen_last_error.restype = ctypes.c_char_p
    L.nwgr_gen_last_error.argtypes = []
    # graph ingestion (include/nwgr_io.h)
    cp = ctypes.c_char_p
    L.nwgr_read_mm.restype = i32
    L.nwgr_read_mm.argtypes = [cp, i32, P(vp)]
    L.nwgr_read_mm_as.restype = i32
    L.nwgr_read_mm_as.argtypes = [cp, i32, i32, i32, P(vp)]
    L.nwgr_read_nw_edge_list.restype = i32
    L.nwgr_read_nw_edge_list.argtypes = [cp, P(vp)]
    L.nwgr_write_nw_edge_list.restype = i32
    L.nwgr_write_nw_edge_list.argtypes = [cp, u32, u64, vp, vp, i32]
    L.nwgr_write_nw_edge_list_attr.restype = i32
    L.nwgr_write_nw_edge_list_attr.argtypes = [cp, u32, u64, vp, vp, vp, i32]
    L.nwgr_write_nw_adjacency.restype = i32
    L.nwgr_write_nw_adjacency.argtypes = [cp, u32, u64, vp, vp]
    L.nwgr_read_nw_adjacency.restype = i32
    L.nwgr_read_nw_adjacency.argtypes = [cp, P(vp)]
    L.nwgr_edge_list_n.restype = u32
    L.nwgr_edge_list_n.argtypes = [vp]
    L.nwgr_edge_list_m.restype = u64
    L.nwgr_edge_list_m.argtypes = [vp]
    L.nwgr_edge_list_directedness.restype = i32
    L.nwgr_edge_list_directedness.argtypes = [vp]
    for f in ("src", "dst"):
        fn = getattr(L, "nwgr_edge_list_" + f)
        fn.restype = ctypes.POINTER(ctypes.c_uint32)
        fn.argtypes = [vp]
    L.nwgr_edge_list_values.restype = ctypes.POINTER(ctypes.c_double)
    L.nwgr_edge_list_values.argtypes = [vp]
    L.nwgr_edge_list_int_values.restype = ctypes.POINTER(ctypes.c_int32)
    L.nwgr_edge_list_int_values.argtypes = [vp]
    L.nwgr_edge_list_free.restype = None
    L.nwgr_edge_list_free.argtypes = [vp]


def graphgen_lib():
    global _gen
    if _gen is None:
        if not os.path.exists(GRAPHGEN_SO):
            raise ImportError(f"{GRAPHGEN_SO} is missing: build it with `python -m nwgr_amd.build`")
        L = ctypes.CDLL(GRAPHGEN_SO)
        _declare_graphgen(L)
        _gen = L
    return _gen


def check(code: int) -> None:
    if code != NWGR_OK:
        msg = pagerank_lib().nwgr_last_error().decode(errors="replace")
        raise NwgrError(code, msg)


def declared_functions(header: str) -> list:
    """Function names declared in include/<header> (used by the ABI tests)."""
    with open(os.path.join(INCLUDE_DIR, header)) as f:
        text = f.read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    text = re.sub(r"//[^\n]*", "", text)
    names = re.findall(r"\b(nwgr_\w+)\s*\(", text)
    return sorted(set(names))
