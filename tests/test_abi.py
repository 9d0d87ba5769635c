"""The C-ABI libraries load and export what include/*.h declares (no GPU needed)."""
import ctypes
import os
import shutil
import subprocess

import pytest

from nwgr_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_pagerank_header_symbols_exported():
    L = _lib.pagerank_lib()
    names = _lib.declared_functions("nwgr_pagerank.h")
    assert "nwgr_pr_run_f32" in names and len(names) >= 12
    for name in names:
        assert hasattr(L, name), name


def test_graphgen_header_symbols_exported():
    L = _lib.graphgen_lib()
    names = _lib.declared_functions("nwgr_graphgen.h")
    for name in names:
        assert hasattr(L, name), name


def test_io_header_symbols_exported():
    L = _lib.graphgen_lib()
    names = _lib.declared_functions("nwgr_io.h")
    assert "nwgr_read_mm" in names and "nwgr_read_nw_adjacency" in names
    for name in names:
        assert hasattr(L, name), name


def test_every_header_is_checked():
    """Every C header under include/ belongs to one of the two libraries."""
    hdrs = sorted(f for f in os.listdir(os.path.join(ROOT, "include")) if f.endswith(".h"))
    assert hdrs == ["nwgr_graphgen.h", "nwgr_io.h", "nwgr_pagerank.h"]
    assert "nwgr_top_k_f32" in _lib.declared_functions("nwgr_pagerank.h")


def test_library_is_gfx950_hip_code():
    """The product library carries gfx950 device code (no other target)."""
    with open(_lib.PAGERANK_SO, "rb") as f:
        blob = f.read()
    assert b"amdgcn-amd-amdhsa--gfx950" in blob
    for other in (b"amdgcn-amd-amdhsa--gfx942", b"amdgcn-amd-amdhsa--gfx90a"):
        assert other not in blob


def test_params_default_matches_pr_driver():
    """bench/pr.cpp:26-27 defaults: 20 iterations, tolerance 1e-4; damping 0.85 (pr.cpp:137)."""
    p = _lib.PrParams()
    _lib.pagerank_lib().nwgr_pr_params_default(ctypes.byref(p))
    assert p.damping == 0.85 and p.tol == 1e-4 and p.max_iters == 20 and p.dangling == 0
    assert p.teleport is None and p.init is None


def test_device_count_never_fails():
    assert _lib.pagerank_lib().nwgr_device_count() >= 0


def test_no_device_is_reported_as_error():
    L = _lib.pagerank_lib()
    if L.nwgr_device_count() > 0:
        pytest.skip("a GPU is present")
    import numpy as np
    off = np.array([0, 1, 1], np.uint32)
    idx = np.array([1], np.uint32)
    h = ctypes.c_void_p()
    rc = L.nwgr_pr_graph_create(off.ctypes.data, idx.ctypes.data, 2, 1, None, 0, ctypes.byref(h))
    assert rc == _lib.NWGR_EDEVICE
    assert b"device" in L.nwgr_last_error()
    assert not h.value


def test_null_arguments_are_einval():
    L = _lib.pagerank_lib()
    h = ctypes.c_void_p()
    rc = L.nwgr_pr_graph_create(None, None, 4, 0, None, 0, ctypes.byref(h))
    assert rc == _lib.NWGR_EINVAL
    res = _lib.PrResult()
    p = _lib.PrParams()
    assert L.nwgr_pr_run_f32(None, ctypes.byref(p), None, ctypes.byref(res)) == _lib.NWGR_EINVAL


def test_cpp_dropin_header_compiles(tmp_path):
    """The C++20 drop-in compiles and links against the library."""
    gxx = shutil.which("g++")
    if not gxx:
        pytest.skip("g++ missing")
    exe = tmp_path / "dropin_pr"
    cmd = [gxx, "-std=c++20", "-O1", "-I", os.path.join(ROOT, "include"),
           os.path.join(ROOT, "tests", "cpp", "dropin_pr.cpp"), "-o", str(exe),
           "-L", os.path.dirname(_lib.PAGERANK_SO), "-lnwgr_pagerank",
           "-Wl,-rpath," + os.path.dirname(_lib.PAGERANK_SO),
           "-Wl,--allow-shlib-undefined"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_cpp_dropin_concept_accepts_nwgraph_like_graph(tmp_path):
    """A type with NWGraph's CSR members binds to the nw::graph::page_rank template."""
    gxx = shutil.which("g++")
    if not gxx:
        pytest.skip("g++ missing")
    src = tmp_path / "concept.cpp"
    src.write_text(r'''
#include <tuple>
#include <vector>
#include "nwgr/page_rank.hpp"
struct fake_adjacency {          // shape of indexed_struct_of_arrays (compressed.hpp:67-69)
  using vertex_id_type = unsigned;
  std::vector<unsigned> indices_;
  std::tuple<std::vector<unsigned>> to_be_indexed_;
  std::size_t size() const { return indices_.size() - 1; }
};
static_assert(nw::graph::nwgr_csr_graph<fake_adjacency>);
void call(const fake_adjacency& g, const std::vector<unsigned>& d, std::vector<float>& r) {
  nw::graph::page_rank(g, d, r, 0.85f, 1e-4f, 20, 1);
}
''')
    r = subprocess.run([gxx, "-std=c++20", "-fsyntax-only", "-I", os.path.join(ROOT, "include"), str(src)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
