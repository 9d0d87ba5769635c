"""The C-ABI libraries load and export what include/*.h declares (no GPU needed)."""
import ctypes
import os
import shutil
import subprocess

import numpy as np
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


# ---- inputs the reference accepts: an empty graph; checks of caller vectors


def _empty_handle(L):
    off = np.zeros(1, np.uint32)
    h = ctypes.c_void_p()
    assert L.nwgr_pr_graph_create(off.ctypes.data, None, 0, 0, None, 0, ctypes.byref(h)) == _lib.NWGR_OK
    assert h.value
    return h


def test_empty_graph_returns_normally_without_a_device():
    """N = 0: the reference's page_rank initialises nothing, reduces an empty
    range (error 0) and returns after its first iteration
    (page_rank.hpp:92-147); the C ABI does the same and needs no GPU."""
    L = _lib.pagerank_lib()
    h = _empty_handle(L)
    try:
        assert L.nwgr_pr_graph_n(h) == 0 and L.nwgr_pr_graph_m(h) == 0
        assert L.nwgr_pr_graph_device_bytes(h) == 0
        p = _lib.PrParams()
        L.nwgr_pr_params_default(ctypes.byref(p))
        for fn in (L.nwgr_pr_run_f32, L.nwgr_pr_run_f64, L.nwgr_pr_delta_run_f32):
            res = _lib.PrResult()
            assert fn(h, ctypes.byref(p), None, ctypes.byref(res)) == _lib.NWGR_OK, L.nwgr_last_error()
            assert (res.iters, res.converged, res.error) == (1, 1, 0.0)
        res = _lib.PrResult()
        assert L.nwgr_pr_run_device_f32(h, ctypes.byref(p), None, ctypes.byref(res), None) == _lib.NWGR_OK
        assert res.iters == 1
        # no sweeps asked for: none performed; tol 0 is never met
        p.max_iters = 0
        assert L.nwgr_pr_run_f32(h, ctypes.byref(p), None, ctypes.byref(res)) == _lib.NWGR_OK
        assert (res.iters, res.converged) == (0, 0)
        p.max_iters, p.tol = 7, 0.0
        assert L.nwgr_pr_run_f32(h, ctypes.byref(p), None, ctypes.byref(res)) == _lib.NWGR_OK
        assert (res.iters, res.converged) == (7, 0)
        # tuning is kept and checked without a layout to re-plan
        t = _lib.PrTuning()
        assert L.nwgr_pr_graph_get_tuning(h, ctypes.byref(t)) == _lib.NWGR_OK
        t.chunk = 0
        assert L.nwgr_pr_graph_set_tuning(h, ctypes.byref(t)) == _lib.NWGR_EINVAL
        # a bad parameter is still an error on an empty graph
        p.damping = float("nan")
        assert L.nwgr_pr_run_f32(h, ctypes.byref(p), None, ctypes.byref(res)) == _lib.NWGR_EINVAL
    finally:
        L.nwgr_pr_graph_destroy(h)


def test_empty_graph_with_entries_is_einval():
    L = _lib.pagerank_lib()
    off = np.zeros(1, np.uint32)
    h = ctypes.c_void_p()
    assert L.nwgr_pr_graph_create(off.ctypes.data, off.ctypes.data, 0, 1, None, 0, ctypes.byref(h)) == _lib.NWGR_EINVAL


def test_entry_count_limit_and_sort_pass_bounds():
    """m up to 2^32 - 2^28 is accepted (the uint32 offsets of
    util/defaults.hpp:23-24 with room for the tile padding); beyond it, and
    n >= 2^31, the create call refuses before reading the arrays or touching
    a device.  The pass size of the device sorts is bounded to [2^16, 2^30]."""
    L = _lib.pagerank_lib()
    off = np.zeros(2, np.uint32)
    h = ctypes.c_void_p()
    too_many = (1 << 32) - (1 << 28) + 1
    assert L.nwgr_pr_graph_create(off.ctypes.data, off.ctypes.data, 1, too_many, None, 0,
                                  ctypes.byref(h)) == _lib.NWGR_EINVAL
    assert "2^32 - 2^28" in L.nwgr_last_error().decode()
    assert L.nwgr_pr_graph_create(off.ctypes.data, off.ctypes.data, 1 << 31, 0, None, 0,
                                  ctypes.byref(h)) == _lib.NWGR_EINVAL
    for bad in (0, (1 << 16) - 1, (1 << 30) + 1):
        assert L.nwgr_pr_set_sort_batch(bad) == _lib.NWGR_EINVAL
    assert L.nwgr_pr_set_sort_batch(1 << 30) == _lib.NWGR_OK


def test_empty_graph_through_the_cache_and_the_mirror():
    import nwgr_amd
    g = nwgr_amd.CSRGraph(np.zeros(1, np.uint32), np.zeros(0, np.uint32))
    deg = np.zeros(0, np.uint32)
    before = nwgr_amd.cache_stats()
    pg, hit = nwgr_amd.PageRankGraph.cached(g, deg)
    assert pg.n == 0 and not hit
    r, res = pg.run(tol=1e-4, max_iters=20)
    assert len(r) == 0 and res.iters == 1 and res.converged
    out = np.zeros(0, np.float64)
    assert nwgr_amd.page_rank(g, deg, out, 0.85, 1e-4, 20) is None
    assert nwgr_amd.cache_stats()["misses"] == before["misses"] + 1
    nwgr_amd.cache_clear()
    assert nwgr_amd.cache_stats()["entries"] == 0


@pytest.mark.parametrize("bad", [float("nan"), -1e-9, float("inf"), -float("inf")])
@pytest.mark.parametrize("field", ["teleport", "init"])
def test_params_check_rejects_bad_vectors(field, bad):
    L = _lib.pagerank_lib()
    n = 1000
    v = np.full(n, 1.0 / n, np.float32)
    p = _lib.PrParams()
    L.nwgr_pr_params_default(ctypes.byref(p))
    setattr(p, field, v.ctypes.data)
    assert L.nwgr_pr_params_check(ctypes.byref(p), n) == _lib.NWGR_OK
    v[n - 1 if bad != bad else 17] = bad
    assert L.nwgr_pr_params_check(ctypes.byref(p), n) == _lib.NWGR_EINVAL
    assert field.encode() in L.nwgr_last_error()


def test_params_check_scalars():
    L = _lib.pagerank_lib()
    p = _lib.PrParams()
    L.nwgr_pr_params_default(ctypes.byref(p))
    assert L.nwgr_pr_params_check(ctypes.byref(p), 0) == _lib.NWGR_OK
    for field, val in (("damping", 1.5), ("damping", float("nan")), ("tol", float("nan")), ("dangling", 3)):
        q = _lib.PrParams()
        L.nwgr_pr_params_default(ctypes.byref(q))
        setattr(q, field, val)
        assert L.nwgr_pr_params_check(ctypes.byref(q), 0) == _lib.NWGR_EINVAL, field
    q = _lib.PrParams()
    L.nwgr_pr_params_default(ctypes.byref(q))
    q.dangling = 2  # TELEPORT needs a vector
    assert L.nwgr_pr_params_check(ctypes.byref(q), 0) == _lib.NWGR_EINVAL
    assert L.nwgr_pr_params_check(None, 0) == _lib.NWGR_EINVAL


def test_cache_capacity_bounds():
    L = _lib.pagerank_lib()
    assert L.nwgr_pr_cache_set_capacity(0) == _lib.NWGR_EINVAL
    assert L.nwgr_pr_cache_set_capacity(65) == _lib.NWGR_EINVAL
    assert L.nwgr_pr_cache_set_capacity(2) == _lib.NWGR_OK


def test_cpp_dropin_empty_graph_and_double(tmp_path):
    """page_rank<Graph, float / double> on an empty graph returns at once, as
    the reference does for N = 0 (no device is touched)."""
    gxx = shutil.which("g++")
    if not gxx:
        pytest.skip("g++ missing")
    src = tmp_path / "empty.cpp"
    src.write_text(r'''
#include <cstdio>
#include <vector>
#include "nwgr/page_rank.hpp"
int main() {
  std::vector<uint32_t> offs{0}, idx, deg;
  nw::graph::csr_view g{offs.data(), idx.data(), 0, 0};
  std::vector<float> rf;
  std::vector<double> rd;
  nw::graph::page_rank(g, deg, rf, 0.85f, 1e-4f, 20, 1);
  nw::graph::page_rank(g, deg, rd, 0.85, 1e-7, 100, 1);
  nw::graph::page_rank_v13(g, deg, rf, 0.85f, 1e-4f, 20, 1);
  std::puts("ok");
  return 0;
}
''')
    exe = tmp_path / "empty"
    r = subprocess.run([gxx, "-std=c++20", "-O1", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe),
                        "-L", os.path.dirname(_lib.PAGERANK_SO), "-lnwgr_pagerank",
                        "-Wl,-rpath," + os.path.dirname(_lib.PAGERANK_SO), "-Wl,--allow-shlib-undefined"],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=60)
    assert out.returncode == 0 and out.stdout.strip() == "ok", out.stderr
