"""Graph ingestion (include/nwgr_io.h): MatrixMarket and the "NW" binary
formats, checked against independent readers (scipy.io.mmread, a plain
line-by-line restatement of mm_fill) on data files the reference's own tests
use (data/*.mtx, copied to tests/golden/mtx/), and byte by byte against the
layout the reference's serialize writes.
"""
import json
import os
import struct

import numpy as np
import pytest
import scipy.io

import nwgr_amd
from tests.conftest import GOLDEN

MTX = os.path.join(GOLDEN, "mtx")
FILES = ["karate.mtx", "USAir97.mtx", "Erdos981.mtx", "dwt___66.mtx", "tree.mmio", "spmatA.mmio"]


def mm_fill_lines(path):
    """mm_fill for directed edge lists (io/mmio.hpp:95-128), line by line."""
    with open(path) as f:
        head = f.readline().split()
        symmetric, pattern = head[4] == "symmetric", head[3] == "pattern"
        line = f.readline()
        while line.startswith("%"):
            line = f.readline()
        n0, n1, nnz = map(int, line.split())
        src, dst, val = [], [], []
        for _ in range(nnz):
            w = f.readline().split()
            i, j = int(w[0]), int(w[1])
            v = 1.0 if pattern else float(w[2])
            src.append(i - 1); dst.append(j - 1); val.append(v)
            if symmetric and i != j:
                src.append(j - 1); dst.append(i - 1); val.append(v)
    n = max([n0] + [max(s, d) + 1 for s, d in zip(src, dst)])
    return n, np.array(src), np.array(dst), np.array(val), pattern


@pytest.mark.parametrize("name", FILES)
def test_read_mm_matches_line_parser_and_scipy(name):
    path = os.path.join(MTX, name)
    el = nwgr_amd.read_mm(path)
    n, src, dst, val, pattern = mm_fill_lines(path)
    assert el.n == n and el.directed
    np.testing.assert_array_equal(el.src, src)
    np.testing.assert_array_equal(el.dst, dst)
    if pattern:
        assert el.values is None
    else:
        np.testing.assert_array_equal(el.values, val)
    # scipy's reader: the same matrix (symmetric files mirrored), as a multiset
    A = scipy.io.mmread(path).tocoo()
    want = sorted(zip(A.row.tolist(), A.col.tolist()))
    got = sorted(zip(el.src.tolist(), el.dst.tolist()))
    assert got == want


def test_karate_mtx_is_the_reference_test_graph():
    """data/karate.mtx gives the 156 directed edges page_rank_test.cpp:42-54 lists."""
    with open(os.path.join(GOLDEN, "karate_pagerank.json")) as f:
        d = json.load(f)
    el = nwgr_amd.read_mm(os.path.join(MTX, "karate.mtx"))
    assert el.n == d["n"]
    assert sorted(zip(el.src.tolist(), el.dst.tolist())) == sorted(zip(d["src"], d["dst"]))


def test_parallel_parse_equals_serial(tmp_path):
    """Per-thread byte ranges cut at line starts: any thread count, same list."""
    g = nwgr_amd.rmat(13, 8)
    src = g.indices.astype(np.int64)
    dst = np.repeat(np.arange(g.n), np.diff(g.offsets))
    p = tmp_path / "g.mtx"
    with open(p, "w") as f:
        f.write("%%MatrixMarket matrix coordinate pattern general\n% comment\n")
        f.write(f"{g.n} {g.n} {len(src)}\n")
        f.write("".join(f"{s + 1} {t + 1}\n" for s, t in zip(src, dst)))
    ref = nwgr_amd.read_mm(str(p), nthreads=1)
    for T in (2, 3, 8, 64):
        el = nwgr_amd.read_mm(str(p), nthreads=T)
        np.testing.assert_array_equal(el.src, ref.src)
        np.testing.assert_array_equal(el.dst, ref.dst)
    np.testing.assert_array_equal(ref.src, src)
    np.testing.assert_array_equal(ref.dst, dst)


@pytest.mark.parametrize("text,msg", [
    ("%%MatrixMarket matrix array real general\n2 2\n1\n2\n3\n4\n", "coordinate"),
    ("%%MatrixMarket matrix coordinate real hermitian\n2 2 1\n1 1 1.0\n", "symmetry"),
    ("%%MatrixMarket matrix coordinate pattern general\n2 2 3\n1 2\n2 1\n", "entries"),
    ("%%MatrixMarket matrix coordinate real general\n2 2 1\n1 2\n", "value"),
    ("%%MatrixMarket matrix coordinate pattern general\n2 2 1\n0 1\n", "malformed"),
    ("not a matrix\n", "MatrixMarket"),
])
def test_read_mm_rejects_bad_files(tmp_path, text, msg):
    p = tmp_path / "bad.mtx"
    p.write_text(text)
    with pytest.raises(ValueError) as e:
        nwgr_amd.read_mm(str(p))
    assert msg in str(e.value)


def test_nw_edge_list_round_trip_and_layout(tmp_path):
    el = nwgr_amd.read_mm(os.path.join(MTX, "Erdos981.mtx"))
    p = tmp_path / "e.el"
    nwgr_amd.write_nw_edge_list(str(p), el)
    raw = p.read_bytes()
    # index_edge_list::serialize: magic[27] (with its NUL), directedness (int),
    # graph_base: size_t N, bool is_open; struct_of_arrays: size_t count,
    # size_t arrays, then per array size_t count, size_t element size, data
    assert raw[:27] == b"NW Graph index_edge_list\0\0\0"
    d, N, is_open, cnt, arrays = struct.unpack_from("<iQ?QQ", raw, 27)
    assert (d, N, is_open, cnt, arrays) == (1, el.n, False, el.m, 2)
    off = 27 + struct.calcsize("<iQ?QQ")
    for a in (el.src, el.dst):
        c, esz = struct.unpack_from("<QQ", raw, off)
        assert (c, esz) == (el.m, 4)
        off += 16
        np.testing.assert_array_equal(np.frombuffer(raw, np.uint32, el.m, off), a)
        off += 4 * el.m
    assert off == len(raw)
    back = nwgr_amd.read_nw_edge_list(str(p))
    assert back.n == el.n and back.directed
    np.testing.assert_array_equal(back.src, el.src)
    np.testing.assert_array_equal(back.dst, el.dst)
    # load_graph tells the formats apart by the first word (bench/common.hpp:82-101)
    assert nwgr_amd.load_graph(str(p)).m == el.m
    assert nwgr_amd.load_graph(os.path.join(MTX, "Erdos981.mtx")).m == el.m


def test_nw_adjacency_round_trip_and_layout(tmp_path):
    el = nwgr_amd.read_mm(os.path.join(MTX, "USAir97.mtx"))
    g = nwgr_amd.adjacency_from_edge_list(el)
    p = tmp_path / "g.adj"
    nwgr_amd.write_nw_adjacency(str(p), g)
    raw = p.read_bytes()
    assert raw[:34] == b"NW GRAPH indexed_struct_of_arrays\0"
    N, cnt, esz = struct.unpack_from("<QQQ", raw, 34)
    assert (N, cnt, esz) == (g.n, g.n + 1, 4)
    back = nwgr_amd.read_nw_adjacency(str(p))
    np.testing.assert_array_equal(back.offsets, g.offsets)
    np.testing.assert_array_equal(back.indices, g.indices)
    p.write_bytes(raw[:-4])  # truncated
    with pytest.raises(ValueError):
        nwgr_amd.read_nw_adjacency(str(p))


def _int_prefix(field):
    """What `stringstream >> int` takes from a value field (0 if no digits)."""
    import re
    m = re.match(r"\s*([+-]?\d+)", field)
    return int(m.group(1)) if m else 0


def test_read_mm_undirected_and_int_attribute(tmp_path):
    text = ("%%MatrixMarket matrix coordinate real symmetric\n% values as the int reader sees them\n"
            "4 4 6\n2 1 3.5\n3 1 -2.7\n3 3 1e3\n4 2 .5\n4 3 +7\n4 1 -.5\n")
    p = tmp_path / "v.mtx"
    p.write_text(text)
    fields = [ln.split(None, 2)[2] for ln in text.splitlines()[3:]]
    want_int = [_int_prefix(f) for f in fields]
    pairs = [tuple(int(x) - 1 for x in ln.split()[:2]) for ln in text.splitlines()[3:]]
    # undirected: every entry once, directedness 0
    u = nwgr_amd.read_mm(str(p), directed=False, int_attribute=True)
    assert not u.directed and u.n == 4
    assert list(zip(u.src.tolist(), u.dst.tolist())) == pairs
    assert u.ints.tolist() == want_int == [3, -2, 1, 0, 7, 0]
    np.testing.assert_array_equal(u.values, [float(f) for f in fields])
    # directed: off-diagonal entries mirrored right after, attribute repeated
    d = nwgr_amd.read_mm(str(p), int_attribute=True)
    exp_pairs, exp_int = [], []
    for (i, j), v in zip(pairs, want_int):
        exp_pairs.append((i, j)); exp_int.append(v)
        if i != j:
            exp_pairs.append((j, i)); exp_int.append(v)
    assert d.directed and list(zip(d.src.tolist(), d.dst.tolist())) == exp_pairs
    assert d.ints.tolist() == exp_int
    # without the flag no attribute is kept
    assert nwgr_amd.read_mm(str(p)).ints is None
    # a pattern file's attribute is 1 (mm_fill's `T v(1.0)`)
    q = tmp_path / "p.mtx"
    q.write_text("%%MatrixMarket matrix coordinate pattern general\n3 3 2\n1 2\n3 1\n")
    assert nwgr_amd.read_mm(str(q), int_attribute=True).ints.tolist() == [1, 1]


def test_process_edge_list_converter_layout(tmp_path):
    """-d / -u outputs: edge_list<..., int>::serialize with three arrays, read
    back by load_graph (the first two arrays are the ids)."""
    from nwgr_amd.process_edge_list import main as convert
    src = os.path.join(MTX, "USAir97.mtx")
    dpath, upath = tmp_path / "d.el", tmp_path / "u.el"
    assert convert(["-d", str(dpath), "-u", str(upath), src]) == 0
    for path, directed in ((dpath, True), (upath, False)):
        el = nwgr_amd.read_mm(src, directed=directed, int_attribute=True)
        raw = path.read_bytes()
        assert raw[:27] == b"NW Graph index_edge_list\0\0\0"
        d, N, is_open, cnt, arrays = struct.unpack_from("<iQ?QQ", raw, 27)
        assert (d, N, is_open, cnt, arrays) == (1 if directed else 0, el.n, False, el.m, 3)
        off = 27 + struct.calcsize("<iQ?QQ")
        for a, dt in ((el.src, np.uint32), (el.dst, np.uint32), (el.ints, np.int32)):
            c, esz = struct.unpack_from("<QQ", raw, off)
            assert (c, esz) == (el.m, 4)
            np.testing.assert_array_equal(np.frombuffer(raw, dt, el.m, off + 16), a)
            off += 16 + 4 * el.m
        assert off == len(raw)
        back = nwgr_amd.load_graph(str(path))
        assert back.directed == directed
        np.testing.assert_array_equal(back.src, el.src)
        np.testing.assert_array_equal(back.ints, el.ints)
    # USAir97 holds real weights below 1: their int attribute is 0
    assert not nwgr_amd.read_nw_edge_list(str(dpath)).ints.any()


def test_nw_edge_list_skips_wider_attributes(tmp_path):
    """A third array of another element size (e.g. double weights) is skipped."""
    el = nwgr_amd.read_mm(os.path.join(MTX, "karate.mtx"))
    p = tmp_path / "w.el"
    nwgr_amd.write_nw_edge_list(str(p), el)
    raw = bytearray(p.read_bytes())
    hdr = 27 + struct.calcsize("<iQ?")
    struct.pack_into("<Q", raw, hdr + 8, 3)  # three arrays
    raw += struct.pack("<QQ", el.m, 8) + np.arange(el.m, dtype=np.float64).tobytes()
    p.write_bytes(bytes(raw))
    back = nwgr_amd.read_nw_edge_list(str(p))
    np.testing.assert_array_equal(back.src, el.src)
    assert back.ints is None
