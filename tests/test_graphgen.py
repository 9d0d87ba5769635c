"""Synthetic generators and the host CSR builder (libnwgr_graphgen)."""
import json
import os

import numpy as np
import pytest

import nwgr_amd
from nwgr_amd.graph import splitmix64_at
from tests.conftest import GOLDEN

MASK = (1 << 64) - 1


def test_splitmix64_known_answers():
    """Vigna's SplitMix64 seeded with 0: published first outputs."""
    assert splitmix64_at(0, 0) == 0xE220A8397B1DCDAF
    assert splitmix64_at(0, 1) == 0x6E789E6AA1B965F4
    assert splitmix64_at(0, 2) == 0x06C45D188009454F


def _sm64(key, k):
    z = (key + (k + 1) * 0x9E3779B97F4A7C15) & MASK
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK
    return z ^ (z >> 31)


def _rmat_edges_py(scale, ef, seed, a=0.57, b=0.19, c=0.19):
    t1, t2, t3 = (int(round(x * 2 ** 32)) for x in (a, a + b, a + b + c))
    W = (scale + 1) // 2
    src, dst = [], []
    for e in range(ef << scale):
        u = v = 0
        for lvl in range(scale):
            word = _sm64(seed, e * W + lvl // 2)
            r = (word & 0xFFFFFFFF) if lvl & 1 else (word >> 32)
            bit = 1 << (scale - 1 - lvl)
            if r < t1:
                pass
            elif r < t2:
                v |= bit
            elif r < t3:
                u |= bit
            else:
                u |= bit
                v |= bit
        src.append(u)
        dst.append(v)
    return np.array(src), np.array(dst)


def _csr_py(n, src, dst, dedup, drop_loops):
    pairs = [(int(v), int(u)) for u, v in zip(src, dst) if not (drop_loops and u == v)]
    pairs = sorted(set(pairs)) if dedup else sorted(pairs)
    off = np.zeros(n + 1, np.int64)
    for v, _ in pairs:
        off[v + 1] += 1
    return np.cumsum(off).astype(np.uint32), np.array([u for _, u in pairs], np.uint32)


def test_sm64_python_matches_c():
    for key, k in [(27491095, 0), (27491095, 12345), (1, 2 ** 40)]:
        assert _sm64(key, k) == splitmix64_at(key, k)


@pytest.mark.parametrize("scale,ef", [(4, 4), (7, 3)])
def test_rmat_matches_spec(scale, ef):
    src, dst = _rmat_edges_py(scale, ef, 27491095)
    off, ind = _csr_py(1 << scale, src, dst, True, True)
    g = nwgr_amd.rmat(scale, ef)
    np.testing.assert_array_equal(g.offsets, off)
    np.testing.assert_array_equal(g.indices, ind)


def test_rmat_quadrant_marginals():
    """Before dedup, a source bit is 1 with prob c+d and a destination bit with b+d."""
    src, dst = _rmat_edges_py(6, 64, 7)
    sb = np.unpackbits(src.astype(np.uint8)[:, None], axis=1)[:, 2:]
    db = np.unpackbits(dst.astype(np.uint8)[:, None], axis=1)[:, 2:]
    assert abs(sb.mean() - 0.24) < 0.01 and abs(db.mean() - 0.24) < 0.01
    assert abs((sb & db).mean() - 0.05) < 0.006


def test_er_matches_spec():
    n, m, seed = 100, 1000, 5
    src, dst = [], []
    for e in range(m):
        w = _sm64(seed, e)
        src.append(((w >> 32) * n) >> 32)
        dst.append(((w & 0xFFFFFFFF) * n) >> 32)
    off, ind = _csr_py(n, src, dst, True, True)
    g = nwgr_amd.erdos_renyi(n, m, seed)
    np.testing.assert_array_equal(g.offsets, off)
    np.testing.assert_array_equal(g.indices, ind)


@pytest.mark.parametrize("dedup,drop", [(False, False), (True, False), (False, True), (True, True)])
def test_csr_builder_brute_force(dedup, drop):
    rng = np.random.default_rng(42)
    n = 50
    src = rng.integers(0, n, 600)
    dst = rng.integers(0, n, 600)
    g = nwgr_amd.build_adjacency(src, dst, n=n, dedup=dedup, drop_self_loops=drop, nthreads=4)
    off, ind = _csr_py(n, src, dst, dedup, drop)
    np.testing.assert_array_equal(g.offsets, off)
    np.testing.assert_array_equal(g.indices, ind)


def test_csr_builder_rejects_bad_ids():
    with pytest.raises(ValueError):
        nwgr_amd.build_adjacency([0, 5], [1, 2], n=4)


def test_build_degrees_counts_out_edges(karate):
    d, g, deg = karate
    np.testing.assert_array_equal(deg, np.bincount(d["src"], minlength=34))


def test_transpose_brute_force():
    g = nwgr_amd.rmat(8, 4)
    off, ind, perm = nwgr_amd.transpose(g)
    rows = np.repeat(np.arange(g.n), np.diff(g.offsets))
    # out-CSR entry k is edge indices[perm[k]] -> rows[perm[k]]
    np.testing.assert_array_equal(ind, rows[perm])
    np.testing.assert_array_equal(np.repeat(np.arange(g.n), np.diff(off)), g.indices[perm])
    assert np.all(np.diff(off) == nwgr_amd.build_degrees(g))


def test_thread_count_independence():
    a = nwgr_amd.rmat(12, 8, nthreads=1)
    b = nwgr_amd.rmat(12, 8, nthreads=7)
    np.testing.assert_array_equal(a.offsets, b.offsets)
    np.testing.assert_array_equal(a.indices, b.indices)


def test_rmat_many_blocks():
    ns = np.array([1024, 1500, 3000, 2048], np.uint32)
    g, vbase = nwgr_amd.rmat_many(ns, 8, seed=99)
    assert g.n == ns.sum() and list(vbase) == [0, 1024, 2524, 5524, 7572]
    rows = np.repeat(np.arange(g.n), np.diff(g.offsets))
    blk_row = np.searchsorted(vbase, rows, side="right") - 1
    blk_src = np.searchsorted(vbase, g.indices, side="right") - 1
    np.testing.assert_array_equal(blk_row, blk_src)  # block diagonal
    assert g.m > 0.8 * 8 * ns.sum()


def _fnv1a(a):
    h = 0xCBF29CE484222325
    for byte in np.ascontiguousarray(a).view(np.uint8).tobytes():
        h = ((h ^ byte) * 0x100000001B3) & MASK
    return h


def test_checksum_fixture_s10():
    """Regression pin of the generated scale-10 graph (fixture written by
    tests/golden/make_graphgen_checksums.py)."""
    with open(os.path.join(GOLDEN, "graphgen_checksums.json")) as f:
        fx = json.load(f)
    g = nwgr_amd.rmat(10, 16)
    e = fx["rmat_s10_ef16_seed27491095"]
    assert g.m == e["m"]
    assert _fnv1a(g.offsets) == int(e["fnv1a_offsets"], 16)
    assert _fnv1a(g.indices) == int(e["fnv1a_indices"], 16)


@pytest.mark.parametrize("scale", [16, 22])
def test_checksum_fixture_configuration_sizes(scale):
    """The R-MAT graphs of the benchmark configurations (scale 16 for the CPU
    path, scale 22 for the headline) are pinned by XXH64 of their arrays."""
    xxhash = pytest.importorskip("xxhash")
    with open(os.path.join(GOLDEN, "graphgen_checksums.json")) as f:
        e = json.load(f)[f"rmat_s{scale}_ef16_seed27491095"]
    g = nwgr_amd.rmat(scale, 16)
    assert (g.n, g.m) == (e["n"], e["m"])
    assert xxhash.xxh64(g.offsets.tobytes()).hexdigest() == e["xxh64_offsets"]
    assert xxhash.xxh64(g.indices.tobytes()).hexdigest() == e["xxh64_indices"]
