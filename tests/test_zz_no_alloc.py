"""The allocation-free top-k (nwgr_top_k_device_f32_noalloc) against the
allocating entry point nwgr_top_k_device_f32, through the C ABI: the same
ids and ranks, bit for bit, on two calls in a row (the second smaller, so
stale scratch contents from the first must not leak into it), plus ties and
the size limit."""
import ctypes
import sys

import numpy as np
import pytest

from nwgr_amd._lib import NWGR_EINVAL, NWGR_OK, pagerank_lib

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
MAX_N = 1 << 22


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def both(ranks_np, k):
    """(ids, vals) of the allocating and of the allocation-free entry point."""
    L = pagerank_lib()
    dev = torch.device("cuda:0")
    r = torch.from_numpy(ranks_np).to(dev)
    n = r.numel()
    out = []
    for fn in (L.nwgr_top_k_device_f32, L.nwgr_top_k_device_f32_noalloc):
        ids = torch.full((max(k, 1),), 0xFFFF, dtype=torch.int32, device=dev)
        vals = torch.full((max(k, 1),), -1.0, dtype=torch.float32, device=dev)
        rc = fn(r.data_ptr(), n, k, ids.data_ptr(), vals.data_ptr(), None)
        assert rc == NWGR_OK, L.nwgr_last_error()
        torch.cuda.synchronize()
        out.append((ids[:k].cpu().numpy().view(np.uint32), vals[:k].cpu().numpy()))
    return out


def test_noalloc_matches_allocating_top_k_twice():
    rng = np.random.default_rng(3)
    # the largest size the scratch holds, then a smaller one
    for n, k in ((MAX_N, 100), (100_003, 1000)):
        say(f"noalloc top-k n={n} k={k}")
        r = rng.random(n).astype(np.float32)
        r[rng.integers(0, n, n // 8)] = 0.5  # many equal ranks: ties by ascending id
        (ia, va), (ib, vb) = both(r, k)
        np.testing.assert_array_equal(ia, ib)
        np.testing.assert_array_equal(va, vb)
        want = np.lexsort((np.arange(n), -r.astype(np.float64)))[:k]
        np.testing.assert_array_equal(ia, want)


def test_noalloc_all_ranks_and_small_inputs():
    for n, k in ((1, 1), (7, 7), (4096, 4096), (5, 0)):
        r = np.random.default_rng(n).random(n).astype(np.float32)
        (ia, va), (ib, vb) = both(r, k)
        np.testing.assert_array_equal(ia, ib)
        np.testing.assert_array_equal(va, vb)


def test_noalloc_rejects_what_it_cannot_hold():
    L = pagerank_lib()
    dev = torch.device("cuda:0")
    r = torch.zeros(8, dtype=torch.float32, device=dev)
    ids = torch.zeros(8, dtype=torch.int32, device=dev)
    assert L.nwgr_top_k_device_f32_noalloc(r.data_ptr(), MAX_N + 1, 1, ids.data_ptr(), None, None) == NWGR_EINVAL
    assert b"4194304" in L.nwgr_last_error()
    assert L.nwgr_top_k_device_f32_noalloc(r.data_ptr(), 8, 9, ids.data_ptr(), None, None) == NWGR_EINVAL
    assert L.nwgr_top_k_device_f32_noalloc(None, 8, 1, ids.data_ptr(), None, None) == NWGR_EINVAL
