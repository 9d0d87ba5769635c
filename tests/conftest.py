import json
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X) and the built HIP library")
    config.addinivalue_line("markers", "slow: full-size workloads")


@pytest.fixture(scope="session")
def karate():
    """Karate known-answer test of test/page_rank_test.cpp:42-54, :72-75."""
    with open(os.path.join(GOLDEN, "karate_pagerank.json")) as f:
        d = json.load(f)
    import nwgr_amd
    g = nwgr_amd.build_adjacency(np.array(d["src"], np.uint32), np.array(d["dst"], np.uint32), n=d["n"])
    deg = nwgr_amd.build_degrees(g)
    return d, g, deg


def l1(a, b):
    return float(np.abs(np.asarray(a, np.float64) - np.asarray(b, np.float64)).sum())


def assert_same_top_k(r, r64, k, tie_rel=1e-12):
    """Top-k order of r equals the FP64 order (vertex-id tie-break), except
    that vertices whose FP64 ranks are equal to `tie_rel` (ties in exact
    arithmetic, e.g. symmetric vertices) may appear in either order.
    Returns the smallest relative gap between distinct consecutive FP64
    ranks within the top k."""
    import oracle
    r64 = np.asarray(r64, np.float64)
    k = min(k, len(r64))
    want = oracle.top_k(r64, k)
    got = oracle.top_k(r, k)
    vals = r64[want]
    # tie groups along the FP64 order
    group = np.zeros(k, np.int64)
    for i in range(1, k):
        same = abs(vals[i] - vals[i - 1]) <= tie_rel * max(abs(vals[i]), 1e-300)
        group[i] = group[i - 1] if same else group[i - 1] + 1
    # the boundary group may extend past position k: take all its members
    last = vals[-1]
    members = set(np.nonzero(np.abs(r64 - last) <= tie_rel * abs(last))[0].tolist())
    for i in range(k):
        if group[i] == group[-1]:
            assert got[i] in members or got[i] in set(want[group == group[-1]].tolist()), (i, got[i], want[i])
        else:
            allowed = set(want[group == group[i]].tolist())
            assert got[i] in allowed, f"position {i}: got vertex {got[i]}, FP64 order has {want[i]}"
    gaps = [abs(vals[i] - vals[i - 1]) / abs(vals[i]) for i in range(1, k) if group[i] != group[i - 1]]
    return min(gaps) if gaps else float("inf")
