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
