"""pr.exe on the MI355X: the reference's PageRank benchmark driver
(bench/pr.cpp:17-177) over this library.

    python -m nwgr_amd.pr [--version ID ...] -f FILE ... [-i NUM] [-t NUM] [-n NUM]
                          [-d] [-v] [-V] [--log FILE] [--log-header] [THREADS ...]

Same options and defaults: version 11, 20 iterations (0 means 1, :63),
tolerance 1e-4, one trial; files are MatrixMarket or the "NW" binary edge
list (load_graph, bench/common.hpp:82-101); the graph is the in-CSR of the
edge list (build_adjacency<1>) and the degrees its out-degrees
(build_degrees).  Versions 1, 2, 10, 11, 12 and 14 (the sweep forms) run the
device Jacobi solve, 3 and 13 the device residual solve; the float ranks
are what the reference's Real = float writes.  The timed region is the
whole drop-in call, as pr.cpp:94-95 times the whole page_rank call: the
first call on a graph also uploads it and builds its device layout (the
library keeps that resident graph for the later trials, versions and thread
rows on the same arrays), every call copies the ranks back to the host.
THREADS only label the rows: the device does not use host threads.  Prints
the Times table (bench/common.hpp:205-231); --verify prints the ten
best-ranked vertices as pr.cpp:47-52; --log appends Log.hpp rows ("Time(s)",
"Tolerance").  The loading phases print life_timer lines
("# [ read mm ]: 12 ms") under the names bench/common.hpp gives them.
"""
from __future__ import annotations

import argparse
import os
import sys
import time

import numpy as np

SWEEP_VERSIONS = {1, 2, 10, 11, 12, 14}
RESIDUAL_VERSIONS = {3, 13}


def parse(argv):
    ap = argparse.ArgumentParser(prog="pr.exe", description="PageRank benchmark driver (MI355X)")
    ap.add_argument("--version", dest="ids", action="append", type=int, metavar="ID",
                    help="algorithm version to run [default: 11]")
    ap.add_argument("-f", dest="files", action="append", required=True, metavar="FILE", help="input file path")
    ap.add_argument("-i", dest="max_iters", type=int, default=20, help="maximum iteration [default: 20]")
    ap.add_argument("-t", dest="tol", type=str, default="1e-4", help="tolerance [default: 1e-4]")
    ap.add_argument("-n", dest="trials", type=int, default=1, help="number of trials [default: 1]")
    ap.add_argument("--log", dest="log", metavar="FILE", help="log times to a file")
    ap.add_argument("--log-header", action="store_true", help="add a header to the log file")
    ap.add_argument("-d", "--debug", action="store_true", help="run in debug mode")
    ap.add_argument("-v", "--verify", action="store_true", help="verify results")
    ap.add_argument("-V", "--verbose", action="store_true", help="run in verbose mode")
    ap.add_argument("threads", nargs="*", type=int, metavar="THREADS")
    a = ap.parse_args(argv)
    a.ids = a.ids or [11]
    a.max_iters = a.max_iters or 1
    a.trials = a.trials or 1
    a.tol = float(np.float32(float(a.tol)))  # std::stof
    a.threads = a.threads or [os.cpu_count() or 1]
    for i in a.ids:
        if i not in SWEEP_VERSIONS | RESIDUAL_VERSIONS:
            ap.error(f"Unknown version id {i}")
    return a


def main(argv=None) -> int:
    import nwgr_amd
    from nwgr_amd.report import Times, life_timer, log_rows
    a = parse(sys.argv[1:] if argv is None else argv)
    times = Times()
    for file in a.files:
        # the phases bench/common.hpp:83-114 times with life_timer
        with open(file, "rb") as f:
            binary = f.read(2) == b"NW"
        if not binary:
            print(f"Reading matrix market input {file} (slow)")
        with life_timer("deserialize" if binary else "read mm"):
            el = nwgr_amd.load_graph(file)
        if a.verbose:
            print(f"edge list: n={el.n} m={el.m} directed={el.directed}")
        with life_timer("build adjacency"):
            graph = nwgr_amd.adjacency_from_edge_list(el)
        if a.verbose:
            print(f"adjacency: n={graph.n} m={graph.m}")
        if a.debug:
            print("indices:", " ".join(map(str, graph.offsets[:min(len(graph.offsets), 64)])))
        with life_timer("degrees"):
            degrees = nwgr_amd.build_degrees(graph)
        rankings = np.zeros(graph.n, np.float32)
        for thread in a.threads:
            for vid in a.ids:
                solve = nwgr_amd.page_rank_delta if vid in RESIDUAL_VERSIONS else nwgr_amd.page_rank
                for _ in range(a.trials):
                    t0 = time.perf_counter()
                    solve(graph, degrees, rankings, 0.85, a.tol, a.max_iters, thread)
                    times.append(file, vid, thread, time.perf_counter() - t0, a.tol)
                if a.verify:
                    print("Verifying")
                    ids, vals = nwgr_amd.top_k(rankings, min(10, graph.n))
                    for v, r in zip(ids, vals):
                        print(f"{int(v)}: {float(r):.6f}")
    times.print(sys.stdout)
    if a.log:
        log_rows(a.log, "pr", times, a.log_header, "Time(s)", "Tolerance")
    return 0


if __name__ == "__main__":
    sys.exit(main())
