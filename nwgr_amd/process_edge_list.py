"""process_edge_list: MatrixMarket -> "NW" binary edge list, as the
reference's converter (bench/ProcessEdgelist.cpp:17-68).

    python -m nwgr_amd.process_edge_list [-d FILE] [-u FILE] INPUT

-d writes the directed edge list (a symmetric file's off-diagonal entries in
both directions), -u the undirected one (every entry once); both carry the
int attribute read_mm<..., int> reads (the value's leading integer, 1 in a
pattern file), as edge_list<..., int>::serialize lays them out.  The
MatrixMarket file is parsed in parallel (the reference reads it line by
line).  Prints the read and write times as the reference does.
"""
from __future__ import annotations

import argparse
import sys
import time


def main(argv=None) -> int:
    import nwgr_amd
    ap = argparse.ArgumentParser(prog="process_edge_list", description="BGL17 graph input handler")
    ap.add_argument("-d", dest="directed", metavar="FILE", help="generate directed graph")
    ap.add_argument("-u", dest="undirected", metavar="FILE", help="generate undirected graph")
    ap.add_argument("input")
    a = ap.parse_args(sys.argv[1:] if argv is None else argv)
    for out, directed in ((a.directed, True), (a.undirected, False)):
        if not out:
            continue
        print(f"Processing {'directed' if directed else 'undirected'} graph")
        print(f"Reading {a.input}")
        t0 = time.perf_counter()
        el = nwgr_amd.read_mm(a.input, directed=directed, int_attribute=True)
        print(f"{time.perf_counter() - t0} seconds")
        print(f"Writing {out}")
        t0 = time.perf_counter()
        nwgr_amd.write_nw_edge_list(out, el, with_ints=True)
        print(f"{time.perf_counter() - t0} seconds")
    return 0


if __name__ == "__main__":
    sys.exit(main())
