"""Extract the karate PageRank known-answer test of NWGraph into a fixture.

Reads (as text) test/page_rank_test.cpp of the reference tree: the 156-entry
directed edge list i -> j (lines 42-54) and the expected ranks for
page_rank_v1 with damping 0.85, threshold 1e-7, max 100 iterations (lines
72-75), checked there to 0.5 % relative tolerance (lines 98-101).  Writes
karate_pagerank.json next to this script.

Usage: python tests/golden/make_karate.py [/root/reference]
"""
import json
import os
import re
import sys


def _vec(src: str, name: str):
    m = re.search(r"std::vector<\w+>\s+(?:const\s+)?" + name + r"\s*=\s*\{([^}]*)\}", src)
    if not m:
        raise SystemExit(f"vector {name} not found")
    return [float(x) if "." in x or "e" in x else int(x) for x in re.findall(r"[-+0-9.eE]+", m.group(1))]


def main():
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    path = os.path.join(ref, "test", "page_rank_test.cpp")
    with open(path) as f:
        src = f.read()
    i, j, answer = _vec(src, "i"), _vec(src, "j"), _vec(src, "answer")
    assert len(i) == len(j) == 156 and len(answer) == 34
    out = {
        "source": "test/page_rank_test.cpp:42-54 (edges), :72-75 (answer), :93 (call), :98-101 (tolerance)",
        "algorithm": "page_rank_v1 (serial Jacobi), adjacency<1> of the directed edge list, degrees = out-degree",
        "n": 34, "damping": 0.85, "threshold": 1e-7, "max_iters": 100, "rel_tolerance": 0.005,
        "src": i, "dst": j, "answer": answer,
    }
    dst = os.path.join(os.path.dirname(os.path.abspath(__file__)), "karate_pagerank.json")
    with open(dst, "w") as f:
        json.dump(out, f, indent=1)
    print("wrote", dst)


if __name__ == "__main__":
    main()
