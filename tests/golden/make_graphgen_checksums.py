"""Write graphgen_checksums.json: edge counts and FNV-1a (scale 10) or XXH64
(scales 16 and 22, the configuration sizes) checksums of the generated graphs, so later changes to the generator are caught.  The generator
itself is pinned against its specification by tests/test_graphgen.py."""
import json
import os
import sys

import numpy as np

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__)))))
import nwgr_amd  # noqa: E402

MASK = (1 << 64) - 1


def fnv1a(a):
    h = 0xCBF29CE484222325
    for byte in np.ascontiguousarray(a).view(np.uint8).tobytes():
        h = ((h ^ byte) * 0x100000001B3) & MASK
    return h


def main():
    out = {}
    for scale in (10,):
        g = nwgr_amd.rmat(scale, 16)
        out[f"rmat_s{scale}_ef16_seed27491095"] = {
            "n": g.n, "m": g.m, "fnv1a_offsets": hex(fnv1a(g.offsets)), "fnv1a_indices": hex(fnv1a(g.indices))}
    import xxhash
    for scale in (16, 22):
        g = nwgr_amd.rmat(scale, 16)
        out[f"rmat_s{scale}_ef16_seed27491095"] = {
            "n": g.n, "m": g.m, "xxh64_offsets": xxhash.xxh64(g.offsets.tobytes()).hexdigest(),
            "xxh64_indices": xxhash.xxh64(g.indices.tobytes()).hexdigest()}
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "graphgen_checksums.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=1)
    print(json.dumps(out, indent=1))


if __name__ == "__main__":
    main()
