"""Writes full_size_coloring.json: the oracle's restated Jones-Plassmann
coloring of the LFR-style 4,194,304-vertex graph (seed 2), as checksums, so
that a GPU test can compare the device's coloring of that graph with it bit
for bit without running the oracle (minutes on one core) on the GPU host.

    python tests/golden/make_full_size_coloring.py
"""
import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402
from grappolo_amd.graph import gen_lfr  # noqa: E402


def sha(*arrays):
    h = hashlib.sha256()
    for a in arrays:
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def main():
    g = gen_lfr(4194304, seed=2)
    colors, mc, rounds = oracle.color_jp(g)
    out = {
        "graph": "gen_lfr(4194304, seed=2)",
        "numVertices": int(g.numVertices),
        "nnz": int(g.nnz),
        "graph_sha256": sha(g.edgeListPtrs, g.tails, g.weights),
        "max_color": int(mc),
        "colors_int32_sha256": sha(colors.astype(np.int32)),
        "class_sizes": np.bincount(colors, minlength=mc + 1).tolist(),
    }
    with open(os.path.join(HERE, "full_size_coloring.json"), "w") as f:
        json.dump(out, f, indent=1)
        f.write("\n")
    print(json.dumps({k: v for k, v in out.items() if k != "class_sizes"}))


if __name__ == "__main__":
    main()
