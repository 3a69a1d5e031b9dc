"""Writes wide_rmat19.json: the oracle's colored multi-phase Louvain
(runMultiPhaseColoring with the engine's coloring, the restatement in
oracle/grappolo_oracle.c of Coloring/runMultiPhaseColoring.cpp:51-238) on the
RMAT graph of scale 19 (524,288 vertices, edge factor 16, a/b/c/d
0.57/0.19/0.19/0.05, unit weights, seed 3), whose first phase is swept by the
whole-GPU persistent kernel (above 262,144 vertices).  The oracle needs about
a minute there, so the SHA-256 of its labels, its final modularity and its
iterations per phase travel instead of the labels.

    python tests/golden/make_wide_fixture.py
"""
import hashlib
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402
from grappolo_amd.graph import gen_rmat  # noqa: E402


def main():
    g = gen_rmat(19, seed=3)
    C, st = oracle.run_colored(g, coloring="jp", frozen=True, minGraphSize=1000, threshold=1e-9,
                               C_threshold=1e-6)
    h = hashlib.sha256()
    for a in (g.edgeListPtrs, g.tails, g.weights):
        h.update(a.tobytes())
    out = {"graph": "gen_rmat(19, seed=3)", "numVertices": int(g.numVertices), "nnz": int(g.nnz),
           "graph_sha256": h.hexdigest(),
           "labels_sha256": hashlib.sha256(C.astype("int64").tobytes()).hexdigest(),
           "final_modularity": st["final_modularity"], "phases": st["phases"],
           "total_iterations": st["total_iterations"],
           "phase_iters": [int(x) for x in st["per_phase"]["phase_iters"]],
           "phase_nv": [int(x) for x in st["per_phase"]["phase_nv"]]}
    with open(os.path.join(HERE, "wide_rmat19.json"), "w") as f:
        json.dump(out, f, indent=1)
        f.write("\n")
    print(json.dumps(out))


if __name__ == "__main__":
    main()
