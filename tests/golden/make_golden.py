"""Writes the golden fixtures of tests/test_golden.py from the compiled
reference (oracle/_ref/libgrappolo_ref.so, built from the reference sources by
oracle/Makefile).  Run in a checkout that has the reference tree:

    python tests/golden/make_golden.py

Each ref_<name>.npz holds a small input graph and what the reference's own
routines return for it at one thread: colors, one colored phase, the plain
phases, renumbering, coarsening and the two multi-phase drivers.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(HERE))

import oracle  # noqa: E402
from grappolo_amd.graph import Graph, gen_planted  # noqa: E402


def graphs():
    from graphs import karate
    g_loops = gen_planted(600, 6, 10, 0.25, seed=31, wlo=0.5, whi=2.0)
    rows = np.repeat(np.arange(g_loops.numVertices), g_loops.degrees())
    keep = rows < g_loops.tails
    pairs = np.stack([rows[keep], g_loops.tails[keep]], 1)
    w = g_loops.weights[keep]
    loops = np.arange(0, 600, 9)
    pairs = np.concatenate([pairs, np.stack([loops, loops], 1)])
    w = np.concatenate([w, np.full(len(loops), 3.0)])
    return {
        "karate": karate(),
        "planted1k": gen_planted(1024, 8, 12, 0.3, seed=21),
        "planted1k_w": gen_planted(1024, 8, 12, 0.3, seed=22, wlo=0.5, whi=2.0),
        "loops650": Graph.from_pairs(650, pairs, w),
    }


def main():
    if not oracle.ref_available():
        oracle.build(ref=True)
    for name, g in graphs().items():
        out = dict(off=g.edgeListPtrs, tails=g.tails, w=g.weights, num_edges=g.numEdges)
        colors, mc = oracle.ref_color(g, 1)
        out.update(colors=colors, max_color=mc)
        for nomap in (0, 1):
            q, C, it = oracle.ref_colored_phase(g, colors, mc + 1, -1.0, 1e-6, nomap=bool(nomap))
            out.update({f"colored{nomap}_C": C, f"colored{nomap}_mod": q, f"colored{nomap}_it": it})
            q, C, it = oracle.ref_plain_phase(g, -1.0, 1e-6, nomap=bool(nomap))
            out.update({f"plain{nomap}_C": C, f"plain{nomap}_mod": q, f"plain{nomap}_it": it})
        Cr, k = oracle.ref_renumber(out["colored0_C"])
        out.update(renum_C=Cr, renum_k=k)
        cg = oracle.canonical(oracle.ref_coarsen(g, Cr, k))
        out.update(coarse_off=cg.edgeListPtrs, coarse_tails=cg.tails, coarse_w=cg.weights,
                   coarse_num_edges=cg.numEdges)
        mgs = 100 if g.numVertices > 100 else 10
        C_orig, final = oracle.ref_run_colored(g, minGraphSize=mgs)
        out.update(run_colored_C=C_orig, run_colored_final=final, run_colored_mgs=mgs)
        for b in (0, 1):
            C_orig, final = oracle.ref_run_basic(g, basicOpt=b)
            out.update({f"run_basic{b}_C": C_orig, f"run_basic{b}_final": final})
        path = os.path.join(HERE, f"ref_{name}.npz")
        np.savez_compressed(path, **out)
        print(path, os.path.getsize(path))


if __name__ == "__main__":
    main()
