"""Pins the oracle (oracle/grappolo_oracle.c) before it is trusted as the
checker of the HIP engine:
  * against the golden fixtures written by the compiled reference
    (tests/golden/make_golden.py) -- runs everywhere;
  * against the compiled reference itself on freshly generated graphs
    (skipped where oracle/_ref was not built);
  * against exact rational brute force on tiny graphs (best move, tie and
    swap rules, modularity closed forms) and the invariants of coarsening.
"""
import glob
import itertools
import os
from fractions import Fraction

import numpy as np
import pytest

import oracle
from grappolo_amd.graph import Graph, gen_lfr, gen_planted

from graphs import GOLDEN, catalog, karate

FIXTURES = sorted(glob.glob(os.path.join(GOLDEN, "ref_*.npz")))
need_ref = pytest.mark.skipif(not oracle.ref_available(), reason="reference checker not built")


def _fixture(path):
    z = np.load(path, allow_pickle=False)
    g = Graph(z["off"], z["tails"], z["w"], numEdges=int(z["num_edges"]))
    return g, z


# ------------------------------------------------------------ golden vectors
@pytest.mark.parametrize("path", FIXTURES, ids=lambda p: os.path.basename(p))
def test_oracle_matches_golden(path):
    g, z = _fixture(path)
    colors, mc = oracle.color_greedy(g)
    np.testing.assert_array_equal(colors, z["colors"])
    assert mc == int(z["max_color"])
    for nomap in (0, 1):
        # both colored variants compute the same targets
        q, C, it, _ = oracle.louvain_colored(g, colors, mc + 1, -1.0, 1e-6, frozen=False)
        assert q == float(z[f"colored{nomap}_mod"]) and it == int(z[f"colored{nomap}_it"])
        np.testing.assert_array_equal(C, z[f"colored{nomap}_C"])
        q, C, it, _ = oracle.louvain_plain(g, -1.0, 1e-6, nomap=bool(nomap))
        assert q == float(z[f"plain{nomap}_mod"]) and it == int(z[f"plain{nomap}_it"])
        np.testing.assert_array_equal(C, z[f"plain{nomap}_C"])
    Cr, k = oracle.renumber(z["colored0_C"])
    np.testing.assert_array_equal(Cr, z["renum_C"])
    assert k == int(z["renum_k"])
    cg = oracle.coarsen(g, Cr, k)
    np.testing.assert_array_equal(cg.edgeListPtrs, z["coarse_off"])
    np.testing.assert_array_equal(cg.tails, z["coarse_tails"])
    np.testing.assert_allclose(cg.weights, z["coarse_w"], rtol=1e-15, atol=0)
    assert cg.numEdges == int(z["coarse_num_edges"])
    C_orig, st = oracle.run_colored(g, coloring="greedy", frozen=False,
                                    minGraphSize=int(z["run_colored_mgs"]))
    np.testing.assert_array_equal(C_orig, z["run_colored_C"])
    assert round(st["final_modularity"], 6) == pytest.approx(float(z["run_colored_final"]),
                                                             abs=1e-6)
    for b in (0, 1):
        C_orig, st = oracle.run_basic(g, basicOpt=b)
        np.testing.assert_array_equal(C_orig, z[f"run_basic{b}_C"])
        assert st["final_modularity"] == pytest.approx(float(z[f"run_basic{b}_final"]), abs=1e-6)


# --------------------------------------------------- the compiled reference
REF_GRAPHS = {
    "planted4k": lambda: gen_planted(4096, 16, 16, 0.3, seed=5),
    "planted4k_w": lambda: gen_planted(4096, 16, 16, 0.3, seed=6, wlo=0.5, whi=2.0),
    "lfr10k_w": lambda: gen_lfr(10000, seed=3, dmax=300, cmax=800),
    "hubs": lambda: catalog()["hubs"],
    "loops": lambda: catalog()["loops_isolated"],
}


@need_ref
@pytest.mark.parametrize("name", list(REF_GRAPHS))
def test_oracle_matches_compiled_reference(name):
    g = REF_GRAPHS[name]()
    colors, mc = oracle.ref_color(g, 1)
    oc, omc = oracle.color_greedy(g)
    np.testing.assert_array_equal(colors, oc)
    for nomap in (False, True):
        q, C, it = oracle.ref_colored_phase(g, colors, mc + 1, -1.0, 1e-6, nomap=nomap)
        qo, Co, ito, _ = oracle.louvain_colored(g, colors, mc + 1, -1.0, 1e-6, frozen=False)
        assert (q, it) == (qo, ito)
        np.testing.assert_array_equal(C, Co)
        q, C, it = oracle.ref_plain_phase(g, -1.0, 1e-6, nomap=nomap)
        qo, Co, ito, _ = oracle.louvain_plain(g, -1.0, 1e-6, nomap=nomap)
        assert (q, it) == (qo, ito)
        np.testing.assert_array_equal(C, Co)
    Cr, k = oracle.ref_renumber(C)
    Co, ko = oracle.renumber(C)
    assert k == ko
    np.testing.assert_array_equal(Cr, Co)
    rc = oracle.canonical(oracle.ref_coarsen(g, Cr, k, nthreads=1))
    oc = oracle.coarsen(g, Co, ko)
    np.testing.assert_array_equal(rc.edgeListPtrs, oc.edgeListPtrs)
    np.testing.assert_array_equal(rc.tails, oc.tails)
    np.testing.assert_array_equal(rc.weights, oc.weights)  # same one-thread sum order
    assert rc.numEdges == oc.numEdges
    C_orig, final = oracle.ref_run_colored(g, minGraphSize=100)
    Co, st = oracle.run_colored(g, coloring="greedy", frozen=False, minGraphSize=100)
    np.testing.assert_array_equal(C_orig, Co)
    assert float(f"{st['final_modularity']:.6f}") == final


@need_ref
def test_coarsen_matches_reference_with_threads():
    """With 8 threads the reference's rows come out in thread order and its
    sums in lock order; the canonical form is the same for integer weights."""
    g = gen_planted(4096, 16, 16, 0.3, seed=9)
    C = np.random.default_rng(0).integers(0, 300, g.numVertices)
    Cr, k = oracle.renumber(C)
    rc = oracle.canonical(oracle.ref_coarsen(g, Cr, k, nthreads=8))
    oc = oracle.coarsen(g, Cr, k)
    for a, b in ((rc.edgeListPtrs, oc.edgeListPtrs), (rc.tails, oc.tails),
                 (rc.weights, oc.weights)):
        np.testing.assert_array_equal(a, b)


# ------------------------------------------------------------- brute force
def _exact_q(g, C):
    """Modularity with rational arithmetic."""
    two_m = Fraction(0)
    a = {}
    e = Fraction(0)
    for i in range(g.numVertices):
        for j in range(g.edgeListPtrs[i], g.edgeListPtrs[i + 1]):
            w = Fraction(float(g.weights[j]))
            two_m += w
            if C[i] >= 0:
                a[C[i]] = a.get(C[i], Fraction(0)) + w
                if C[g.tails[j]] == C[i]:
                    e += w
    return e / two_m - sum(x * x for x in a.values()) / (two_m * two_m)


def _tiny_graphs():
    rng = np.random.default_rng(42)
    out = []
    for t in range(25):
        n = int(rng.integers(4, 11))
        pairs = [(u, v) for u, v in itertools.combinations(range(n), 2) if rng.random() < 0.45]
        if t % 3 == 0:
            pairs += [(u, u) for u in range(n) if rng.random() < 0.3]
        if not pairs:
            pairs = [(0, 1)]
        w = rng.integers(1, 4, len(pairs)).astype(float) if t % 2 else None
        out.append(Graph.from_pairs(n, pairs, w))
    return out


@pytest.mark.parametrize("gi", range(25))
def test_best_move_is_exact_argmax_of_modularity_gain(gi):
    """max() of utilityClusteringFunctions.cpp:226-259: the target maximises the
    exact modularity gain among the neighbours' communities, needs a strictly
    positive gain, prefers the smaller id on exact ties, and a singleton never
    moves to a larger singleton."""
    g = _tiny_graphs()[gi]
    rng = np.random.default_rng(gi)
    n = g.numVertices
    for trial in range(4):
        C = rng.integers(0, max(2, n // 2), n).astype(np.int64)
        deg = g.degrees()
        C[deg == 0] = -1
        q0 = _exact_q(g, C)
        for v in range(n):
            if deg[v] == 0:
                continue
            colors = np.ones(n, dtype=np.int32)
            colors[v] = 0
            got, _ = oracle.color_step(g, C, colors, 0)
            s = C[v]
            cands = {int(C[t]) for t in g.tails[g.edgeListPtrs[v]:g.edgeListPtrs[v + 1]]} - {s}
            gains = {}
            for y in cands:
                C2 = C.copy()
                C2[v] = y
                gains[y] = _exact_q(g, C2) - q0
            best = s
            pos = {y: d for y, d in gains.items() if d > 0}
            if pos:
                top = max(pos.values())
                best = min(y for y, d in pos.items() if d == top)
                size = np.bincount(C[C >= 0], minlength=n)
                if size[best] == 1 and size[s] == 1 and best > s:
                    best = s
            assert got[v] == best, (gi, trial, v, gains)


def test_modularity_closed_forms():
    # two disjoint triangles in two clusters: 1 - 2 * (1/2)^2
    t = Graph.from_pairs(6, [(0, 1), (1, 2), (0, 2), (3, 4), (4, 5), (3, 5)])
    assert oracle.modularity(t, np.array([0, 0, 0, 1, 1, 1])) == pytest.approx(0.5, abs=1e-15)
    # everything in one cluster: 0
    assert oracle.modularity(t, np.zeros(6, np.int64)) == pytest.approx(0.0, abs=1e-15)
    # ring of 4 cliques K5 joined by single edges, one cluster per clique
    pairs = []
    for c in range(4):
        pairs += [(5 * c + i, 5 * c + j) for i, j in itertools.combinations(range(5), 2)]
        pairs.append((5 * c + 4, (5 * c + 5) % 20))
    g = Graph.from_pairs(20, pairs)
    m = 4 * 10 + 4
    exp = 4 * (10 / m - (22 / (2 * m)) ** 2)  # a_c = 2*10 internal + 2 ring endpoints
    assert oracle.modularity(g, np.repeat(np.arange(4), 5)) == pytest.approx(exp, abs=1e-15)
    # against exact arithmetic on random labellings
    for gi, g in enumerate(_tiny_graphs()[:10]):
        C = np.random.default_rng(gi).integers(0, 3, g.numVertices).astype(np.int64)
        C[g.degrees() == 0] = -1
        assert oracle.modularity(g, C) == pytest.approx(float(_exact_q(g, C)), abs=1e-15)


def test_renumber_is_first_appearance_order():
    C = np.array([5, 5, -1, 2, 9, 2, 5, -1, 0, 9], np.int64)
    got, k = oracle.renumber(C)
    np.testing.assert_array_equal(got, [0, 0, -1, 1, 2, 1, 0, -1, 3, 2])
    assert k == 4


@pytest.mark.parametrize("name", ["karate", "planted_w", "loops_isolated", "hubs"])
def test_coarsen_invariants(name):
    g = catalog()[name]
    C = np.random.default_rng(3).integers(0, max(2, g.numVertices // 10), g.numVertices)
    C[g.degrees() == 0] = -1
    Cr, k = oracle.renumber(C)
    cg = oracle.coarsen(g, Cr, k)
    assert cg.numVertices == k
    rows = np.repeat(np.arange(k), cg.degrees())
    # every cluster has exactly one self-loop entry; rows sorted, no repeats
    loops = rows == cg.tails
    assert np.bincount(rows[loops], minlength=k).tolist() == [1] * k
    key = rows * k + cg.tails
    assert np.all(np.diff(key) > 0)
    # symmetric with equal weights
    order = np.lexsort((rows, cg.tails))
    np.testing.assert_array_equal(cg.tails[order], rows)
    np.testing.assert_array_equal(cg.weights[order], cg.weights)
    # total weight kept; numEdges = NE_out + k (buildNextPhase.cpp:237)
    assert cg.weights.sum() == pytest.approx(g.weights.sum(), rel=1e-12)
    assert cg.numEdges == (cg.nnz - k) // 2 + k
    # contraction keeps modularity: Q(G, C) == Q(G', singletons)
    assert oracle.modularity(cg, np.arange(k, dtype=np.int64)) == pytest.approx(
        oracle.modularity(g, Cr), abs=1e-12)


@pytest.mark.parametrize("name", ["karate", "planted_unit", "lfr_w", "rmat12", "hubs"])
def test_colorings_valid(name):
    g = catalog()[name]
    for colors, mc in (oracle.color_greedy(g), oracle.color_jp(g)[:2]):
        assert colors.min() >= 0 and colors.max() == mc
        assert oracle.color_conflicts(g, colors) == 0


@pytest.mark.parametrize("name", ["planted_unit", "lfr_w", "rmat12", "hubs"])
def test_colorings_are_first_fit(name):
    """Any first-fit coloring (the reference's greedy, and the engine's
    Jones-Plassmann rounds plus one-at-a-time tail) gives every vertex the
    smallest color absent from the neighbours colored before it, so every
    smaller color is present among its neighbours."""
    g = catalog()[name]
    for colors in (oracle.color_greedy(g)[0], oracle.color_jp(g)[0]):
        for v in range(g.numVertices):
            nb = g.tails[g.edgeListPtrs[v]:g.edgeListPtrs[v + 1]]
            have = set(colors[nb[nb != v]].tolist())
            assert all(c in have for c in range(colors[v])), v


def _jp_python(g, tail=1024):
    """Jones-Plassmann written out in Python, independently of the C
    restatement: in each round every uncoloured vertex whose key (degree
    capped at 2^23-1, top 40 bits of the id hash, id) beats all uncoloured
    neighbours takes the smallest colour its coloured neighbours leave free;
    once at most `tail` vertices remain, they are coloured first-fit in
    increasing id."""
    nv = g.numVertices
    deg = np.diff(g.edgeListPtrs)
    key = [(min(int(deg[v]), (1 << 23) - 1), int(oracle.jp_priority(v)) >> 24, v) for v in range(nv)]
    nbr = [[int(u) for u in g.tails[g.edgeListPtrs[v]:g.edgeListPtrs[v + 1]] if u != v] for v in range(nv)]
    col = [-1] * nv

    def first_fit(v):
        used = {col[u] for u in nbr[v] if col[u] >= 0}
        c = 0
        while c in used:
            c += 1
        return c

    work = list(range(nv))
    rounds = 0
    while work:
        win = [v for v in work if all(col[u] >= 0 or key[u] < key[v] for u in nbr[v])]
        for v in win:
            col[v] = first_fit(v)
        work = [v for v in work if col[v] < 0]
        rounds += 1
        if 0 < len(work) <= tail:
            break
    for v in range(nv):
        if col[v] < 0:
            col[v] = first_fit(v)
    return np.array(col, np.int32), rounds


@pytest.mark.parametrize("name", ["karate", "planted_unit", "hubs", "rmat12"])
def test_jp_matches_python_rounds(name):
    """The C restatement of the engine's coloring equals the rounds written
    out in Python, and a vertex whose key beats all its neighbours gets 0."""
    g = catalog()[name]
    colors, mc, rounds = oracle.color_jp(g)
    pc, prounds = _jp_python(g)
    np.testing.assert_array_equal(colors, pc)
    assert rounds == prounds and mc == pc.max()
    deg = np.diff(g.edgeListPtrs)
    key = [(min(int(deg[v]), (1 << 23) - 1), int(oracle.jp_priority(v)) >> 24, v) for v in range(g.numVertices)]
    for v in range(g.numVertices):
        nb = [u for u in g.tails[g.edgeListPtrs[v]:g.edgeListPtrs[v + 1]].tolist() if u != v]
        if nb and all(key[u] < key[v] for u in nb):
            assert colors[v] == 0


def test_frozen_equals_sequential_with_valid_coloring():
    g = catalog()["lfr_w"]
    colors, mc = oracle.color_greedy(g)
    a = oracle.louvain_colored(g, colors, mc + 1, -1.0, 1e-6, frozen=True)
    b = oracle.louvain_colored(g, colors, mc + 1, -1.0, 1e-6, frozen=False)
    assert a[0] == b[0] and a[2] == b[2]
    np.testing.assert_array_equal(a[1], b[1])


def test_karate_modularity_near_known_optimum():
    """Zachary's karate club: the best partition has Q ~ 0.4198; Louvain runs
    land just below it."""
    C, st = oracle.run_colored(karate(), coloring="greedy", minGraphSize=10)
    assert 0.40 <= st["final_modularity"] <= 0.4198 + 1e-4


# ------------------------------------------------------- vertex following
def _pairs_graph(nv, pairs, w=None):
    return Graph.from_pairs(nv, np.array(pairs, np.int64).reshape(-1, 2), w)


def test_vertex_following_closed_forms():
    """vertexFollowing.cpp:46-89 on hand-built cases: a star's leaves follow
    the centre, in a lone pair the larger id follows the smaller, a vertex
    whose only entry is its own loop and an isolated vertex (-1) do not."""
    # 0: centre of leaves 1..4; 5-6 a pair; 7 self-loop only; 8 isolated; 9-10-11 a path
    g = _pairs_graph(12, [(0, 1), (0, 2), (0, 3), (0, 4), (5, 6), (7, 7), (9, 10), (10, 11)])
    C, n = oracle.vertex_following(g)
    np.testing.assert_array_equal(C, [0, 0, 0, 0, 0, 5, 5, 7, -1, 10, 10, 10])
    assert n == 4 + 1 + 1 + 2
    Cr, k = oracle.renumber(C)
    assert k == 4
    vf = oracle.build_vf(g, Cr, k)
    # star -> one vertex with loop 2*4; pair -> loop 2; self-loop kept; path -> loop 4
    assert vf.numVertices == 4 and vf.numEdges == 4
    np.testing.assert_array_equal(vf.edgeListPtrs, [0, 1, 2, 3, 4])
    np.testing.assert_array_equal(vf.tails, [0, 1, 2, 3])
    np.testing.assert_array_equal(vf.weights, [8.0, 2.0, 1.0, 4.0])


def test_build_vf_invariants():
    """buildNewGraphVF: no zero self-loop seeds, every weight truncated to an
    integer before summing, numEdges = NE_out + NE_self."""
    from graphs import pendants
    g = pendants(seed=3, nv=3000)
    C, n = oracle.vertex_following(g)
    assert n > 0
    Cr, k = oracle.renumber(C)
    vf = oracle.build_vf(g, Cr, k)
    assert vf.numVertices == k
    rows = np.repeat(np.arange(k), vf.degrees())
    loops = rows == vf.tails
    ne_self = int(loops.sum())
    assert vf.numEdges == (vf.nnz - ne_self) // 2 + ne_self
    assert np.all(np.diff(rows * k + vf.tails) > 0)
    order = np.lexsort((rows, vf.tails))
    np.testing.assert_array_equal(vf.tails[order], rows)
    np.testing.assert_array_equal(vf.weights[order], vf.weights)
    # total truncated weight is kept: every input entry lands in exactly one
    # output entry (self pairs counted from both endpoints, as in the input)
    tw = np.trunc(g.weights)
    assert vf.weights[~loops].sum() + vf.weights[loops].sum() == tw[Cr[np.repeat(
        np.arange(g.numVertices), g.degrees())] >= 0].sum()
    assert np.all(vf.weights == np.trunc(vf.weights))
    # no seeds: a cluster without internal entries has no loop
    Cs = Cr[np.repeat(np.arange(g.numVertices), g.degrees())]
    internal = np.zeros(k, bool)
    internal[Cs[Cs == Cr[g.tails]]] = True
    assert np.array_equal(np.bincount(rows[loops], minlength=k) > 0, internal)


@need_ref
@pytest.mark.parametrize("name", ["karate", "loops_isolated", "rmat12", "hubs", "pendants_w", "pendants"])
def test_vertex_following_matches_reference(name):
    from graphs import pendants
    g = {"pendants_w": lambda: pendants(seed=5), "pendants": lambda: pendants(seed=6, weighted=False)}.get(
        name, lambda: catalog()[name])()
    C, n = oracle.vertex_following(g)
    Cr, nr = oracle.ref_vertex_following(g)
    assert n == nr
    np.testing.assert_array_equal(C, Cr)
    C2, k = oracle.renumber(C)
    a = oracle.build_vf(g, C2, k)
    b = oracle.canonical(oracle.ref_build_vf(g, C2, k))
    np.testing.assert_array_equal(a.edgeListPtrs, b.edgeListPtrs)
    np.testing.assert_array_equal(a.tails, b.tails)
    np.testing.assert_array_equal(a.weights, b.weights)  # integer sums: order-free
    assert a.numEdges == b.numEdges


# ------------------------------------------------------ cluster comparison
def _brute_compare(C1, C2):
    """Pair counts by enumerating every vertex pair after the reference's -1
    relabelling (utilityClusterComparisonMetrics.cpp:62-107)."""
    a = np.array(C1, np.int64)
    b = np.array(C2, np.int64)
    n = len(a)
    nC1 = a.max() + (1 if (a == 0).any() else 0)
    if (a == -1).any():
        a[a == -1] = nC1
    b[b == -1] = n + 1
    ss = sd = ds = 0
    for i, j in itertools.combinations(range(n), 2):
        s1, s2 = a[i] == a[j], b[i] == b[j]
        ss += s1 and s2
        sd += s1 and not s2
        ds += s2 and not s1
    return int(ss), int(sd), int(ds), a, b


def _gini_exact(sizes):
    s = sorted(int(x) for x in sizes)
    n = len(s)
    return Fraction(2 * sum((i + 1) * v for i, v in enumerate(s)), n * sum(s)) - Fraction(n + 1, n)


@pytest.mark.parametrize("seed", range(6))
def test_compare_brute_force(seed):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(2, 60))
    C1 = rng.integers(seed % 2, 6, n)  # odd seeds: one-based truth
    C2 = rng.integers(0, 9, n)
    if seed >= 2:
        C1[rng.random(n) < 0.2] = -1
        C2[rng.random(n) < 0.2] = -1
    if (C1 == -1).all():
        C1[0] = 1
    out, a, b = oracle.compare(C1, C2)
    ss, sd, ds, ea, eb = _brute_compare(C1, C2)
    assert (out["same_same"], out["same_diff"], out["diff_same"], out["agree"]) == (ss, sd, ds, ss)
    np.testing.assert_array_equal(a, ea)
    np.testing.assert_array_equal(b, eb)
    assert out["precision"] == ss / (ss + ds) if ss + ds else np.isnan(out["precision"])
    assert out["recall"] == ss / (ss + sd) if ss + sd else np.isnan(out["recall"])


def test_gini_closed_forms():
    """Equal sizes give 0; one non-empty community among n gives (n-1)/n;
    random sizes match exact rational arithmetic; the input is sorted in place."""
    g, s = oracle.gini([7] * 9)
    assert g == pytest.approx(0.0, abs=1e-15)
    g, s = oracle.gini([0, 0, 0, 12, 0])
    assert g == pytest.approx(4 / 5, abs=1e-15)
    np.testing.assert_array_equal(s, [0, 0, 0, 0, 12])
    rng = np.random.default_rng(1)
    for _ in range(5):
        sz = rng.integers(0, 10000, int(rng.integers(2, 500)))
        g, s = oracle.gini(sz)
        assert g == pytest.approx(float(_gini_exact(sz)), rel=1e-14, abs=1e-15)
        np.testing.assert_array_equal(s, np.sort(sz))


@need_ref
@pytest.mark.parametrize("case", range(8))
def test_compare_matches_reference(case):
    """Counts exactly, the floats as the reference prints them (%lf), the
    relabelled vectors, for zero- and one-based labels with and without -1."""
    rng = np.random.default_rng(100 + case)
    n = int(rng.integers(10, 4000))
    C1 = rng.integers(case % 2, int(rng.integers(2, 80)), n)
    C2 = rng.integers((case // 2) % 2, int(rng.integers(2, 120)), n)
    if case >= 4:
        C1[rng.random(n) < 0.05] = -1
    if case >= 6:
        C2[rng.random(n) < 0.05] = -1
    out, a, b = oracle.compare(C1, C2)
    ref, ra, rb = oracle.ref_compare(C1, C2, nthreads=1)
    np.testing.assert_array_equal(a, ra)
    np.testing.assert_array_equal(b, rb)
    for k, v in ref.items():
        if isinstance(v, int):
            assert out[k] == v, k
        else:
            assert f"{out[k]:f}" == f"{v:f}", k


@need_ref
def test_reference_compare_counts_scale_with_threads():
    """The reference's pair loops sit in a nested parallel region whose every
    outer thread walks all communities: with t threads it prints t times the
    pair counts, with the same ratios.  The restatement keeps the one-thread
    counts."""
    rng = np.random.default_rng(7)
    C1 = rng.integers(0, 30, 2000)
    C2 = rng.integers(0, 50, 2000)
    one, _, _ = oracle.ref_compare(C1, C2, nthreads=1)
    four, _, _ = oracle.ref_compare(C1, C2, nthreads=4)
    for k in ("same_same", "same_diff", "diff_same"):
        assert four[k] == 4 * one[k]
    for k in ("precision", "recall", "f_score", "gini1", "gini2"):
        assert four[k] == one[k]


@need_ref
def test_gini_matches_reference():
    rng = np.random.default_rng(3)
    for n in (1, 2, 17, 1000):
        sz = rng.integers(-5, 100000, n)
        g, s = oracle.gini(sz)
        rg, rs = oracle.ref_gini(sz)
        assert g == rg
        np.testing.assert_array_equal(s, rs)


# ------------------------------------------------------- coloring 2 and 3
def _graphs_for_coloring():
    from graphs import pendants
    c = catalog()
    return [karate(), c["planted_unit"], c["loops_isolated"], c["rmat12"], c["hubs"], pendants(seed=4)]


@need_ref
@pytest.mark.parametrize("nItrs", [1, 2, 4])
def test_multihash_matches_reference_with_its_random_values(nItrs):
    """coloringMultiHashMaxMin.cpp:60-226: given the rand() values the
    reference drew, the restatement colors every vertex the same and returns
    the same maxColor (including the early-exit value of :165-170)."""
    for g in _graphs_for_coloring():
        c, mc, vals = oracle.ref_multihash(g, 2, nItrs, seed=11)
        o, omc = oracle.color_multihash_vals(g, 2, nItrs, vals)
        assert omc == mc
        np.testing.assert_array_equal(o, c)


def test_multihash_classes_are_independent():
    """Every class but the uncoloured one (maxColor) is an independent set;
    a vertex stays uncoloured only if every pass saw a neighbour on both sides."""
    for g in _graphs_for_coloring():
        for nItrs in (1, 3):
            c, mc = oracle.color_multihash(g, 2, nItrs)
            assert c.min() >= 0 and c.max() <= mc
            rows = np.repeat(np.arange(g.numVertices), g.degrees())
            bad = (rows != g.tails) & (c[rows] == c[g.tails]) & (c[rows] != mc)
            assert not bad.any()


def _balance_invariants(g, base, out, ncolors, type):
    import math
    nv = g.numVertices
    avg = math.ceil(nv / ncolors)
    f0 = np.bincount(base, minlength=ncolors)
    f1 = np.bincount(out, minlength=ncolors)
    over = f0 > avg
    rows = np.repeat(np.arange(nv), g.degrees())
    nl = rows != g.tails
    assert not np.any(out[rows[nl]] == out[g.tails[nl]])          # proper
    assert np.all(f1[~over] <= avg)                                # non-oversize capped
    assert np.all(f1[over] <= f0[over]) and np.all(f1[over] >= avg)
    moved = out != base
    assert np.all(over[base[moved]]) and not np.any(over[out[moved]])
    elig = (~over) & (f1 < avg)
    for v in np.flatnonzero(over[base] & (f1[out] > avg)):        # nothing left to move
        held = set(out[g.tails[g.edgeListPtrs[v]:g.edgeListPtrs[v + 1]]].tolist())
        assert not any(elig[c] and c not in held for c in range(ncolors)), v


@pytest.mark.parametrize("type", [0, 1])
def test_balance_invariants(type):
    """vBase.cpp:6-168: what its rules fix -- a proper coloring, colors that
    were not oversize end at most avg, oversize classes neither grow nor drop
    below avg, only vertices of oversize classes move, and no vertex that
    could still move is left.  (The compiled reference's vBaseRedistribution
    cannot be run as a checker: when a vertex finds no eligible color it reads
    and writes one past its count arrays, and on karate at one thread it
    aborts with a corrupted heap.)"""
    for g in _graphs_for_coloring():
        for base, mc in (oracle.color_greedy(g), oracle.color_jp(g)[:2]):
            out, rounds = oracle.color_balance(g, base, mc + 1, type)
            _balance_invariants(g, base, out, mc + 1, type)
            assert rounds >= 1 or np.array_equal(out, base)


def test_balance_least_used_picks_smallest_class():
    """type 1 on a star whose leaves are all in one oversize class: the
    leaves go to the emptiest eligible color first."""
    g = Graph.from_pairs(9, np.array([(0, i) for i in range(1, 9)], np.int64))
    out, _ = oracle.color_balance(g, np.array([0, 1, 1, 1, 1, 1, 1, 1, 2], np.int32), 3, 1)
    # avg = 3; class 1 (7 leaves) is oversize; color 0 is the centre's, so
    # color 2 takes leaves until it holds 3
    assert np.bincount(out, minlength=3).tolist() == [1, 5, 3]
