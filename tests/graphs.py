"""Small seeded graphs shared by the tests (sizes the oracle finishes in seconds)."""
import os

import numpy as np

from grappolo_amd.graph import Graph, gen_lfr, gen_planted, gen_rmat, load_metis

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def karate():
    return load_metis(os.path.join(GOLDEN, "karate.graph"))


def hubs(seed=7, nv=12000, hub_degrees=(300, 1500, 2500, 6000)):
    """Planted background plus a few hub vertices: every degree bin is used,
    including the HBM-hash one (> 2048 entries)."""
    g = gen_planted(nv, 32, 12, 0.3, seed=seed)
    rng = np.random.default_rng(seed)
    rows = np.repeat(np.arange(nv), g.degrees())
    keep = rows < g.tails
    pairs = [np.stack([rows[keep], g.tails[keep]], 1)]
    for h, d in enumerate(hub_degrees):
        others = rng.choice(np.arange(len(hub_degrees), nv), size=d, replace=False)
        pairs.append(np.stack([np.full(d, h), others], 1))
    p = np.concatenate(pairs)
    p = np.unique(np.sort(p, 1), axis=0)
    return Graph.from_pairs(nv, p)


def with_loops_and_isolated(seed=11, nv=3000):
    """Self-loops (weight 2), isolated vertices and a weighted planted core."""
    g = gen_planted(nv - 100, 10, 10, 0.25, seed=seed, wlo=0.5, whi=2.0)
    rows = np.repeat(np.arange(g.numVertices), g.degrees())
    keep = rows < g.tails
    pairs = np.stack([rows[keep], g.tails[keep]], 1)
    w = g.weights[keep]
    loops = np.arange(0, g.numVertices, 7)
    pairs = np.concatenate([pairs, np.stack([loops, loops], 1)])
    w = np.concatenate([w, np.full(len(loops), 2.0)])
    return Graph.from_pairs(nv, pairs, w)  # vertices nv-100.. have no edges


def pendants(seed=17, nv=5000, weighted=True):
    """What vertex following folds: a planted core with pendant (degree-one)
    vertices hung on it, disjoint pairs of degree-one vertices, vertices whose
    only entry is a self-loop, and isolated vertices.  Weighted graphs carry
    fractional weights, which the folded graph truncates."""
    rng = np.random.default_rng(seed)
    core = nv // 2
    g = gen_planted(core, 8, 8, 0.3, seed=seed)
    rows = np.repeat(np.arange(g.numVertices), g.degrees())
    keep = rows < g.tails
    pairs = [np.stack([rows[keep], g.tails[keep]], 1)]
    v = core
    npend = nv // 4
    pairs.append(np.stack([rng.integers(0, core, npend), np.arange(v, v + npend)], 1))
    v += npend
    npair = (nv // 10) // 2
    a = np.arange(v, v + 2 * npair, 2)
    pairs.append(np.stack([a, a + 1], 1))
    v += 2 * npair
    nloop = nv // 20
    pairs.append(np.stack([np.arange(v, v + nloop), np.arange(v, v + nloop)], 1))
    v += nloop  # the rest stay isolated
    p = np.concatenate(pairs)
    w = rng.uniform(0.4, 3.7, len(p)) if weighted else None
    return Graph.from_pairs(nv, p, w)


def cliques(seed=13, nv=20000, size=200, inter=8, hub_degrees=(5000, 4500), weighted=False):
    """Dense near-cliques (every coloring needs ~size colors, so classes hold
    ~nv/size vertices) joined by random inter-clique edges, plus hubs: a graph
    above the dense-accumulator limit that the persistent sweep serves with its
    LDS hash (and, past kPHashMaxDeg, the dense view in slabs or, above two
    slabs of vertices, the HBM table)."""
    rng = np.random.default_rng(seed)
    blocks = nv // size
    pairs = []
    for b in range(blocks):
        base = b * size
        iu = np.triu_indices(size, 1)
        keep = rng.random(len(iu[0])) < 0.97
        pairs.append(np.stack([iu[0][keep] + base, iu[1][keep] + base], 1))
    m = nv * inter // 2
    u, v = rng.integers(0, nv, m), rng.integers(0, nv, m)
    pairs.append(np.stack([u, v], 1)[u != v])
    for h, d in enumerate(hub_degrees):
        others = rng.choice(np.arange(nv), size=d, replace=False)
        others = others[others != h]
        pairs.append(np.stack([np.full(len(others), h), others], 1))
    p = np.unique(np.sort(np.concatenate(pairs), 1), axis=0)
    w = rng.uniform(0.5, 2.0, len(p)) if weighted else None
    return Graph.from_pairs(nv, p, w)


def bipartite_blocks(seed=19, sides=((350, 350, 0.9), (300, 330, 0.95)), extra=200, weighted=False):
    """Dense near-complete bipartite blocks plus a few random edges: the
    first-coloured side of a block takes one color and the other side the
    next, so a vertex of the second side has about 300 neighbours in the
    preceding color class (more than the persistent dense sweep sets aside
    while it gathers ahead of a color step)."""
    rng = np.random.default_rng(seed)
    pairs, base = [], 0
    for nx, ny, dens in sides:
        xs, ys = np.meshgrid(np.arange(nx), np.arange(ny), indexing="ij")
        keep = rng.random(xs.shape) < dens
        pairs.append(np.stack([xs[keep] + base, ys[keep] + base + nx], 1))
        base += nx + ny
    u, v = rng.integers(0, base, extra), rng.integers(0, base, extra)
    pairs.append(np.stack([u, v], 1)[u != v])
    p = np.unique(np.sort(np.concatenate(pairs), 1), axis=0)
    w = rng.uniform(0.5, 2.0, len(p)) if weighted else None
    return Graph.from_pairs(base, p, w)


def catalog():
    return {
        "karate": karate(),
        "planted_unit": gen_planted(4096, 16, 16, 0.3, seed=5),
        "planted_w": gen_planted(4096, 16, 16, 0.3, seed=6, wlo=0.5, whi=2.0),
        "lfr_w": gen_lfr(20000, seed=3, dmax=600, cmax=1500),
        "rmat12": gen_rmat(12, seed=3),
        "hubs": hubs(),
        # above the dense-accumulator limit: degree > 2048 uses the HBM hash
        "hubs_big": hubs(seed=8, nv=40000, hub_degrees=(3000, 9000, 2100, 700)),
        # degree > 16384: several workgroups per vertex (sweep, modularity, coloring)
        "hubs_giant": hubs(seed=9, nv=60000, hub_degrees=(20000, 31000, 17000, 5000)),
        "loops_isolated": with_loops_and_isolated(),
        # vertices with ~300 neighbours in the preceding color class
        "bipartite": bipartite_blocks(),
    }


def unit_weights(g):
    return bool(np.all(g.weights == np.round(g.weights)))
