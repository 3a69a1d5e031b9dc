"""Graph container and the seeded generators / file formats (libglv_host.so).

``Graph`` mirrors the reference's ``graph`` record (DefineStructure/defs.h:85-92):
``numVertices``, ``numEdges`` and the CSR ``edgeListPtrs``; the reference's
array-of-structs ``edgeList`` of ``{head, tail, weight}`` is held as two
arrays ``tails`` (int64) and ``weights`` (float64), which is what the engine
uploads.  ``edge_list()`` rebuilds the 24-byte AoS records when a caller needs
the reference layout.
"""
import ctypes
import os

import numpy as np

from ._native import host_lib

EDGE_DTYPE = np.dtype([("head", np.int64), ("tail", np.int64), ("weight", np.float64)])
assert EDGE_DTYPE.itemsize == 24


class _CSR(ctypes.Structure):
    _fields_ = [("nv", ctypes.c_int64), ("nnz", ctypes.c_int64), ("num_edges", ctypes.c_int64),
                ("off", ctypes.POINTER(ctypes.c_int64)), ("tails", ctypes.POINTER(ctypes.c_int64)),
                ("w", ctypes.POINTER(ctypes.c_double))]


class Graph:
    """Symmetric CSR graph: rows ``tails[edgeListPtrs[i]:edgeListPtrs[i+1]]``."""

    def __init__(self, edgeListPtrs, tails, weights, numEdges=None):
        self.edgeListPtrs = np.ascontiguousarray(edgeListPtrs, dtype=np.int64)
        self.tails = np.ascontiguousarray(tails, dtype=np.int64)
        self.weights = np.ascontiguousarray(weights, dtype=np.float64)
        self.numVertices = int(self.edgeListPtrs.shape[0] - 1)
        nnz = int(self.edgeListPtrs[-1]) if self.numVertices >= 0 else 0
        if self.tails.shape[0] != nnz or self.weights.shape[0] != nnz:
            raise ValueError("tails/weights length must equal edgeListPtrs[-1]")
        self.numEdges = int(numEdges) if numEdges is not None else nnz // 2

    @property
    def nnz(self):
        return int(self.edgeListPtrs[-1])

    @property
    def sVertices(self):
        return self.numVertices

    def degrees(self):
        return np.diff(self.edgeListPtrs)

    def edge_list(self):
        """The reference's AoS ``edge`` records (head, tail, weight)."""
        e = np.empty(self.nnz, dtype=EDGE_DTYPE)
        e["head"] = np.repeat(np.arange(self.numVertices, dtype=np.int64), self.degrees())
        e["tail"] = self.tails
        e["weight"] = self.weights
        return e

    @classmethod
    def from_edge_list(cls, edgeListPtrs, edgeList, numEdges=None):
        edgeList = np.asarray(edgeList, dtype=EDGE_DTYPE)
        return cls(edgeListPtrs, edgeList["tail"].copy(), edgeList["weight"].copy(), numEdges)

    @classmethod
    def from_pairs(cls, nv, pairs, weights=None):
        """Symmetric CSR (rows sorted by tail) from undirected pairs; self-loops
        are stored once, other pairs in both rows."""
        pairs = np.asarray(pairs, dtype=np.int64).reshape(-1, 2)
        w = np.ones(len(pairs)) if weights is None else np.asarray(weights, dtype=np.float64)
        u, v = pairs[:, 0], pairs[:, 1]
        loop = u == v
        src = np.concatenate([u, v[~loop]])
        dst = np.concatenate([v, u[~loop]])
        ww = np.concatenate([w, w[~loop]])
        order = np.lexsort((dst, src))
        src, dst, ww = src[order], dst[order], ww[order]
        ptr = np.zeros(nv + 1, dtype=np.int64)
        np.add.at(ptr, src + 1, 1)
        ptr = np.cumsum(ptr)
        return cls(ptr, dst, ww, numEdges=int((~loop).sum() + loop.sum()))

    def copy(self):
        return Graph(self.edgeListPtrs.copy(), self.tails.copy(), self.weights.copy(),
                     self.numEdges)

    def __repr__(self):
        return f"Graph(numVertices={self.numVertices}, nnz={self.nnz}, numEdges={self.numEdges})"


def union(graphs):
    """Block-diagonal union of independent graphs: (Graph, vbase) where graph
    i owns the vertices vbase[i]..vbase[i+1]-1 (the batched entry points'
    input)."""
    vb = np.zeros(len(graphs) + 1, dtype=np.int64)
    eb = np.zeros(len(graphs) + 1, dtype=np.int64)
    for i, g in enumerate(graphs):
        vb[i + 1] = vb[i] + g.numVertices
        eb[i + 1] = eb[i] + g.nnz
    ptr = np.zeros(vb[-1] + 1, dtype=np.int64)
    tails = np.empty(eb[-1], dtype=np.int64)
    w = np.empty(eb[-1], dtype=np.float64)
    for i, g in enumerate(graphs):
        ptr[vb[i] + 1:vb[i + 1] + 1] = g.edgeListPtrs[1:] + eb[i]
        tails[eb[i]:eb[i + 1]] = g.tails + vb[i]
        w[eb[i]:eb[i + 1]] = g.weights
    return Graph(ptr, tails, w, numEdges=sum(g.numEdges for g in graphs)), vb


def _fn(name, *argtypes):
    f = getattr(host_lib(), name)
    f.argtypes = list(argtypes)
    f.restype = ctypes.c_int
    return f


def _take(csr):
    nv, nnz = csr.nv, csr.nnz
    off = np.ctypeslib.as_array(csr.off, shape=(nv + 1,)).copy()
    tails = np.ctypeslib.as_array(csr.tails, shape=(max(nnz, 1),))[:nnz].copy()
    w = np.ctypeslib.as_array(csr.w, shape=(max(nnz, 1),))[:nnz].copy()
    g = Graph(off, tails, w, numEdges=csr.num_edges)
    free = host_lib().glv_csr_free
    free.argtypes = [ctypes.POINTER(_CSR)]
    free.restype = None
    free(ctypes.byref(csr))
    return g


def _check(rc, what):
    if rc != 0:
        raise RuntimeError(f"{what} failed with status {rc}")


_I64P = ctypes.POINTER(ctypes.c_int64)


def gen_planted(nv, nblocks, mean_degree, mu, seed, wlo=1.0, whi=1.0, return_blocks=False):
    """Planted partition (the workload of the reference's CPU-runnable config:
    nv=65536, 64 blocks, mean degree 16, mu 0.3, unit weights, seed 1)."""
    f = _fn("glv_gen_planted", ctypes.c_int64, ctypes.c_int64, ctypes.c_double, ctypes.c_double,
            ctypes.c_double, ctypes.c_double, ctypes.c_uint64, _I64P, ctypes.POINTER(_CSR))
    blocks = np.empty(nv, dtype=np.int64)
    csr = _CSR()
    _check(f(nv, nblocks, mean_degree, mu, wlo, whi, seed,
             blocks.ctypes.data_as(_I64P), ctypes.byref(csr)), "glv_gen_planted")
    g = _take(csr)
    return (g, blocks) if return_blocks else g


def gen_planted_many(specs, workers=None):
    """gen_planted over a list of (nv, nblocks, mean_degree, mu, seed), the
    graphs generated concurrently one per host thread (the generator's output
    does not depend on its thread count)."""
    from concurrent.futures import ThreadPoolExecutor
    workers = workers or min(16, os.cpu_count() or 1)
    set_threads(1)
    try:
        with ThreadPoolExecutor(max_workers=workers) as ex:
            return list(ex.map(lambda sp: gen_planted(*sp[:4], seed=sp[4]), specs))
    finally:
        set_threads(0)


def gen_lfr(nv, seed, deg_exp=2.5, dmin=10, dmax=1000, comm_exp=1.5, cmin=20, cmax=5000, mu=0.4,
            wlo=0.5, whi=2.0, return_communities=False):
    """LFR-style graph (defaults: the 4M-vertex config's parameters)."""
    f = _fn("glv_gen_lfr", ctypes.c_int64, ctypes.c_double, ctypes.c_int64, ctypes.c_int64,
            ctypes.c_double, ctypes.c_int64, ctypes.c_int64, ctypes.c_double, ctypes.c_double,
            ctypes.c_double, ctypes.c_uint64, _I64P, ctypes.POINTER(_CSR))
    comm = np.empty(nv, dtype=np.int64)
    csr = _CSR()
    _check(f(nv, deg_exp, dmin, dmax, comm_exp, cmin, cmax, mu, wlo, whi, seed,
             comm.ctypes.data_as(_I64P), ctypes.byref(csr)), "glv_gen_lfr")
    g = _take(csr)
    return (g, comm) if return_communities else g


def gen_rmat(scale, seed, edge_factor=16, a=0.57, b=0.19, c=0.19):
    f = _fn("glv_gen_rmat", ctypes.c_int, ctypes.c_int64, ctypes.c_double, ctypes.c_double,
            ctypes.c_double, ctypes.c_uint64, ctypes.POINTER(_CSR))
    csr = _CSR()
    _check(f(scale, edge_factor, a, b, c, seed, ctypes.byref(csr)), "glv_gen_rmat")
    return _take(csr)


def write_binary(path, g):
    """The reference's -f 9 file (writeBinary.cpp:2-29)."""
    f = _fn("glv_write_binary", ctypes.c_char_p, ctypes.POINTER(_CSR))
    csr = _CSR(g.numVertices, g.nnz, g.numEdges, g.edgeListPtrs.ctypes.data_as(_I64P),
               g.tails.ctypes.data_as(_I64P),
               g.weights.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    _check(f(str(path).encode(), ctypes.byref(csr)), "glv_write_binary")


def read_binary(path):
    f = _fn("glv_read_binary", ctypes.c_char_p, ctypes.POINTER(_CSR))
    csr = _CSR()
    _check(f(str(path).encode(), ctypes.byref(csr)), "glv_read_binary")
    return _take(csr)


def load_metis(path):
    f = _fn("glv_load_metis", ctypes.c_char_p, ctypes.POINTER(_CSR))
    csr = _CSR()
    _check(f(str(path).encode(), ctypes.byref(csr)), "glv_load_metis")
    return _take(csr)


def set_threads(n):
    f = host_lib().glv_gen_set_threads
    f.argtypes = [ctypes.c_int]
    f.restype = None
    f(int(n))
