"""CPU checkers for the HIP engine.  TEST INFRASTRUCTURE ONLY.

``oracle`` wraps ``build/liboracle.so`` (grappolo_oracle.c, the C restatement
of the reference's colored multi-phase Louvain) and, where it was built,
``_ref/libgrappolo_ref.so`` (the reference's own sources compiled by
oracle/Makefile).  Only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg import this package.
"""
import ctypes
import os
import subprocess
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_LIB = os.path.join(HERE, "build", "liboracle.so")
REF_LIB = os.path.join(HERE, "_ref", "libgrappolo_ref.so")
REF_BENCH = os.path.join(HERE, "_ref", "ref_bench")
REF_DIR = os.environ.get("GRAPPOLO_REF_DIR", "/root/reference")

_I64P = ctypes.POINTER(ctypes.c_int64)
_I32P = ctypes.POINTER(ctypes.c_int32)
_F64P = ctypes.POINTER(ctypes.c_double)


def build(ref=True, quiet=True):
    """Compile the restatement, and the reference checker when its sources exist."""
    out = subprocess.DEVNULL if quiet else None
    subprocess.run(["make", "-C", HERE, "-j8"], check=True, stdout=out)
    if ref and os.path.isdir(REF_DIR):
        subprocess.run(["make", "-C", HERE, "-j8", "ref", f"REF={REF_DIR}"], check=True, stdout=out)


_lib = None
_ref = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_LIB):
            build(ref=False)
        _lib = ctypes.CDLL(ORACLE_LIB)
        _declare(_lib)
    return _lib


def ref_available():
    return os.path.exists(REF_LIB)


def ref():
    global _ref
    if _ref is None:
        if not ref_available():
            raise RuntimeError("reference checker not built (oracle/_ref/libgrappolo_ref.so)")
        _ref = ctypes.CDLL(REF_LIB)
        _declare_ref(_ref)
    return _ref


class Comparison(ctypes.Structure):
    _fields_ = [("nC1", ctypes.c_int64), ("nC2", ctypes.c_int64), ("same_same", ctypes.c_int64),
                ("same_diff", ctypes.c_int64), ("diff_same", ctypes.c_int64), ("agree", ctypes.c_int64),
                ("precision", ctypes.c_double), ("recall", ctypes.c_double), ("f_score", ctypes.c_double),
                ("gini1", ctypes.c_double), ("gini2", ctypes.c_double)]


def comparison_dict(c):
    return {f: getattr(c, f) for f, _ in c._fields_}


class Stats(ctypes.Structure):
    _fields_ = [("phases", ctypes.c_int32), ("total_iterations", ctypes.c_int32),
                ("final_clusters", ctypes.c_int64), ("final_modularity", ctypes.c_double),
                ("last_modularity", ctypes.c_double), ("edges_visited", ctypes.c_int64),
                ("colors_first", ctypes.c_int32), ("n_phase", ctypes.c_int32),
                ("phase_cap", ctypes.c_int32), ("phase_nv", _I64P), ("phase_nnz", _I64P),
                ("phase_iters", _I32P), ("phase_colors", _I32P), ("phase_colored", _I32P),
                ("phase_mod", _F64P), ("phase_clusters", _I64P)]


def _declare(L):
    G = [ctypes.c_int64, _I64P, _I64P, _F64P]
    L.orc_constant.argtypes = [ctypes.c_int64, _F64P]
    L.orc_constant.restype = ctypes.c_double
    L.orc_degrees.argtypes = [ctypes.c_int64, _I64P, _F64P, _F64P]
    L.orc_degrees.restype = None
    L.orc_color_greedy.argtypes = [ctypes.c_int64, _I64P, _I64P, _I32P]
    L.orc_color_greedy.restype = ctypes.c_int
    L.orc_color_jp.argtypes = [ctypes.c_int64, _I64P, _I64P, _I32P, _I32P]
    L.orc_color_jp.restype = ctypes.c_int
    L.orc_jp_priority.argtypes = [ctypes.c_int64]
    L.orc_jp_priority.restype = ctypes.c_uint64
    L.orc_color_conflicts.argtypes = [ctypes.c_int64, _I64P, _I64P, _I32P]
    L.orc_color_conflicts.restype = ctypes.c_int64
    L.orc_louvain_colored.argtypes = G + [_I64P, _I32P, ctypes.c_int, ctypes.c_double,
                                          ctypes.c_double, ctypes.c_int, _I32P, _F64P,
                                          ctypes.c_int32]
    L.orc_louvain_colored.restype = ctypes.c_double
    L.orc_louvain_plain.argtypes = G + [_I64P, ctypes.c_double, ctypes.c_double, ctypes.c_int,
                                        _I32P, _F64P, ctypes.c_int32]
    L.orc_louvain_plain.restype = ctypes.c_double
    L.orc_color_step.argtypes = G + [_I64P, _I32P, ctypes.c_int32]
    L.orc_color_step.restype = ctypes.c_int64
    L.orc_modularity.argtypes = G + [_I64P]
    L.orc_modularity.restype = ctypes.c_double
    L.orc_renumber.argtypes = [_I64P, ctypes.c_int64]
    L.orc_renumber.restype = ctypes.c_int64
    L.orc_coarsen.argtypes = G + [_I64P, ctypes.c_int64, _I64P, _I64P, _F64P, _I64P]
    L.orc_coarsen.restype = ctypes.c_int64
    L.orc_build_vf.argtypes = G + [_I64P, ctypes.c_int64, _I64P, _I64P, _F64P, _I64P]
    L.orc_build_vf.restype = ctypes.c_int64
    L.orc_vertex_following.argtypes = [ctypes.c_int64, _I64P, _I64P, _I64P]
    L.orc_vertex_following.restype = ctypes.c_int64
    L.orc_compare.argtypes = [_I64P, _I64P, ctypes.c_int64, ctypes.POINTER(Comparison)]
    L.orc_compare.restype = ctypes.c_int
    L.orc_gini.argtypes = [_I64P, ctypes.c_int]
    L.orc_gini.restype = ctypes.c_double
    L.orc_run_colored.argtypes = G + [_I64P, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                      ctypes.c_int64, ctypes.c_double, ctypes.c_double,
                                      ctypes.POINTER(Stats)]
    L.orc_mh_value.argtypes = [ctypes.c_int, ctypes.c_int64]
    L.orc_mh_value.restype = ctypes.c_double
    L.orc_color_multihash.argtypes = [ctypes.c_int64, _I64P, _I64P, ctypes.c_int, ctypes.c_int, _I32P]
    L.orc_color_multihash.restype = ctypes.c_int
    L.orc_color_multihash_vals.argtypes = [ctypes.c_int64, _I64P, _I64P, ctypes.c_int, ctypes.c_int, _F64P,
                                           _I32P]
    L.orc_color_multihash_vals.restype = ctypes.c_int
    L.orc_color_balance.argtypes = [ctypes.c_int64, _I64P, _I64P, _I32P, ctypes.c_int, ctypes.c_int]
    L.orc_color_balance.restype = ctypes.c_int
    L.orc_run_colored.restype = ctypes.c_int
    L.orc_run_basic.argtypes = G + [_I64P, ctypes.c_int, ctypes.c_double, ctypes.POINTER(Stats)]
    L.orc_run_basic.restype = ctypes.c_int


def _declare_ref(R):
    G = [ctypes.c_int64, ctypes.c_int64, _I64P, _I64P, _F64P]
    R.ref_color.argtypes = G + [_I32P, ctypes.c_int]
    R.ref_color.restype = ctypes.c_int
    R.ref_colored_phase.argtypes = G + [_I64P, _I32P, ctypes.c_int, ctypes.c_double,
                                        ctypes.c_double, ctypes.c_int, ctypes.c_int, _I32P]
    R.ref_colored_phase.restype = ctypes.c_double
    R.ref_plain_phase.argtypes = G + [_I64P, ctypes.c_double, ctypes.c_double, ctypes.c_int,
                                      ctypes.c_int, _I32P]
    R.ref_plain_phase.restype = ctypes.c_double
    R.ref_renumber.argtypes = [_I64P, ctypes.c_int64]
    R.ref_renumber.restype = ctypes.c_int64
    R.ref_coarsen.argtypes = G + [_I64P, ctypes.c_int64, ctypes.c_int, ctypes.POINTER(_I64P),
                                  ctypes.POINTER(_I64P), ctypes.POINTER(_F64P), _I64P]
    R.ref_coarsen.restype = ctypes.c_int64
    R.ref_build_vf.argtypes = G + [_I64P, ctypes.c_int64, ctypes.POINTER(_I64P), ctypes.POINTER(_I64P),
                                   ctypes.POINTER(_F64P), _I64P]
    R.ref_build_vf.restype = ctypes.c_int64
    R.ref_vertex_following.argtypes = [ctypes.c_int64, ctypes.c_int64, _I64P, _I64P, _I64P]
    R.ref_vertex_following.restype = ctypes.c_int64
    R.ref_compare.argtypes = [_I64P, _I64P, ctypes.c_int64, ctypes.c_int, ctypes.c_char_p]
    R.ref_compare.restype = None
    R.ref_gini.argtypes = [_I64P, ctypes.c_int]
    R.ref_gini.restype = ctypes.c_double
    R.ref_multihash.argtypes = [ctypes.c_int64, ctypes.c_int64, _I64P, _I64P, ctypes.c_int, ctypes.c_int,
                                ctypes.c_uint, _I32P, _F64P]
    R.ref_multihash.restype = ctypes.c_int
    R.ref_free.argtypes = [ctypes.c_void_p]
    R.ref_free.restype = None
    R.ref_run_colored.argtypes = G + [_I64P, ctypes.c_int, ctypes.c_int64, ctypes.c_double,
                                      ctypes.c_double, ctypes.c_int, ctypes.c_char_p]
    R.ref_run_colored.restype = None
    R.ref_run_basic.argtypes = G + [_I64P, ctypes.c_int, ctypes.c_double, ctypes.c_int,
                                    ctypes.c_char_p]
    R.ref_run_basic.restype = None
    for name in ("ref_load_binary", "ref_load_metis"):
        f = getattr(R, name)
        f.argtypes = [ctypes.c_char_p, ctypes.POINTER(_I64P), ctypes.POINTER(_I64P),
                      ctypes.POINTER(_F64P), _I64P]
        f.restype = ctypes.c_int64
    R.ref_write_binary.argtypes = [ctypes.c_char_p] + G
    R.ref_write_binary.restype = None


def _p(a, t):
    return a.ctypes.data_as(t)


def _g(g):
    return [g.numVertices, _p(g.edgeListPtrs, _I64P), _p(g.tails, _I64P), _p(g.weights, _F64P)]


def _rg(g):
    return [g.numVertices, g.numEdges, _p(g.edgeListPtrs, _I64P), _p(g.tails, _I64P),
            _p(g.weights, _F64P)]


# ---------------------------------------------------------------- restatement
def degrees(g):
    k = np.empty(g.numVertices)
    lib().orc_degrees(g.numVertices, _p(g.edgeListPtrs, _I64P), _p(g.weights, _F64P), _p(k, _F64P))
    return k


def color_greedy(g):
    c = np.empty(g.numVertices, dtype=np.int32)
    mc = lib().orc_color_greedy(g.numVertices, _p(g.edgeListPtrs, _I64P), _p(g.tails, _I64P),
                                _p(c, _I32P))
    if mc < -1:
        raise RuntimeError("coloring needs more than 4096 colors")
    return c, mc


def color_jp(g):
    c = np.empty(g.numVertices, dtype=np.int32)
    rounds = ctypes.c_int32(0)
    mc = lib().orc_color_jp(g.numVertices, _p(g.edgeListPtrs, _I64P), _p(g.tails, _I64P),
                            _p(c, _I32P), ctypes.byref(rounds))
    if mc < -1:
        raise RuntimeError("coloring needs more than 4096 colors")
    return c, mc, rounds.value


def jp_priority(v):
    return lib().orc_jp_priority(int(v))


def color_conflicts(g, colors):
    colors = np.ascontiguousarray(colors, dtype=np.int32)
    return lib().orc_color_conflicts(g.numVertices, _p(g.edgeListPtrs, _I64P), _p(g.tails, _I64P),
                                     _p(colors, _I32P))


def louvain_colored(g, colors, ncolors, lower=-1.0, thresh=1e-6, frozen=True, trace_cap=0):
    C = np.empty(g.numVertices, dtype=np.int64)
    colors = np.ascontiguousarray(colors, dtype=np.int32)
    it = ctypes.c_int32(0)
    tr = np.zeros(max(trace_cap, 1))
    mod = lib().orc_louvain_colored(*_g(g), _p(C, _I64P), _p(colors, _I32P), int(ncolors),
                                    lower, thresh, int(frozen), ctypes.byref(it), _p(tr, _F64P),
                                    trace_cap)
    return mod, C, it.value, tr[:min(trace_cap, it.value)]


def louvain_plain(g, lower=-1.0, thresh=1e-6, nomap=False, trace_cap=0):
    C = np.empty(g.numVertices, dtype=np.int64)
    it = ctypes.c_int32(0)
    tr = np.zeros(max(trace_cap, 1))
    mod = lib().orc_louvain_plain(*_g(g), _p(C, _I64P), lower, thresh, int(nomap),
                                  ctypes.byref(it), _p(tr, _F64P), trace_cap)
    return mod, C, it.value, tr[:min(trace_cap, it.value)]


def color_step(g, C, colors, which):
    C = np.array(C, dtype=np.int64)
    colors = np.ascontiguousarray(colors, dtype=np.int32)
    moves = lib().orc_color_step(*_g(g), _p(C, _I64P), _p(colors, _I32P), int(which))
    return C, moves


def modularity(g, C):
    C = np.ascontiguousarray(C, dtype=np.int64)
    return lib().orc_modularity(*_g(g), _p(C, _I64P))


def renumber(C):
    C = np.array(C, dtype=np.int64)
    k = lib().orc_renumber(_p(C, _I64P), C.shape[0])
    return C, k


def _contract(fn, g, C, k):
    from grappolo_amd.graph import Graph
    C = np.ascontiguousarray(C, dtype=np.int64)
    ne = ctypes.c_int64(0)
    tot = fn(*_g(g), _p(C, _I64P), k, None, None, None, ctypes.byref(ne))
    off = np.empty(k + 1, dtype=np.int64)
    tails = np.empty(max(tot, 1), dtype=np.int64)
    w = np.empty(max(tot, 1))
    fn(*_g(g), _p(C, _I64P), k, _p(off, _I64P), _p(tails, _I64P), _p(w, _F64P), ctypes.byref(ne))
    return Graph(off, tails[:tot], w[:tot], numEdges=ne.value)


def coarsen(g, C, k):
    return _contract(lib().orc_coarsen, g, C, k)


def build_vf(g, C, k):
    """buildNewGraphVF restated; rows sorted by tail."""
    return _contract(lib().orc_build_vf, g, C, k)


def vertex_following(g):
    """vertexFollowing restated: (C, numNode)."""
    C = np.empty(g.numVertices, np.int64)
    n = lib().orc_vertex_following(g.numVertices, _p(g.edgeListPtrs, _I64P), _p(g.tails, _I64P),
                                   _p(C, _I64P))
    return C, n


def _stats(cap=256):
    st = Stats()
    arrays = dict(phase_nv=np.zeros(cap, np.int64), phase_nnz=np.zeros(cap, np.int64),
                  phase_iters=np.zeros(cap, np.int32), phase_colors=np.zeros(cap, np.int32),
                  phase_colored=np.zeros(cap, np.int32), phase_mod=np.zeros(cap),
                  phase_clusters=np.zeros(cap, np.int64))
    st.phase_cap = cap
    for name, arr in arrays.items():
        t = dict(Stats._fields_)[name]
        setattr(st, name, arr.ctypes.data_as(t))
    return st, arrays


def _stats_dict(st, arrays):
    n = min(st.n_phase, st.phase_cap)
    d = {f: getattr(st, f) for f, _ in Stats._fields_[:8]}
    d["per_phase"] = {k: v[:n].copy() for k, v in arrays.items()}
    return d


_COLORING_MODES = {"greedy": 0, "jp": 1, "balanced": 2, "multihash": 3}


def run_colored(g, coloring="greedy", frozen=False, replaceMap=0, minGraphSize=1000,
                threshold=1e-9, C_threshold=1e-6, numColors=16):
    """runMultiPhaseColoring.  coloring='greedy' with frozen=False is the
    reference at one thread (coloring=1); 'jp', 'balanced' and 'multihash'
    with frozen=True are the engine's coloring=1, 2 and 3."""
    C_orig = np.empty(g.numVertices, dtype=np.int64)
    st, arrays = _stats()
    rc = lib().orc_run_colored(*_g(g), _p(C_orig, _I64P), _COLORING_MODES[coloring], int(numColors),
                               int(frozen), int(replaceMap), int(minGraphSize), threshold,
                               C_threshold, ctypes.byref(st))
    if rc != 0:
        raise RuntimeError(f"orc_run_colored failed ({rc})")
    return C_orig, _stats_dict(st, arrays)


def color_multihash(g, nHash=2, nItrs=4):
    """(colors, maxColor) of the restated incomplete coloring."""
    c = np.empty(g.numVertices, np.int32)
    mc = lib().orc_color_multihash(g.numVertices, _p(g.edgeListPtrs, _I64P), _p(g.tails, _I64P), nHash,
                                   nItrs, _p(c, _I32P))
    return c, mc


def color_balance(g, colors, ncolors, type=0):
    """(recolored copy, rounds) of the restated balanced recoloring."""
    c = np.array(colors, dtype=np.int32)
    r = lib().orc_color_balance(g.numVertices, _p(g.edgeListPtrs, _I64P), _p(g.tails, _I64P), _p(c, _I32P),
                                int(ncolors), int(type))
    if r < 0:
        raise ValueError(f"orc_color_balance failed ({r})")
    return c, r


def color_multihash_vals(g, nHash, nItrs, vals):
    c = np.empty(g.numVertices, np.int32)
    vals = np.ascontiguousarray(vals, dtype=np.float64)
    mc = lib().orc_color_multihash_vals(g.numVertices, _p(g.edgeListPtrs, _I64P), _p(g.tails, _I64P), nHash,
                                        nItrs, _p(vals, _F64P), _p(c, _I32P))
    return c, mc


def ref_multihash(g, nHash, nItrs, seed=1):
    """(colors, maxColor, the rand() values drawn) of the compiled reference."""
    c = np.empty(g.numVertices, np.int32)
    vals = np.empty(max(nHash * g.numVertices, 1))
    mc = ref().ref_multihash(g.numVertices, g.numEdges, _p(g.edgeListPtrs, _I64P), _p(g.tails, _I64P), nHash,
                             nItrs, seed, _p(c, _I32P), _p(vals, _F64P))
    return c, mc, vals[:nHash * g.numVertices]


def mh_value(h, v):
    return lib().orc_mh_value(h, v)


def run_basic(g, basicOpt=0, threshold=1e-9):
    C_orig = np.empty(g.numVertices, dtype=np.int64)
    st, arrays = _stats()
    rc = lib().orc_run_basic(*_g(g), _p(C_orig, _I64P), int(basicOpt), threshold,
                             ctypes.byref(st))
    if rc != 0:
        raise RuntimeError(f"orc_run_basic failed ({rc})")
    return C_orig, _stats_dict(st, arrays)


def canonical(g):
    """Rows sorted by (tail, weight): the order-free form of a CSR graph."""
    from grappolo_amd.graph import Graph
    rows = np.repeat(np.arange(g.numVertices), g.degrees())
    order = np.lexsort((g.weights, g.tails, rows))
    return Graph(g.edgeListPtrs.copy(), g.tails[order], g.weights[order], g.numEdges)


# ------------------------------------------------------------ compiled reference
def ref_color(g, nthreads=1):
    c = np.empty(g.numVertices, dtype=np.int32)
    mc = ref().ref_color(*_rg(g), _p(c, _I32P), nthreads)
    return c, mc


def ref_colored_phase(g, colors, ncolors, lower=-1.0, thresh=1e-6, nomap=False, nthreads=1):
    C = np.empty(g.numVertices, dtype=np.int64)
    colors = np.array(colors, dtype=np.int32)
    it = ctypes.c_int32(0)
    mod = ref().ref_colored_phase(*_rg(g), _p(C, _I64P), _p(colors, _I32P), int(ncolors), lower,
                                  thresh, int(nomap), nthreads, ctypes.byref(it))
    return mod, C, it.value


def ref_plain_phase(g, lower=-1.0, thresh=1e-6, nomap=False, nthreads=1):
    C = np.empty(g.numVertices, dtype=np.int64)
    it = ctypes.c_int32(0)
    mod = ref().ref_plain_phase(*_rg(g), _p(C, _I64P), lower, thresh, int(nomap), nthreads,
                                ctypes.byref(it))
    return mod, C, it.value


def compare(C1, C2):
    """computeCommunityComparisons restated: (dict, relabelled C1, C2)."""
    a = np.array(C1, dtype=np.int64)
    b = np.array(C2, dtype=np.int64)
    out = Comparison()
    if lib().orc_compare(_p(a, _I64P), _p(b, _I64P), a.shape[0], ctypes.byref(out)) != 0:
        raise ValueError("labels the reference cannot compare")
    return comparison_dict(out), a, b


def gini(sizes):
    """computeGiniCoefficient restated: (gini, sorted sizes)."""
    s = np.array(sizes, dtype=np.int64)
    return lib().orc_gini(_p(s, _I64P), s.shape[0]), s


_REPORT = {"Same-Same (True positive)": ("same_same", int), "Same-Diff (False negative)": ("same_diff", int),
           "Diff-Same (False positive)": ("diff_same", int), "Num communities in C1": ("nC1", int),
           "Num communities in C2": ("nC2", int), "Precision": ("precision", float),
           "Recall": ("recall", float), "F-score": ("f_score", float),
           "Gini coefficient, C1": ("gini1", float), "Gini coefficient, C2": ("gini2", float)}


def ref_compare(C1, C2, nthreads=1):
    """The compiled reference's computeCommunityComparisons; its printed
    report parsed into a dict (floats as printed, %lf), plus the relabelled
    vectors."""
    import tempfile
    a = np.array(C1, dtype=np.int64)
    b = np.array(C2, dtype=np.int64)
    with tempfile.NamedTemporaryFile("r", suffix=".log") as f:
        ref().ref_compare(_p(a, _I64P), _p(b, _I64P), a.shape[0], nthreads, f.name.encode())
        text = open(f.name).read()
    out = {}
    for line in text.splitlines():
        if ":" not in line:
            continue
        key, val = line.split(":", 1)
        key = key.strip()
        if key in _REPORT:
            name, typ = _REPORT[key]
            out[name] = typ(val.split()[0])
    return out, a, b


def ref_gini(sizes):
    s = np.array(sizes, dtype=np.int64)
    return ref().ref_gini(_p(s, _I64P), s.shape[0]), s


def ref_renumber(C):
    C = np.array(C, dtype=np.int64)
    k = ref().ref_renumber(_p(C, _I64P), C.shape[0])
    return C, k


def ref_vertex_following(g):
    C = np.empty(g.numVertices, np.int64)
    n = ref().ref_vertex_following(g.numVertices, g.numEdges, _p(g.edgeListPtrs, _I64P),
                                   _p(g.tails, _I64P), _p(C, _I64P))
    return C, n


def ref_build_vf(g, C, k):
    return _ref_contract(lambda *a: ref().ref_build_vf(*a), g, C, k, ())


def ref_coarsen(g, C, k, nthreads=1):
    return _ref_contract(lambda *a: ref().ref_coarsen(*a), g, C, k, (nthreads,))


def _ref_contract(fn, g, C, k, extra):
    from grappolo_amd.graph import Graph
    C = np.array(C, dtype=np.int64)
    po, pt, pw = _I64P(), _I64P(), _F64P()
    ne = ctypes.c_int64(0)
    tot = fn(*_rg(g), _p(C, _I64P), k, *extra, ctypes.byref(po), ctypes.byref(pt), ctypes.byref(pw),
             ctypes.byref(ne))
    off = np.ctypeslib.as_array(po, shape=(k + 1,)).copy()
    tails = np.ctypeslib.as_array(pt, shape=(max(tot, 1),))[:tot].copy()
    w = np.ctypeslib.as_array(pw, shape=(max(tot, 1),))[:tot].copy()
    for p in (po, pt, pw):
        ref().ref_free(ctypes.cast(p, ctypes.c_void_p))
    return Graph(off, tails, w, numEdges=ne.value)


def _ref_load(fn, path):
    from grappolo_amd.graph import Graph
    po, pt, pw = _I64P(), _I64P(), _F64P()
    ne = ctypes.c_int64(0)
    nv = getattr(ref(), fn)(str(path).encode(), ctypes.byref(po), ctypes.byref(pt),
                            ctypes.byref(pw), ctypes.byref(ne))
    off = np.ctypeslib.as_array(po, shape=(nv + 1,)).copy()
    nnz = int(off[-1])
    tails = np.ctypeslib.as_array(pt, shape=(max(nnz, 1),))[:nnz].copy()
    w = np.ctypeslib.as_array(pw, shape=(max(nnz, 1),))[:nnz].copy()
    for p in (po, pt, pw):
        ref().ref_free(ctypes.cast(p, ctypes.c_void_p))
    return Graph(off, tails, w, numEdges=ne.value)


def ref_load_binary(path):
    """The reference's parse_EdgeListBinaryNew (InputsOutput/loadBinary.cpp:6-48)."""
    return _ref_load("ref_load_binary", path)


def ref_load_metis(path):
    """The reference's loadMetisFileFormat (InputsOutput/loadMetis.cpp:19-232)."""
    return _ref_load("ref_load_metis", path)


def ref_write_binary(path, g):
    """The reference's writeGraphBinaryFormatNew (InputsOutput/writeBinary.cpp:2-29)."""
    ref().ref_write_binary(str(path).encode(), *_rg(g))


def _parse_final(log):
    final = None
    with open(log) as f:
        for line in f:
            if line.startswith("Final modularity"):
                final = float(line.split(":")[1])
    return final


def ref_run_colored(g, replaceMap=0, minGraphSize=1000, threshold=1e-9, C_threshold=1e-6,
                    nthreads=1):
    C_orig = np.empty(g.numVertices, dtype=np.int64)
    with tempfile.NamedTemporaryFile(suffix=".log", delete=False) as t:
        log = t.name
    try:
        ref().ref_run_colored(*_rg(g), _p(C_orig, _I64P), int(replaceMap), int(minGraphSize),
                              threshold, C_threshold, nthreads, log.encode())
        final = _parse_final(log)
    finally:
        os.unlink(log)
    return C_orig, final


def ref_run_basic(g, basicOpt=0, threshold=1e-9, nthreads=1):
    C_orig = np.empty(g.numVertices, dtype=np.int64)
    with tempfile.NamedTemporaryFile(suffix=".log", delete=False) as t:
        log = t.name
    try:
        ref().ref_run_basic(*_rg(g), _p(C_orig, _I64P), int(basicOpt), threshold, nthreads,
                            log.encode())
        final = _parse_final(log)
    finally:
        os.unlink(log)
    return C_orig, final
